// Error state, device-memory helpers and device queries of the C ABI.
#include "common.h"

namespace hz {
namespace capi {

namespace {
thread_local std::string last_error;
thread_local bool has_error = false;
}  // namespace

void set_error(const std::string& message) {
    last_error = message;
    has_error = true;
}

void clear_error() {
    has_error = false;
    last_error.clear();
}

int compute_units() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    int cus = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    return cus > 0 ? cus : 1;
}

}  // namespace capi
}  // namespace hz

using hz::capi::check;
using hz::capi::guarded;

extern "C" const char* hz_last_error(void) {
    return hz::capi::has_error ? hz::capi::last_error.c_str() : nullptr;
}

extern "C" void* harmonize_alloc_device_bytes(size_t size) {
    void* p = nullptr;
    guarded([&] { check(hipMalloc(&p, size), "hipMalloc"); });
    return p;
}

extern "C" void* harmonize_alloc_managed_bytes(size_t size) {
    void* p = nullptr;
    guarded([&] { check(hipMallocManaged(&p, size), "hipMallocManaged"); });
    return p;
}

extern "C" void harmonize_free_device_bytes(void* ptr) {
    guarded([&] { check(hipFree(ptr), "hipFree"); });
}

extern "C" void harmonize_memcpy_host_to_device(void* dev_ptr, void* host_ptr, size_t byte_count) {
    guarded([&] { check(hipMemcpy(dev_ptr, host_ptr, byte_count, hipMemcpyHostToDevice), "hipMemcpy H2D"); });
}

extern "C" void harmonize_memcpy_device_to_host(void* host_ptr, void* dev_ptr, size_t byte_count) {
    guarded([&] { check(hipMemcpy(host_ptr, dev_ptr, byte_count, hipMemcpyDeviceToHost), "hipMemcpy D2H"); });
}

extern "C" int hz_device_synchronize(void) {
    return guarded([&] { check(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

extern "C" int hz_device_copy(void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (bytes && (!dst || !src)) throw std::invalid_argument("hz_device_copy: null pointer");
        check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
    });
}
