// Per-label state accessors of the C ABI, named as the reference's generated
// binding names them (harmonize/python/templates.py:85-152; labels from
// runtime.py:562-595 and :1348-1371): one field of a program's device state
// at a time, the group and thread labels, and the wall-clock rate.
#include <cstring>

#include "common.h"

using hz::capi::check;
using hz::capi::guarded;

namespace {

void copy_out(void* host, const void* dev_state, size_t offset, size_t bytes, const char* what) {
    if (!host || !dev_state) throw std::invalid_argument(std::string(what) + ": null pointer");
    check(hipMemcpy(host, (const char*)dev_state + offset, bytes, hipMemcpyDeviceToHost), what);
}

void copy_in(void* dev_state, const void* host, size_t offset, size_t bytes, const char* what) {
    if (!host || !dev_state) throw std::invalid_argument(std::string(what) + ": null pointer");
    check(hipMemcpy((char*)dev_state + offset, host, bytes, hipMemcpyHostToDevice), what);
}

// The reference's message for a pointer store into a field that is not an
// array (templates.py:146-150).
[[noreturn]] void not_an_array() {
    throw std::runtime_error("Attempted to set indirection of non-array state field.");
}

}  // namespace

#define HZ_DEFINE_STATE_FIELD_FOR(STATE, suffix, field, is_pointer)                                             \
    extern "C" void load_state_device_##field##_##suffix(void* host_ptr, void* dev_ptr) {                       \
        guarded([&] {                                                                                           \
            copy_out(host_ptr, dev_ptr, offsetof(STATE, field), sizeof(((STATE*)nullptr)->field),              \
                     "load_state_device_" #field "_" #suffix);                                                  \
        });                                                                                                     \
    }                                                                                                           \
    extern "C" void store_state_device_##field##_##suffix(void* dev_ptr, void* host_ptr) {                      \
        guarded([&] {                                                                                           \
            copy_in(dev_ptr, host_ptr, offsetof(STATE, field), sizeof(((STATE*)nullptr)->field),               \
                    "store_state_device_" #field "_" #suffix);                                                  \
        });                                                                                                     \
    }                                                                                                           \
    extern "C" void store_pointer_state_device_##field##_##suffix(void* dev_ptr, void* host_ptr) {              \
        guarded([&] {                                                                                           \
            if (!(is_pointer)) not_an_array();                                                                  \
            static_assert(!(is_pointer) || sizeof(((STATE*)nullptr)->field) == sizeof(void*),                 \
                          "pointer fields hold one pointer");                                                   \
            copy_in(dev_ptr, &host_ptr, offsetof(STATE, field), sizeof(void*),                                  \
                    "store_pointer_state_device_" #field "_" #suffix);                                          \
        });                                                                                                     \
    }

#define HZ_DEFINE_STATE_LABELS(STATE, suffix, FIELDS)                                                           \
    FIELDS(HZ_DEFINE_STATE_FIELD_##STATE, suffix)                                                               \
    extern "C" void store_pointer_state_device_##suffix(void*, void*) { guarded([] { not_an_array(); }); }     \
    extern "C" void load_state_group_##suffix(void*, void*) { guarded([] {}); }                                \
    extern "C" void store_state_group_##suffix(void*, void*) { guarded([] {}); }                               \
    extern "C" void store_pointer_state_group_##suffix(void*, void*) { guarded([] { not_an_array(); }); }      \
    extern "C" void load_state_thread_##suffix(void*, void*) { guarded([] {}); }                               \
    extern "C" void store_state_thread_##suffix(void*, void*) { guarded([] {}); }                              \
    extern "C" void store_pointer_state_thread_##suffix(void*, void*) { guarded([] { not_an_array(); }); }

#define HZ_DEFINE_STATE_FIELD_hz_collatz_state(suffix, field, p) HZ_DEFINE_STATE_FIELD_FOR(hz_collatz_state, suffix, field, p)
#define HZ_DEFINE_STATE_FIELD_hz_photon_state(suffix, field, p) HZ_DEFINE_STATE_FIELD_FOR(hz_photon_state, suffix, field, p)
#define HZ_DEFINE_STATE_FIELD_hz_bfs_state(suffix, field, p) HZ_DEFINE_STATE_FIELD_FOR(hz_bfs_state, suffix, field, p)
#define HZ_DEFINE_STATE_FIELD_hz_barrier_state(suffix, field, p) HZ_DEFINE_STATE_FIELD_FOR(hz_barrier_state, suffix, field, p)

HZ_DEFINE_STATE_LABELS(hz_collatz_state, collatz_event, HZ_COLLATZ_STATE_FIELDS)
HZ_DEFINE_STATE_LABELS(hz_collatz_state, collatz_async, HZ_COLLATZ_STATE_FIELDS)
HZ_DEFINE_STATE_LABELS(hz_photon_state, photon_event, HZ_PHOTON_STATE_FIELDS)
HZ_DEFINE_STATE_LABELS(hz_photon_state, photon_async, HZ_PHOTON_STATE_FIELDS)
HZ_DEFINE_STATE_LABELS(hz_bfs_state, bfs_event, HZ_BFS_STATE_FIELDS)
HZ_DEFINE_STATE_LABELS(hz_bfs_state, bfs_async, HZ_BFS_STATE_FIELDS)
HZ_DEFINE_STATE_LABELS(hz_barrier_state, barrier_async, HZ_BARRIER_STATE_FIELDS)

extern "C" long long wall_clock_rate(void) {
    long long rate = 0;
    guarded([&] {
        int device = 0;
        check(hipGetDevice(&device), "hipGetDevice");
        int khz = 0;
        check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device), "wall clock rate");
        rate = (long long)khz * 1000;
    });
    return rate;
}
