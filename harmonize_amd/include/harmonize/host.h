// Host- and device-side utilities that specifications and host mains use
// through `using namespace util;` (reference util/mod.h:16-133 and
// util/host.h:10-279): wave helpers, the LCG used for scheduling choices,
// a hipEvent stopwatch, and owning device buffers.
#pragma once

#include "preamble.h"
#include "wave.h"

namespace util {

// Number of active lanes below this one (reference util/mod.h:16-23, made wave64).
__device__ HZ_INLINE unsigned int warp_inc_scan() { return hz::wave::rank(hz::wave::active()); }
// Number of active lanes (util/mod.h:29-31).
__device__ HZ_INLINE unsigned int active_count() { return hz::wave::popc(hz::wave::active()); }
// True on the lowest active lane (util/mod.h:39-41).
__device__ HZ_INLINE bool current_leader() { return hz::wave::leader(); }

__device__ HZ_INLINE unsigned int pop_count(unsigned int v) { return __popc(v); }
__device__ HZ_INLINE unsigned long long int pop_count(unsigned long long int v) { return __popcll(v); }
__device__ HZ_INLINE unsigned int leading_zeros(unsigned int v) { return __clz(v); }
__device__ HZ_INLINE unsigned long long int leading_zeros(unsigned long long int v) { return __clzll(v); }

// Linear congruential generator for scheduling choices (util/mod.h:66-77).
// Not for simulation randomness.
HZ_HD inline unsigned int random_uint(unsigned int& state) {
    state = 0x10DCDu * state + 1u;
    return state;
}
HZ_HD inline unsigned long long int random_uint(unsigned long long int& state) {
    state = 2971215073ull * state + 12345ull;
    return state;
}

// Times a span of GPU work with hipEvents on the null stream (util/mod.h:88-133).
struct Stopwatch {
    hipEvent_t beg;
    hipEvent_t end;
    float duration = 0.0f;

    Stopwatch() {
        if (hipEventCreate(&beg) != hipSuccess || hipEventCreate(&end) != hipSuccess) {
            throw std::runtime_error("Stopwatch: failed to create hip events");
        }
    }
    ~Stopwatch() {
        (void)hipEventDestroy(beg);
        (void)hipEventDestroy(end);
    }
    Stopwatch(const Stopwatch&) = delete;
    Stopwatch& operator=(const Stopwatch&) = delete;

    bool start() { return hipEventRecord(beg, nullptr) == hipSuccess; }
    bool stop() {
        if (hipEventRecord(end, nullptr) != hipSuccess) return false;
        if (hipEventSynchronize(end) != hipSuccess) return false;
        return hipEventElapsedTime(&duration, beg, end) == hipSuccess;
    }
    float ms_duration() { return duration; }
};

namespace host {

// Prints and reports any pending HIP error (util/host.h:10-23).
inline bool check_error() {
    hipError_t status = hipGetLastError();
    if (status != hipSuccess) {
        printf("ERROR: \"%s\"\n", hipGetErrorString(status));
    }
    return status != hipSuccess;
}

// Turns a HIP error into std::runtime_error (util/host.h:26-32).
inline void auto_throw(hipError_t status) {
    if (status != hipSuccess) {
        std::string message = "GPU Runtime Error: ";
        message += hipGetErrorString(status);
        throw std::runtime_error(message);
    }
}

template <typename T>
T* hardMalloc(size_t size) {
    T* result = nullptr;
    if (size == 0) return nullptr;
    auto_throw(hipMalloc((void**)&result, sizeof(T) * size));
    return result;
}

// A reference-counted device array (util/host.h:45-196).
template <typename T>
class DevBuf {
   protected:
    struct Inner {
        size_t size;
        T* adr;
        Inner(T* a, size_t s) : size(s), adr(a) {}
        ~Inner() {
            if (adr != nullptr) (void)hipFree(adr);
        }
    };
    std::shared_ptr<Inner> inner;

   public:
    operator T*&() { return inner->adr; }
    T* data() const { return inner->adr; }
    size_t size() const { return inner->size; }

    void resize(size_t s) {
        T* new_adr = hardMalloc<T>(s);
        size_t copy_count = (s < inner->size) ? s : inner->size;
        if (copy_count > 0) {
            auto_throw(hipMemcpy(new_adr, inner->adr, sizeof(T) * copy_count, hipMemcpyDeviceToDevice));
        }
        if (inner->adr != nullptr) auto_throw(hipFree(inner->adr));
        inner->size = s;
        inner->adr = new_adr;
    }

    void operator<<(std::vector<T>& other) {
        if (other.size() != inner->size) resize(other.size());
        if (inner->size > 0) {
            auto_throw(hipMemcpy(inner->adr, other.data(), sizeof(T) * inner->size, hipMemcpyHostToDevice));
        }
    }

    void operator>>(std::vector<T>& other) {
        if (other.size() != inner->size) other.resize(inner->size);
        if (inner->size > 0) {
            auto_throw(hipMemcpy(other.data(), inner->adr, sizeof(T) * inner->size, hipMemcpyDeviceToHost));
        }
    }

    void operator<<(const T& other) {
        if (inner->size != 1) resize(1);
        auto_throw(hipMemcpy(inner->adr, &other, sizeof(T), hipMemcpyHostToDevice));
    }

    void operator>>(T& other) {
        auto_throw(hipMemcpy(&other, inner->adr, sizeof(T), hipMemcpyDeviceToHost));
    }

    DevBuf(T* adr, size_t size) : inner(new Inner(adr, size)) {}
    DevBuf() : DevBuf((T*)nullptr, (size_t)0) {}
    // Same overload set as the reference: a size_t argument sizes the buffer,
    // a T argument becomes its single element.
    DevBuf(size_t size) : DevBuf(hardMalloc<T>(size), size) {}
    DevBuf(T& value) : DevBuf() { (*this) << value; }
    DevBuf(T&& value) : DevBuf() { (*this) << value; }

    template <typename... ARGS>
    static DevBuf<T> make(ARGS... args) {
        return DevBuf<T>(T(args...));
    }
};

// A single device object with a host mirror; T may define host_init/host_free
// (util/host.h:201-279).
template <typename T>
class DevObj {
   protected:
    template <typename U>
    static auto call_init(U& u, int) -> decltype(u.host_init(), void()) { u.host_init(); }
    template <typename U>
    static void call_init(U&, long) {}
    template <typename U>
    static auto call_free(U& u, int) -> decltype(u.host_free(), void()) { u.host_free(); }
    template <typename U>
    static void call_free(U&, long) {}

    struct Inner {
        T* adr;
        T host_copy;
        void push_data() { auto_throw(hipMemcpy(adr, &host_copy, sizeof(T), hipMemcpyHostToDevice)); }
        void pull_data() { auto_throw(hipMemcpy(&host_copy, adr, sizeof(T), hipMemcpyDefault)); }
        template <typename... ARGS>
        Inner(T* a, ARGS... args) : adr(a), host_copy(args...) {
            call_init(host_copy, 0);
            push_data();
        }
        ~Inner() {
            if (adr != nullptr) {
                (void)hipMemcpy(&host_copy, adr, sizeof(T), hipMemcpyDefault);
                call_free(host_copy, 0);
                (void)hipFree(adr);
            }
        }
    };
    std::shared_ptr<Inner> inner;

   public:
    void push_data() { inner->push_data(); }
    void pull_data() { inner->pull_data(); }
    operator T*() { return inner->adr; }
    operator T() { return inner->host_copy; }
    T& host_copy() { return inner->host_copy; }

    template <typename... ARGS>
    DevObj(ARGS... args) : inner(new Inner(hardMalloc<T>(1), args...)) {}
};

}  // namespace host
}  // namespace util
