// ORACLE — test infrastructure only. Nothing in the product links or calls
// anything under oracle/; only tests, the smoke check and the bench's CPU
// baseline do.
//
// Host stand-ins for the few device facilities that specification sources
// use, so that the same specification headers compile with g++ and run in
// the sequential and OpenMP host executors (host_program.h). Atomics are
// real (GCC __atomic builtins) so the OpenMP executor is race-free.
#pragma once

#if defined(__HIPCC__) || defined(__HIP__)
#error "host_shim.h is for host compilers only"
#endif

#include <cmath>
#include <cstdint>
#include <cstring>

#define __host__
#define __device__
#define __global__
#define __forceinline__ inline __attribute__((always_inline))

// The host executors run one group per thread; this is that thread's index.
inline int& hz_host_group_slot() {
    static thread_local int slot = 0;
    return slot;
}
inline unsigned int hz_group_index() { return (unsigned int)hz_host_group_slot(); }
inline unsigned int hz_block_thread() { return 0u; }
inline unsigned int hz_block_threads() { return 1u; }
inline void __syncthreads() {}
#define __shared__ static

// Ordering around barrier semaphore updates (the atomics below are
// sequentially consistent already; the fences keep plain stores ordered).
inline void hz_fence_release() { __atomic_thread_fence(__ATOMIC_RELEASE); }
inline void hz_fence_acquire() { __atomic_thread_fence(__ATOMIC_ACQUIRE); }

template <typename T>
inline T hz_fetch_add(T* counter, T amount) {
    return __atomic_fetch_add(counter, amount, __ATOMIC_SEQ_CST);
}
template <typename T>
inline T hz_load(const T* p) {
    return __atomic_load_n(p, __ATOMIC_SEQ_CST);
}

// Device atomic names used by specifications.
inline unsigned int atomicAdd(unsigned int* p, unsigned int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline double atomicAdd(double* p, double v) {
    double old;
    __atomic_load(p, &old, __ATOMIC_SEQ_CST);
    double next;
    do {
        next = old + v;
    } while (!__atomic_compare_exchange(p, &old, &next, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
    return old;
}
inline float atomicAdd(float* p, float v) {
    float old;
    __atomic_load(p, &old, __ATOMIC_SEQ_CST);
    float next;
    do {
        next = old + v;
    } while (!__atomic_compare_exchange(p, &old, &next, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
    return old;
}
inline int atomicMin(int* p, int v) {
    int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline unsigned int atomicMin(unsigned int* p, unsigned int v) {
    unsigned int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline unsigned int atomicOr(unsigned int* p, unsigned int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned int atomicExch(unsigned int* p, unsigned int v) {
    return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned int atomicCAS(unsigned int* p, unsigned int expected, unsigned int v) {
    __atomic_compare_exchange_n(p, &expected, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return expected;
}
