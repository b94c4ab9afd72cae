// Wave64 primitives for gfx950 (CDNA4). Every runtime decision that the
// reference makes per warp with 32-bit lane masks (util/mod.h:16-41) is made
// here per wavefront with the 64-bit EXEC mask: ballots, lane ranks from
// mbcnt, wave-uniform values through readfirstlane, and wave-aggregated
// atomics (one global atomic per wavefront instead of one per lane).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace hz {
namespace wave {

constexpr uint32_t SIZE = 64;

// Lanes currently executing this instruction.
__device__ HZ_INLINE uint64_t active() { return __builtin_amdgcn_read_exec(); }

// Index of this lane within its wavefront.
__device__ HZ_INLINE uint32_t lane() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` that belong to lanes below this one.
__device__ HZ_INLINE uint32_t rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ HZ_INLINE uint32_t popc(uint64_t mask) { return (uint32_t)__builtin_popcountll(mask); }

__device__ HZ_INLINE uint64_t ballot(bool pred) { return __builtin_amdgcn_ballot_w64(pred); }

// True on the lowest active lane only.
__device__ HZ_INLINE bool leader() { return rank(active()) == 0; }

// Broadcast the value held by the lowest active lane.
__device__ HZ_INLINE uint32_t first(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ HZ_INLINE int32_t first(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ HZ_INLINE uint64_t first(uint64_t v) {
    uint32_t lo = first((uint32_t)v);
    uint32_t hi = first((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Value of `v` held by lane `src` (any lane, active or not at the time it was written).
__device__ HZ_INLINE uint32_t read_lane(uint32_t v, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)src);
}

// Orders this wavefront's LDS accesses; it does not synchronise other waves.
__device__ HZ_INLINE void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ HZ_INLINE void sleep_short() { __builtin_amdgcn_s_sleep(2); }

// Device-scope atomics on HBM (relaxed unless stated).
template <typename T>
__device__ HZ_INLINE T atomic_add(T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE T atomic_add_acqrel(T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE T atomic_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE T atomic_load_acquire(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE void atomic_store(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE void atomic_store_release(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE T atomic_or(T* p, T v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE T atomic_and(T* p, T v) {
    return __hip_atomic_fetch_and(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ HZ_INLINE void fence_acq_rel() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent"); }
__device__ HZ_INLINE void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ HZ_INLINE void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Fetch-and-add where every active lane asks for `amount`: one atomic per
// wavefront, each lane receiving the start of its own disjoint range in lane
// order. Lanes asking for different amounts fall back to one atomic each.
template <typename T>
__device__ HZ_INLINE T aggregated_fetch_add(T* counter, T amount) {
    uint64_t act = active();
    T lead = (T)first((uint64_t)amount);
    bool uniform = ballot(amount == lead) == act;
    if (uniform) {
        uint32_t r = rank(act);
        T base = 0;
        if (r == 0) base = atomic_add(counter, (T)(lead * (T)popc(act)));
        base = (T)first((uint64_t)base);
        return base + (T)r * lead;
    }
    return atomic_add(counter, amount);
}

}  // namespace wave
}  // namespace hz

// Index of the executing group (a wavefront's workgroup). Specifications use
// it to spread contended atomics over replicas of a tally.
__device__ HZ_INLINE unsigned int hz_group_index() { return blockIdx.x; }
// Thread index within the workgroup and the workgroup's thread count: the
// threads that share a specification's __shared__ data, so that loops over
// such data (zeroing, flushing) can be split among them.
__device__ HZ_INLINE unsigned int hz_block_thread() { return threadIdx.x; }
__device__ HZ_INLINE unsigned int hz_block_threads() { return blockDim.x; }

// Platform hooks used by the portable barrier layer: device-wide release and
// acquire ordering around semaphore updates.
__device__ HZ_INLINE void hz_fence_release() { hz::wave::fence_release(); }
__device__ HZ_INLINE void hz_fence_acquire() { hz::wave::fence_acquire(); }

// Platform hooks used by the portable iterator layer.
template <typename T>
__device__ HZ_INLINE T hz_fetch_add(T* counter, T amount) {
    return hz::wave::aggregated_fetch_add(counter, amount);
}
template <typename T>
__device__ HZ_INLINE T hz_load(const T* p) {
    return hz::wave::atomic_load(p);
}
