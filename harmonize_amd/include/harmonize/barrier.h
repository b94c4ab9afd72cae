// RemappingBarrier and UnitBarrier (reference sync/barrier.h:15-641,
// :648-699).
//
// A RemappingBarrier holds awaited promises until its semaphore reaches
// zero, coalescing them by operation into work links; on release every held
// promise becomes runnable and later awaits run their promise directly.
// UnitBarrier does the same for a single promise.
//
// The barrier is a plain struct that lives in device memory (set up on the
// host with blank()). Its methods delegate the holding and releasing to the
// program type, so the same specification runs on the GPU runtime and on the
// host executors:
//  - AsyncProgram keeps held promises in arena links: a lane reserves a slot
//    in its operation's partial link with one 64-bit atomic on
//    partial_table[op] = (link << 32) | reserved, writes its promise into
//    the link's structure-of-arrays columns, and counts itself in the link's
//    (target << 16) | filled word; whoever brings `filled` to `target` hands
//    the link on: onto the barrier's full list (a lock-free stack in
//    full_head) or, once released, straight to the work queues. Release
//    marks every partial word and the list head as released, closes partial
//    links at their reserved count and moves the full list to the queues.
//  - the host executors hold the promises in a side table.
//
// A semaphore decrement is a release and the decrement that reaches zero an
// acquire (hz_fence_release / hz_fence_acquire, provided by the platform
// layer), so whatever a decrementer wrote before (a future's value, say) is
// visible to the promises the barrier releases.
#pragma once

#include "promise.h"

template <typename OP_SET, typename ADR_TYPE>
struct RemappingBarrier {
    using AdrType = ADR_TYPE;
    // Indexed like the program's operation set, which has ReturnOp appended
    // when the operations return values.
    using UnionType = PromiseUnion<typename hz::ExecOpSet<OP_SET>::type>;
    static const size_t TYPE_COUNT = UnionType::Info::COUNT;

    TaggedSemaphore semaphore;
    unsigned int count;      // promises awaited on this barrier (statistic)
    unsigned int full_head;  // list of full held links, or BARRIER_RELEASED
    unsigned long long int partial_table[TYPE_COUNT > 0 ? TYPE_COUNT : 1];

    RemappingBarrier() = default;

    // A barrier holding nothing, released when `sem_val` decrements bring it
    // to zero (reference barrier.h:60-70).
    static HZ_HD RemappingBarrier blank(unsigned int sem_val) {
        RemappingBarrier b;
        b.semaphore = TaggedSemaphore(sem_val, hz::BARRIER_TAG);
        b.count = 0u;
        b.full_head = hz::BARRIER_NO_LINK;
        for (size_t i = 0; i < (TYPE_COUNT > 0 ? TYPE_COUNT : 1); ++i) {
            b.partial_table[i] = (unsigned long long)hz::BARRIER_NO_LINK << 32;
        }
        return b;
    }

    // Holds `promise` until release; runs it at once if already released
    // (reference barrier.h:460-609).
    template <bool CAN_RELEASE = true, typename PROGRAM, typename OP_TYPE>
    __device__ HZ_INLINE void await(PROGRAM program, Promise<OP_TYPE> promise) {
        static_assert(UnionType::template Lookup<OP_TYPE>::type::CONTAINED,
                      "the barrier cannot hold promises of this operation");
        program.barrier_await(*this, promise);
    }

    // Releases the barrier regardless of its semaphore (barrier.h:614-621).
    template <typename PROGRAM>
    __device__ HZ_INLINE void release(PROGRAM program) {
        hz_fence_release();
        unsigned int before = atomicExch(&semaphore.sem, 0u);
        if (before != 0u) {
            hz_fence_acquire();
            program.barrier_release(*this);
        }
    }

    // Adds `delta` (wrapping, so (unsigned)-1 decrements) and releases when
    // the semaphore reaches zero (barrier.h:624-629).
    template <typename PROGRAM>
    __device__ HZ_INLINE void add_semaphore(PROGRAM program, unsigned int delta) {
        hz_fence_release();
        unsigned int before = atomicAdd(&semaphore.sem, delta);
        if (before + delta == 0u) {
            hz_fence_acquire();
            program.barrier_release(*this);
        }
    }

    template <typename PROGRAM>
    __device__ HZ_INLINE void sub_semaphore(PROGRAM program, unsigned int delta) {
        add_semaphore(program, 0u - delta);
    }
};

// A barrier for one promise (reference barrier.h:648-699): when its
// semaphore reaches zero the promise it was built with is queued. A unit
// barrier built without a promise only counts (a future's readiness).
template <typename OP_SET, typename ADR_TYPE>
struct UnitBarrier {
    using AdrType = ADR_TYPE;
    // The program's operation set (ReturnOp appended for value-returning
    // operations), so that every UnitBarrier<S> of one specification has the
    // same layout whichever of the two sets S names.
    using EnumType = PromiseEnum<typename hz::ExecOpSet<OP_SET>::type>;
    static constexpr OpDisc NO_PROMISE = 0xFFFFFFFFu;

    TaggedSemaphore semaphore;
    EnumType promise;

    UnitBarrier() = default;
    HZ_HD UnitBarrier(EnumType promise_value, unsigned int semaphore_value)
        : semaphore(semaphore_value, hz::UNIT_BARRIER_TAG), promise(promise_value) {}
    template <typename OP_TYPE>
    HZ_HD UnitBarrier(Promise<OP_TYPE> promise_value, unsigned int semaphore_value)
        : semaphore(semaphore_value, hz::UNIT_BARRIER_TAG), promise(promise_value) {}
    explicit HZ_HD UnitBarrier(unsigned int semaphore_value) : semaphore(semaphore_value, hz::UNIT_BARRIER_TAG) {
        promise.disc = NO_PROMISE;
    }

    template <typename PROGRAM>
    __device__ HZ_INLINE void release(PROGRAM program) {
        hz_fence_release();
        unsigned int old_val = atomicExch(&semaphore.sem, 0u);
        if (old_val != 0u) fire(program);
    }

    template <typename PROGRAM>
    __device__ HZ_INLINE void add_semaphore(PROGRAM program, unsigned int delta) {
        hz_fence_release();
        unsigned int old_val = atomicAdd(&semaphore.sem, delta);
        if ((old_val + delta) == 0u && old_val != 0u) fire(program);
    }

    template <typename PROGRAM>
    __device__ HZ_INLINE void sub_semaphore(PROGRAM program, unsigned int delta) {
        add_semaphore(program, 0u - delta);
    }

   private:
    template <typename PROGRAM>
    __device__ HZ_INLINE void fire(PROGRAM program) {
        hz_fence_acquire();
        if (promise.disc != NO_PROMISE) program.async_call(promise);
    }
};

#include "future.h"
