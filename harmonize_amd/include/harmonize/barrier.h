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
//  - AsyncProgram keeps held promises in arena links, filled by tickets: an
//    awaiting lane draws ticket t from its operation's counter tickets[op]
//    with one atomic add; ticket t is slot t % 64 of block t / 64, and each
//    block fills one arena link. The lane that draws a block's slot 0
//    installs the block's link and publishes it in ring[op][block % RING];
//    the other lanes of the block wait for that entry (reading only), write
//    their promise into their slot and count themselves in the link's
//    (target << 16) | filled word. Whoever brings `filled` to `target` marks
//    the ring entry done (so block + RING may take it) and hands the link
//    on: onto the barrier's full list (a lock-free stack in full_head) or,
//    once released, straight to the work queues. Release sets the RELEASED
//    bit of every ticket counter (later awaits run their promise at once),
//    closes the last block of each operation at the tickets it handed out,
//    and moves the full list to the queues. Every await costs one atomic on
//    the counter however many lanes wait on the barrier, and no lane retries
//    a reservation.
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

    // Blocks of 64 tickets per operation whose links may be in flight at once.
    // A block's slot-0 lane waits for the block RING before it to fill, so
    // RING x 64 should cover the awaits a device can have outstanding on one
    // barrier at once (4 KiB of ring per operation; 64 barriers with 32,768
    // awaits each took 362, 60 and 40 ms with rings of 8, 128 and 512).
#ifndef HZ_BARRIER_RING
#define HZ_BARRIER_RING 512
#endif
    static constexpr unsigned int RING = HZ_BARRIER_RING;
    // Set in a ticket counter once the barrier is released.
    static constexpr unsigned int RELEASED_BIT = 0x80000000u;

    // The three words every await or decrement touches sit on 128-byte lines
    // of their own, so that the semaphore's decrements, the ticket draws and
    // the full list's pushes of one barrier do not queue behind each other on
    // one line. The semaphore stays first: a future's ReturnOp finds the
    // barrier from its address.
    TaggedSemaphore semaphore;
    alignas(128) unsigned int full_head;  // list of full held links, or BARRIER_RELEASED
    // Tickets handed out per operation (at most 2^31 before release); their
    // sum, without RELEASED_BIT, is the number of promises awaited so far.
    alignas(128) unsigned int tickets[TYPE_COUNT > 0 ? TYPE_COUNT : 1];
    // ring[op][block % RING] = (block << 32) | link: the link of a block
    // being filled; (block << 32) | BARRIER_NO_LINK once it is full.
    unsigned long long int ring[TYPE_COUNT > 0 ? TYPE_COUNT : 1][RING];

    RemappingBarrier() = default;

    // A barrier holding nothing, released when `sem_val` decrements bring it
    // to zero (reference barrier.h:60-70).
    static HZ_HD RemappingBarrier blank(unsigned int sem_val) {
        RemappingBarrier b;
        b.semaphore = TaggedSemaphore(sem_val, hz::BARRIER_TAG);
        b.full_head = hz::BARRIER_NO_LINK;
        for (size_t i = 0; i < (TYPE_COUNT > 0 ? TYPE_COUNT : 1); ++i) {
            b.tickets[i] = 0u;
            // Entry k starts as the finished block k - RING, so block k may install.
            for (unsigned int k = 0; k < RING; ++k) {
                b.ring[i][k] = ((unsigned long long)(unsigned int)(k - RING) << 32) | hz::BARRIER_NO_LINK;
            }
        }
        return b;
    }

    // Promises awaited on this barrier so far, released or held (a statistic,
    // read when no await is in flight).
    HZ_HD unsigned long long awaited() const {
        unsigned long long n = 0;
        for (size_t i = 0; i < (TYPE_COUNT > 0 ? TYPE_COUNT : 1); ++i) n += tickets[i] & ~RELEASED_BIT;
        return n;
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
