// Collatz step-count specification.
//
// Same program as the reference's example (examples/cpp/collaz/collaz_hrm.cpp:
// Even :9-32, Odd :36-61, make_work :127-144): two async operations that
// carry (value, step, index); Even halves an even value, Odd maps an odd value
// v to 3v+1, each then re-dispatches on the parity of the new value, and a
// value <= 1 writes its step count. Differences, both forced by the workloads:
//  - starts come from a device array (configurations draw them at random),
//    with start = index when the array is absent, as in the reference;
//  - 3v+1 is checked for overflow of 64 bits (starts below 2^40 exist whose
//    trajectories pass 2^64); such a start records COLLATZ_OVERFLOW.
//
// Written against the spec API only; compiled for gfx950 by the runtime and
// for the host by the oracle's sequential executor.
#pragma once

#include <cstdint>

constexpr unsigned int COLLATZ_OVERFLOW = 0xFFFFFFFFu;

struct CollatzState {
    const unsigned long long* starts;
    unsigned int* steps;
    util::iter::AtomicIter<unsigned int>* iterator;
};

struct CollatzEven;
struct CollatzOdd;

template <typename PROGRAM>
__device__ inline void collatz_dispatch(PROGRAM& prog, unsigned long long val, unsigned int step,
                                        unsigned int index) {
    if ((val & 1ull) == 0) {
        prog.template async<CollatzEven>(val, step, index);
    } else {
        prog.template async<CollatzOdd>(val, step, index);
    }
}

struct CollatzEven {
    using Type = void (*)(unsigned long long val, unsigned int step, unsigned int index);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned long long val, unsigned int step, unsigned int index) {
        if (val <= 1) {
            prog.device.steps[index] = step;
            return;
        }
        collatz_dispatch(prog, val >> 1, step + 1, index);
    }
};

struct CollatzOdd {
    using Type = void (*)(unsigned long long val, unsigned int step, unsigned int index);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned long long val, unsigned int step, unsigned int index) {
        if (val <= 1) {
            prog.device.steps[index] = step;
            return;
        }
        // 3v + 1 <= 2^64 - 1  <=>  v <= (2^64 - 2) / 3
        if (val > 0x5555555555555554ull) {
            prog.device.steps[index] = COLLATZ_OVERFLOW;
            return;
        }
        collatz_dispatch(prog, val * 3ull + 1ull, step + 1, index);
    }
};

// Stash links per group, and start indices each thread claims from the
// shared iterator at once (see make_work). Seven links is the smallest stash
// that does not spill (two open links plus the three kept free before each
// evaluation); fewer links per wavefront mean more resident wavefronts to
// hide LDS and memory latency: 2^26 starts take 79 ms with 7 links against
// 145 ms with 16, while 6 links spill on nearly every cycle.
#ifndef HZ_COLLATZ_STASH
#define HZ_COLLATZ_STASH 7
#endif
#ifndef HZ_COLLATZ_CLAIM
#define HZ_COLLATZ_CLAIM 8
#endif
// Below this many unclaimed starts a thread claims one at a time, so that a
// small job (or the end of a large one) spreads its chains over more
// wavefronts: a chain advances once per scheduler cycle of its wavefront, so
// a wavefront holding fewer chains finishes them sooner.
#ifndef HZ_COLLATZ_SPREAD_BELOW
#define HZ_COLLATZ_SPREAD_BELOW (1u << 22)
#endif
// Wavefront groups per workgroup.
#ifndef HZ_COLLATZ_BLOCK
#define HZ_COLLATZ_BLOCK 1
#endif

// Start indices a thread has claimed and not yet launched.
struct CollatzClaim {
    unsigned int next;
    unsigned int end;
};

template <typename PROGRAM>
__device__ inline void collatz_launch(PROGRAM& prog, unsigned int index) {
    unsigned long long start = prog.device.starts ? prog.device.starts[index] : (unsigned long long)index;
    collatz_dispatch(prog, start, 0u, index);
}

struct CollatzSpec {
    typedef OpUnion<CollatzEven, CollatzOdd> OpSet;
    typedef CollatzState DeviceState;
    typedef CollatzClaim ThreadState;

    static const size_t STASH_SIZE = HZ_COLLATZ_STASH;
    static const size_t GROUPS_PER_BLOCK = HZ_COLLATZ_BLOCK;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM prog) {
        prog.thread.next = 0u;
        prog.thread.end = 0u;
    }

    // Claimed starts not launched yet are queued now, so a pass that ends
    // early loses none of them.
    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM prog) {
        while (prog.thread.next < prog.thread.end) {
            collatz_launch(prog, prog.thread.next);
            prog.thread.next += 1u;
        }
    }

    // Launches one start per thread per call, as the reference's make_work
    // does (collaz_hrm.cpp:127-144), but claims HZ_COLLATZ_CLAIM starts from
    // the shared iterator at a time: that iterator is one atomic counter for
    // the whole device, and same-address atomics serialise across the XCDs.
    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        if (prog.thread.next >= prog.thread.end) {
            util::iter::AtomicIter<unsigned int>& starts = *prog.device.iterator;
            unsigned int at = hz_load(&starts.value);
            unsigned int left = at < starts.limit ? starts.limit - at : 0u;
            util::iter::Iter<unsigned int> it = starts.leap(left < HZ_COLLATZ_SPREAD_BELOW ? 1u : HZ_COLLATZ_CLAIM);
            prog.thread.next = it.value;
            prog.thread.end = it.limit;
        }
        if (prog.thread.next < prog.thread.end) {
            collatz_launch(prog, prog.thread.next);
            prog.thread.next += 1u;
            return true;
        }
        return !prog.device.iterator->done();
    }
};
