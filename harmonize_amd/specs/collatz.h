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

struct CollatzSpec {
    typedef OpUnion<CollatzEven, CollatzOdd> OpSet;
    typedef CollatzState DeviceState;

    static const size_t STASH_SIZE = 16;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM) {}

    // Claims one start per thread (reference collaz_hrm.cpp:127-144).
    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        util::iter::Iter<unsigned int> it = prog.device.iterator->leap(1u);
        unsigned int index;
        while (it.step(index)) {
            unsigned long long start =
                prog.device.starts ? prog.device.starts[index] : (unsigned long long)index;
            collatz_dispatch(prog, start, 0u, index);
        }
        return !prog.device.iterator->done();
    }
};
