// Remapping-barrier specification, after the reference's barrier demo
// (examples/cpp/barrier_demo/demo.cpp:26-44, :153-162): make_work evaluates
// Step(id) for ids [0, dep + non_dep) in place; a dependent step (id < dep)
// awaits barrier[id % barriers] with a held Step and then decrements that
// barrier's semaphore, which starts at the number of dependents mapped to it.
// The reference's expected count is dep * 2 + non_dep (demo.cpp:214).
//
// Held steps carry HELD_BASE + id so the run can prove that every held promise
// ran exactly once (hits[id]) and only after its barrier was released
// (violations).
#pragma once

#include <cstdint>

struct BarrierStep;

constexpr unsigned int BARRIER_HELD_BASE = 0x80000000u;

struct BarrierDemoState {
    unsigned int dep_count;
    unsigned int non_dep_count;
    unsigned int barrier_count;
    unsigned int pad;
    util::iter::AtomicIter<unsigned int>* source_id_iter;
    unsigned int* counter;     // Step evaluations
    unsigned int* hits;        // [dep_count] runs of each held step
    unsigned int* violations;  // held steps that ran before their barrier was released
    RemappingBarrier<OpUnion<BarrierStep>>* barrier;
};

struct BarrierStep {
    using Type = void (*)(unsigned int);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned int arg) {
        const BarrierDemoState& st = prog.device;
        atomicAdd(st.counter, 1u);
        if (arg >= BARRIER_HELD_BASE) {
            unsigned int id = arg - BARRIER_HELD_BASE;
            atomicAdd(&st.hits[id], 1u);
            RemappingBarrier<OpUnion<BarrierStep>>& b = st.barrier[id % st.barrier_count];
            if (hz_load(&b.semaphore.sem) != 0u) atomicAdd(st.violations, 1u);
            return;
        }
        if (arg < st.dep_count) {
            RemappingBarrier<OpUnion<BarrierStep>>& b = st.barrier[arg % st.barrier_count];
            b.await(prog, Promise<BarrierStep>(BARRIER_HELD_BASE + arg));
            b.add_semaphore(prog, (unsigned int)-1);
        }
    }
};

struct BarrierDemoSpec {
    using OpSet = OpUnion<BarrierStep>;
    using DeviceState = BarrierDemoState;

    static const size_t STASH_SIZE = 8;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        util::iter::Iter<unsigned int> it = prog.device.source_id_iter->leap(1u);
        unsigned int id;
        while (it.step(id)) prog.template sync<BarrierStep>(id);
        return !prog.device.source_id_iter->done();
    }
};
