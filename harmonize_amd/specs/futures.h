// Value-returning operations and futures: a reduction tree over Collatz step
// counts, with readers of the total.
//
// - Leaf(i) returns a promise, Steps(starts[i], 0), which takes over Leaf's
//   return address. Steps(v, s) returns the value s when v <= 1 and otherwise
//   the promise Steps(next(v), s + 1), so each leaf's chain resolves its
//   address once, with steps(starts[i]) (collaz_hrm.cpp:9-61 counts the same
//   steps with void operations).
// - The leaves' results meet in a binary tree of unit futures, in heap order:
//   node k (1 <= k < leaves, leaves a power of two) has children 2k and 2k+1
//   and leaf i is node leaves + i. A node's result goes to slot[node] and
//   decrements join[node / 2], a UnitBarrier holding Join(node / 2). Join(k)
//   returns slot[2k] + slot[2k+1] as a plain value.
// - The root's result resolves `total`, a RemappingBarrier future.
// - Reader r queues Echo(r) on `total` (held until it is ready; it then
//   writes echo[r] = total + r), and runs Forward(r), whose eval returns
//   `total` itself (the future form). The runtime forwards the total to
//   forwarded[r], a unit future, through a ReturnOp held on `total`.
//
// Program types that cannot hold promises on a barrier (EventProgram) run
// the tree only; the readers need AsyncProgram or the host executors.
#pragma once

#include <cstdint>

struct FutLeaf;
struct FutSteps;
struct FutJoin;
struct FutEcho;
struct FutForward;

using FuturesOps = OpUnion<FutLeaf, FutSteps, FutJoin, FutEcho, FutForward>;

struct FuturesState {
    const unsigned long long* starts;                                  // [leaves]
    unsigned long long* slot;                                          // [2 * leaves]: result of node k at slot[k]
    UnitBarrier<FuturesOps>* join;                                     // [leaves]: join[k], 1 <= k < leaves
    Future<unsigned long long, RemappingBarrier<FuturesOps>>* total;   // the root's result
    Future<unsigned long long, UnitBarrier<FuturesOps>>* forwarded;    // [readers]
    unsigned long long* echo;                                          // [readers]
    util::iter::AtomicIter<unsigned int>* iter;                        // ids [0, leaves + readers)
    unsigned int leaves;                                               // a power of two
    unsigned int readers;
};

// Readers write this instead of total + r if they ran before the total was ready.
constexpr unsigned long long FUTURES_EARLY = ~0ull;

struct FutSteps {
    using Type = unsigned long long (*)(unsigned long long, unsigned int);

    template <typename PROGRAM>
    __device__ static Return<PROGRAM, unsigned long long> eval(PROGRAM, unsigned long long v, unsigned int s) {
        if (v <= 1ull) return (unsigned long long)s;
        return Promise<FutSteps>((v & 1ull) ? 3ull * v + 1ull : (v >> 1), s + 1u);
    }
};

struct FutLeaf {
    using Type = unsigned long long (*)(unsigned int);

    template <typename PROGRAM>
    __device__ static Return<PROGRAM, unsigned long long> eval(PROGRAM prog, unsigned int i) {
        return Promise<FutSteps>(prog.device.starts[i], 0u);
    }
};

struct FutJoin {
    using Type = unsigned long long (*)(unsigned int);

    template <typename PROGRAM>
    __device__ static unsigned long long eval(PROGRAM prog, unsigned int k) {
        const FuturesState& st = prog.device;
        return st.slot[2u * k] + st.slot[2u * k + 1u];
    }
};

struct FutEcho {
    using Type = void (*)(unsigned int);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned int r) {
        // Typed through PROGRAM, so that Future's members are looked up only
        // when eval is instantiated, with every operation complete.
        const auto& st = prog.device;
        st.echo[r] = st.total->ready() ? st.total->data + r : FUTURES_EARLY;
    }
};

struct FutForward {
    using Type = unsigned long long (*)(unsigned int);

    template <typename PROGRAM>
    __device__ static Return<PROGRAM, unsigned long long> eval(PROGRAM prog, unsigned int) {
        return *prog.device.total;
    }
};

// Where the result of tree node `node` goes (defined after the operations,
// which Future's barrier types need complete).
HZ_HD inline ReturnAdr<unsigned long long> futures_parent(const FuturesState& st, unsigned int node) {
    if (node == 1u) return st.total->adr();
    return ReturnAdr<unsigned long long>{&st.join[node >> 1].semaphore, &st.slot[node]};
}

// Sets up node k's join barrier (1 <= k < leaves) — its Join promise carries
// the parent's return address — and reader r's forwarded future.
HZ_HD inline void futures_setup_node(const FuturesState& st, unsigned int k) {
    if (k >= 1u && k < st.leaves) {
        Promise<FutJoin> join(k);
        join.ret = futures_parent(st, k);
        st.join[k] = UnitBarrier<FuturesOps>(join, 2u);
    }
    if (k < st.readers) {
        st.forwarded[k] = Future<unsigned long long, UnitBarrier<FuturesOps>>(1u);
        st.echo[k] = 0ull;
    }
}

struct FuturesSpec {
    using OpSet = FuturesOps;
    using DeviceState = FuturesState;

    // Six operations with ReturnOp: the async scheduler keeps OPS + 1 links
    // free for a link's appends, so a stash of 8 would hold only two links of
    // work and spill nearly every link it fills.
    static const size_t STASH_SIZE = 16;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;
    // Link forwarding slows the tree (2^20 leaves: 4.9 -> 8.2 ms alone,
    // 8.5 -> 11.5 ms with 2^14 readers), so it is off here.
    static const size_t FORWARD_LINKS = 0;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        const FuturesState& st = prog.device;
        util::iter::Iter<unsigned int> it = st.iter->leap(1u);
        unsigned int id;
        while (it.step(id)) {
            if (id < st.leaves) {
                Promise<FutLeaf> leaf(id);
                leaf.ret = futures_parent(st, st.leaves + id);
                prog.async_call(leaf);
            } else {
                if constexpr (PROGRAM::REMAPPING_AWAIT) {
                    unsigned int r = id - st.leaves;
                    st.total->await(prog, Promise<FutEcho>(r));
                    Promise<FutForward> forward(r);
                    forward.ret = st.forwarded[r].adr();
                    prog.async_call(forward);
                }
            }
        }
        return !st.iter->done();
    }
};
