// Depth-level demonstration: a recursive tree of promises.
//
// make_work starts one Node per root; a Node at height d < H queues its
// `fanout` children one level deeper (async_call_cast with depth delta +1,
// the reference's depth-aware async call, program/async.h:1627-1660), and a
// leaf (d == H) counts itself for its root. Each node also checks that the
// depth level the program reports for it (prog.depth()) is its height, clamped
// to the program's levels; a mismatch is counted as a violation.
//
// Root r therefore counts fanout^H leaves, and the run evaluates
// roots * (fanout^(H+1) - 1) / (fanout - 1) promises. The order in which the
// levels run decides how many promises are alive at once: deepest-first (the
// reference's pull order) finishes subtrees before opening new ones, while
// flat or shallowest-first order expands the tree breadth-first and parks far
// more links in the arena.
//
// Written against the spec API only; compiled for gfx950 by the runtime and
// for the host by the oracle's executors.
#pragma once

#include <cstdint>

struct TreeState {
    unsigned int roots;
    unsigned int fanout;
    unsigned int height;
    unsigned int levels;  // depth levels of the program running it (1 = flat)
    util::iter::AtomicIter<unsigned int>* root_iter;
    unsigned int* leaves;      // [roots] leaves counted per root
    unsigned int* violations;  // nodes whose reported depth was wrong
};

struct TreeNode {
    using Type = void (*)(unsigned int root, unsigned int height);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned int root, unsigned int height) {
        const TreeState& st = prog.device;
        const unsigned int want = st.levels > 1 ? (height < st.levels ? height : st.levels - 1) : 0u;
        if (prog.depth() != want) atomicAdd(st.violations, 1u);
        if (height >= st.height) {
            atomicAdd(&st.leaves[root], 1u);
            return;
        }
        for (unsigned int k = 0; k < st.fanout; ++k) {
            prog.template async_call_cast<TreeNode>(1, Promise<TreeNode>(root, height + 1));
        }
    }
};

// Depth-level variants of the specification: STACK_SIZE levels, run deepest
// first (the reference's order) or, with SHALLOW, shallowest first.
template <unsigned int LEVELS_, bool SHALLOW>
struct TreeSpecT {
    using OpSet = OpUnion<TreeNode>;
    using DeviceState = TreeState;

    static const size_t STACK_SIZE = LEVELS_;
    static const size_t SHALLOW_FIRST = SHALLOW ? 1 : 0;
    static const size_t STASH_SIZE = 16;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        util::iter::Iter<unsigned int> it = prog.device.root_iter->leap(1u);
        unsigned int root;
        while (it.step(root)) prog.template async<TreeNode>(root, 0u);
        return !prog.device.root_iter->done();
    }
};

// The depth-level count the demo uses: one level per tree height up to 8.
constexpr unsigned int TREE_LEVELS = 8;
using TreeFlatSpec = TreeSpecT<0, false>;
using TreeDeepSpec = TreeSpecT<TREE_LEVELS, false>;
using TreeShallowSpec = TreeSpecT<TREE_LEVELS, true>;
