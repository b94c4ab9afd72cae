// Asynchronous argument loads, demonstrated on pointer chasing: many linked
// lists whose nodes lie scattered through device memory.
//
// make_work starts one Step per list at its head node; a Step adds its node's
// value to the list's running sum and queues a Step for the next node, until
// the list's end, where the sum and length are written. The node each Step
// consumes is its LazyLoad argument (reference async/promise.h:204-240):
//  - ChaseLazySpec declares the parameter as LazyLoad<ChaseNode>, so the
//    promise carries the node's address and AsyncProgram loads the nodes of a
//    whole link while the link before it runs (the reference's ASYNC_LOADS);
//  - ChaseEagerSpec declares a plain ChaseNode, so the same call site reads
//    the node when the promise is built (the reference without ASYNC_LOADS),
//    a dependent load in the middle of every Step.
// Both compute identical sums and lengths.
//
// Written against the spec API only; compiled for gfx950 by the runtime and
// for the host by the oracle's executors.
#pragma once

#include <cstdint>

constexpr unsigned int CHASE_END = 0xFFFFFFFFu;

// One list node (8 bytes): the index of the next node (CHASE_END at the end
// of the list) and a value.
struct ChaseNode {
    unsigned int next;
    unsigned int value;
};

struct ChaseState {
    unsigned int lists;
    unsigned int pad;
    const unsigned int* heads;  // [lists] first node of each list
    const ChaseNode* nodes;
    unsigned long long* sums;   // [lists]
    unsigned int* lengths;      // [lists]
    util::iter::AtomicIter<unsigned int>* list_iter;
};

template <bool LAZY>
struct ChaseStepT {
    using NodeArg = typename std::conditional<LAZY, LazyLoad<ChaseNode>, ChaseNode>::type;
    using Type = void (*)(unsigned int list, unsigned int length, unsigned long long sum, NodeArg node);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned int list, unsigned int length, unsigned long long sum,
                                NodeArg node_arg) {
        const ChaseNode node = node_arg;
        sum += node.value;
        length += 1u;
        if (node.next == CHASE_END) {
            prog.device.sums[list] = sum;
            prog.device.lengths[list] = length;
            return;
        }
        prog.template async<ChaseStepT<LAZY>>(list, length, sum, LazyLoad<ChaseNode>(prog.device.nodes[node.next]));
    }
};

template <bool LAZY>
struct ChaseSpecT {
    using OpSet = OpUnion<ChaseStepT<LAZY>>;
    using DeviceState = ChaseState;

    static const size_t STASH_SIZE = 8;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM) {}

    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        util::iter::Iter<unsigned int> it = prog.device.list_iter->leap(1u);
        unsigned int list;
        while (it.step(list)) {
            const unsigned int head = prog.device.heads[list];
            if (head == CHASE_END) {
                prog.device.sums[list] = 0ull;
                prog.device.lengths[list] = 0u;
            } else {
                prog.template async<ChaseStepT<LAZY>>(list, 0u, 0ull, LazyLoad<ChaseNode>(prog.device.nodes[head]));
            }
        }
        return !prog.device.list_iter->done();
    }
};

using ChaseLazySpec = ChaseSpecT<true>;
using ChaseEagerSpec = ChaseSpecT<false>;
