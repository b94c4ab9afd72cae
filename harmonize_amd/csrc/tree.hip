// Depth-level demonstration program of the C ABI (specs/tree.h): a recursive
// tree of promises run flat, deepest level first (the reference's order,
// program/async.h:1164-1181) or shallowest level first, under AsyncProgram,
// and flat under EventProgram.
#include "common.h"
#include "tree.h"

using hz::capi::check;
using hz::capi::guarded;

namespace {

using TreeFlatAsync = AsyncProgram<TreeFlatSpec>;
using TreeDeepAsync = AsyncProgram<TreeDeepSpec>;
using TreeShallowAsync = AsyncProgram<TreeShallowSpec>;
using TreeEvent = EventProgram<TreeFlatSpec>;

static_assert(TreeDeepAsync::LEVELS == TREE_LEVELS && TreeFlatAsync::LEVELS == 1, "tree depth levels");

template <typename P>
void run_tree_async(TreeState st, const hz_run_config& cfg, hz_run_stats* stats) {
    uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::resident_groups_of<P>();
    uint64_t arena = cfg.arena ? cfg.arena : (1ull << 18);
    uint32_t cycles = cfg.cycles ? cfg.cycles : (1u << 24);
    typename P::Instance inst(arena, st, cfg.queues, cfg.coalesce ? cfg.coalesce : 64u);
    init<P>(inst, groups);
    check(hipDeviceSynchronize(), "init");
    hz::capi::run_async<P>(inst, groups, cycles, cfg.max_launches, stats, (cfg.flags & HZ_RUN_AUDIT) != 0);
}

}  // namespace

extern "C" int hz_tree_run(int kind, int order, uint32_t roots, uint32_t fanout, uint32_t height, uint32_t* d_leaves,
                           uint32_t* d_violations, const hz_run_config* config, hz_run_stats* stats) {
    return guarded([&] {
        if (!d_violations || (roots && !d_leaves)) throw std::invalid_argument("hz_tree_run: null output");
        if (fanout < 1 || fanout > 64) throw std::invalid_argument("hz_tree_run: fanout must be in [1, 64]");
        double leaves = 1.0;
        for (uint32_t i = 0; i < height; ++i) leaves *= fanout;
        if (height > 64 || leaves >= 4294967296.0) {
            throw std::invalid_argument("hz_tree_run: fanout^height must stay below 2^32 leaves per root");
        }
        if (order < HZ_TREE_FLAT || order > HZ_TREE_SHALLOWEST) throw std::invalid_argument("hz_tree_run: bad order");
        if (kind == HZ_EVENT && order != HZ_TREE_FLAT) {
            throw std::invalid_argument("hz_tree_run: EventProgram runs levels together (order must be flat)");
        }
        hz_run_config cfg{};
        if (config) cfg = *config;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        check(hipMemset(d_violations, 0, sizeof(uint32_t)), "zero violations");
        if (roots) check(hipMemset(d_leaves, 0, (size_t)roots * sizeof(uint32_t)), "zero leaves");
        if (roots == 0) return;
        hz::capi::DeviceIter iter(roots);
        TreeState st{roots, fanout, height, order == HZ_TREE_FLAT ? 1u : TREE_LEVELS, iter.ptr, d_leaves,
                     d_violations};
        if (kind == HZ_EVENT) {
            uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::compute_units() * 16u;
            uint64_t io = cfg.arena ? cfg.arena : (1ull << 24);
            TreeEvent::Instance inst(io, st);
            hz::capi::run_event<TreeEvent>(inst, groups, cfg.cycles ? cfg.cycles : 4u, cfg.max_launches, stats);
        } else if (kind == HZ_ASYNC) {
            switch (order) {
                case HZ_TREE_FLAT: run_tree_async<TreeFlatAsync>(st, cfg, stats); break;
                case HZ_TREE_DEEPEST: run_tree_async<TreeDeepAsync>(st, cfg, stats); break;
                default: run_tree_async<TreeShallowAsync>(st, cfg, stats); break;
            }
        } else {
            throw std::invalid_argument("hz_tree_run: kind must be HZ_EVENT or HZ_ASYNC");
        }
    });
}
