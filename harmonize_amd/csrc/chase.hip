// Asynchronous argument loads of the C ABI (specs/chase.h): linked-list
// pointer chasing with the node as a LazyLoad argument, loaded ahead of the
// link that consumes it (ChaseLazySpec), or read where the promise is built
// (ChaseEagerSpec), under AsyncProgram or EventProgram.
#include "common.h"
#include "chase.h"

using hz::capi::check;
using hz::capi::guarded;

namespace {

using ChaseLazyAsync = AsyncProgram<ChaseLazySpec>;
using ChaseEagerAsync = AsyncProgram<ChaseEagerSpec>;
using ChaseLazyEvent = EventProgram<ChaseLazySpec>;
using ChaseEagerEvent = EventProgram<ChaseEagerSpec>;

static_assert(ChaseLazyAsync::HAS_LAZY && ChaseLazyAsync::LAZY_WORDS == 2, "the lazy spec prefetches 8-byte nodes");
static_assert(!ChaseEagerAsync::HAS_LAZY, "the eager spec carries node values");
static_assert(sizeof(hz_chase_node) == sizeof(ChaseNode), "C and C++ nodes must match");

template <typename P>
void run_async_chase(ChaseState st, const hz_run_config& cfg, hz_run_stats* stats) {
    uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::resident_groups_of<P>();
    uint64_t arena = cfg.arena ? cfg.arena : (1ull << 16);
    uint32_t cycles = cfg.cycles ? cfg.cycles : (1u << 24);
    typename P::Instance inst(arena, st, cfg.queues, cfg.coalesce ? cfg.coalesce : 64u);
    init<P>(inst, groups);
    check(hipDeviceSynchronize(), "init");
    hz::capi::run_async<P>(inst, groups, cycles, cfg.max_launches, stats, (cfg.flags & HZ_RUN_AUDIT) != 0);
}

template <typename P>
void run_event_chase(ChaseState st, uint32_t lists, const hz_run_config& cfg, hz_run_stats* stats) {
    uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::compute_units() * 16u;
    uint64_t io = cfg.arena;
    if (io == 0) {
        io = 1ull << 16;
        while (io < 2ull * lists && io < (1ull << 26)) io <<= 1;
    }
    typename P::Instance inst(io, st);
    hz::capi::run_event<P>(inst, groups, cfg.cycles ? cfg.cycles : 4u, cfg.max_launches, stats);
}

}  // namespace

extern "C" int hz_chase_run(int kind, int lazy, uint32_t lists, const uint32_t* d_heads, const hz_chase_node* d_nodes,
                            uint64_t* d_sums, uint32_t* d_lengths, const hz_run_config* config, hz_run_stats* stats) {
    return guarded([&] {
        if (lists && (!d_heads || !d_nodes || !d_sums || !d_lengths)) {
            throw std::invalid_argument("hz_chase_run: null array");
        }
        hz_run_config cfg{};
        if (config) cfg = *config;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (lists == 0) return;
        hz::capi::DeviceIter iter(lists);
        ChaseState st{lists, 0u, d_heads, (const ChaseNode*)d_nodes, (unsigned long long*)d_sums, d_lengths, iter.ptr};
        if (kind == HZ_ASYNC) {
            if (lazy) {
                run_async_chase<ChaseLazyAsync>(st, cfg, stats);
            } else {
                run_async_chase<ChaseEagerAsync>(st, cfg, stats);
            }
        } else if (kind == HZ_EVENT) {
            if (lazy) {
                run_event_chase<ChaseLazyEvent>(st, lists, cfg, stats);
            } else {
                run_event_chase<ChaseEagerEvent>(st, lists, cfg, stats);
            }
        } else {
            throw std::invalid_argument("hz_chase_run: kind must be HZ_EVENT or HZ_ASYNC");
        }
    });
}
