// Remapping-barrier program of the C ABI: AsyncProgram over BarrierDemoSpec.
#include "common.h"
#include "barrier_demo.h"

using BarrierAsync = AsyncProgram<BarrierDemoSpec>;

static_assert(sizeof(hz_barrier_state) == sizeof(BarrierDemoState), "C and C++ barrier states must match");

HZ_DEFINE_PROGRAM_ABI(barrier_async, BarrierAsync, BarrierDemoState)

using hz::capi::check;
using hz::capi::guarded;

extern "C" int hz_barrier_run(uint32_t dep_count, uint32_t non_dep_count, uint32_t barrier_count, uint32_t* d_counter,
                              uint32_t* d_hits, uint32_t* d_violations, const hz_run_config* config,
                              hz_run_stats* stats) {
    return guarded([&] {
        if (barrier_count == 0) throw std::invalid_argument("hz_barrier_run: barrier_count must be positive");
        if (!d_counter || !d_violations || (dep_count && !d_hits)) {
            throw std::invalid_argument("hz_barrier_run: null output");
        }
        if ((uint64_t)dep_count + non_dep_count >= BARRIER_HELD_BASE) {
            throw std::invalid_argument("hz_barrier_run: too many steps");
        }
        hz_run_config cfg{};
        if (config) cfg = *config;
        if (stats) std::memset(stats, 0, sizeof(*stats));

        using Barrier = RemappingBarrier<OpUnion<BarrierStep>>;
        std::vector<Barrier> host(barrier_count);
        for (uint32_t i = 0; i < barrier_count; ++i) {
            uint32_t waits = dep_count / barrier_count + (i < dep_count % barrier_count ? 1u : 0u);
            host[i] = Barrier::blank(waits);
        }
        util::host::DevBuf<Barrier> barriers((size_t)barrier_count);
        barriers << host;
        check(hipMemset(d_counter, 0, sizeof(uint32_t)), "zero counter");
        check(hipMemset(d_violations, 0, sizeof(uint32_t)), "zero violations");
        if (dep_count) check(hipMemset(d_hits, 0, (size_t)dep_count * sizeof(uint32_t)), "zero hits");
        hz::capi::DeviceIter iter(dep_count + non_dep_count);

        BarrierDemoState st{};
        st.dep_count = dep_count;
        st.non_dep_count = non_dep_count;
        st.barrier_count = barrier_count;
        st.source_id_iter = iter.ptr;
        st.counter = d_counter;
        st.hits = d_hits;
        st.violations = d_violations;
        st.barrier = barriers.data();

        uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)hz::capi::resident_groups_of<BarrierAsync>();
        uint64_t arena = cfg.arena ? cfg.arena : (1ull << 16);
        uint32_t cycles = cfg.cycles ? cfg.cycles : (1u << 24);
        BarrierAsync::Instance inst(arena, st, cfg.queues, cfg.coalesce ? cfg.coalesce : 64u);
        init<BarrierAsync>(inst, groups);
        check(hipDeviceSynchronize(), "init");
        hz::capi::run_async<BarrierAsync>(inst, groups, cycles, cfg.max_launches, stats,
                                          (cfg.flags & HZ_RUN_AUDIT) != 0, true);
    });
}
