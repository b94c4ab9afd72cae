// Autotuning entry points of the C ABI: per-program coordinate descent over
// workgroup size, stash, group count, queue count, coalescing threshold,
// arena size and the program's own knob, timing the program's C entry point.
#include "common.h"
#include "tune.h"

using hz::capi::check;
using hz::capi::guarded;
using hz::tune::axis32;
using hz::tune::axis64;

namespace {

void ok(int status, const char* what) {
    if (status != HZ_OK) {
        const char* msg = hz_last_error();
        throw std::runtime_error(std::string(what) + ": " + (msg ? msg : "failed"));
    }
}

std::vector<uint64_t> group_counts() {
    uint64_t cu = (uint64_t)hz::capi::compute_units();
    return {0, cu * 2, cu * 4, cu * 8, cu * 16};
}

}  // namespace

extern "C" int hz_collatz_tune(const uint64_t* d_starts, uint32_t count, uint32_t* d_steps, uint32_t rounds,
                               uint32_t reps, hz_tune_result* result, hz_tune_trial* log, uint32_t log_capacity) {
    return guarded([&] {
        if (!d_steps || count == 0) throw std::invalid_argument("hz_collatz_tune: empty workload");
        hz_run_config start{};
        start.arena = 1ull << 16;
        std::vector<hz::tune::Axis> axes = {
            axis32("groups", &hz_run_config::groups, group_counts()),
            axis32("queues", &hz_run_config::queues, {1, 2, 4, 8, 16}),
            axis32("coalesce", &hz_run_config::coalesce, {16, 32, 48, 64}),
            axis64("arena", &hz_run_config::arena, {1ull << 14, 1ull << 16, 1ull << 18}, true),
        };
        hz::tune::coordinate_descent(
            start, axes,
            [&](const hz_run_config& c) {
                hz_run_stats st{};
                ok(hz_collatz_run(HZ_ASYNC, d_starts, count, d_steps, &c, &st), "collatz run");
                return st.elapsed_ms;
            },
            rounds, reps, result, log, log_capacity);
    });
}

extern "C" int hz_photon_tune(const hz_photon_params* params, double* d_tally, uint32_t* d_events, uint32_t rounds,
                              uint32_t reps, hz_tune_result* result, hz_tune_trial* log, uint32_t log_capacity) {
    return guarded([&] {
        if (!params || !d_tally || !d_events) throw std::invalid_argument("hz_photon_tune: null argument");
        hz_run_config start{};
        start.arena = 1ull << 16;
        std::vector<hz::tune::Axis> axes = {
            axis32("groups", &hz_run_config::groups, group_counts()),
            axis32("queues", &hz_run_config::queues, {1, 2, 4, 8, 16}),
            axis32("coalesce", &hz_run_config::coalesce, {16, 32, 48, 64}),
            axis64("arena", &hz_run_config::arena, {1ull << 14, 1ull << 16, 1ull << 18}, true),
        };
        hz::tune::coordinate_descent(
            start, axes,
            [&](const hz_run_config& c) {
                hz_run_stats st{};
                ok(hz_photon_run(HZ_ASYNC, params, d_tally, d_events, &c, &st), "photon run");
                return st.elapsed_ms;
            },
            rounds, reps, result, log, log_capacity);
    });
}

extern "C" int hz_bfs_batch_tune(hz_bfs_batch* batch, uint32_t rounds, uint32_t reps, hz_tune_result* result,
                                 hz_tune_trial* log, uint32_t log_capacity) {
    return guarded([&] {
        if (!batch) throw std::invalid_argument("hz_bfs_batch_tune: null batch");
        hz_bfs_info info{};
        ok(hz_bfs_batch_info(batch, &info), "bfs info");
        // Arena sizes around one link per 64 visits plus one per 64 x 64 edges.
        uint64_t base = (info.total_vertices + info.total_edges / 64 + 63) / 64;
        if (base < (1ull << 14)) base = 1ull << 14;
        hz_run_config start{};
        start.flags = HZ_RUN_GRAPH;
        start.max_launches = 1u << 16;
        start.arena = 2 * base;
        std::vector<hz::tune::Axis> axes = {
            axis32("block_groups", &hz_run_config::block_groups, {1, 2, 4}),
            axis32("stash", &hz_run_config::stash, {16, 8}),
            axis32("groups", &hz_run_config::groups, group_counts()),
            axis32("queues", &hz_run_config::queues, {1, 4, 8, 16}),
            axis32("coalesce", &hz_run_config::coalesce, {16, 32, 64}),
            axis32("split", &hz_run_config::spec_param, {16, 32, 64, 128, 256}),
            axis64("arena", &hz_run_config::arena, {base / 2, base, 2 * base}, true),
        };
        hz::tune::coordinate_descent(
            start, axes,
            [&](const hz_run_config& c) {
                hz_run_stats st{};
                int status = hz_bfs_batch_run(batch, HZ_ASYNC, &c, &st);
                // The batch keeps an instance (arena, queues, graph) per
                // configuration; free it so trials do not pile them up.
                int released = hz_bfs_batch_release(batch);
                ok(status, "bfs run");
                ok(released, "bfs release");
                return st.elapsed_ms;
            },
            rounds, reps, result, log, log_capacity);
    });
}
