// Launch autotuner: coordinate descent over the fields of hz_run_config.
//
// Starting from a configuration, each round visits every axis in turn, runs
// the workload at each of the axis's values with the other fields fixed, and
// keeps the fastest (minimum over `reps` runs). A value whose run fails (an
// exhausted arena, an overflowing queue, a launch bound) scores as infinite
// and is never chosen. On axes marked prefer_small, a smaller value within
// 3% of the fastest wins, so the arena is not sized beyond what pays.
#pragma once

#include <cmath>
#include <limits>
#include <vector>

#include "common.h"

namespace hz {
namespace tune {

struct Axis {
    const char* name;
    uint32_t hz_run_config::*u32 = nullptr;
    uint64_t hz_run_config::*u64 = nullptr;
    std::vector<uint64_t> values;
    bool prefer_small = false;

    uint64_t get(const hz_run_config& c) const { return u32 ? c.*u32 : c.*u64; }
    void set(hz_run_config& c, uint64_t v) const {
        if (u32) {
            c.*u32 = (uint32_t)v;
        } else {
            c.*u64 = v;
        }
    }
};

inline Axis axis32(const char* name, uint32_t hz_run_config::*f, std::vector<uint64_t> v, bool small = false) {
    Axis a;
    a.name = name;
    a.u32 = f;
    a.values = std::move(v);
    a.prefer_small = small;
    return a;
}

inline Axis axis64(const char* name, uint64_t hz_run_config::*f, std::vector<uint64_t> v, bool small = false) {
    Axis a;
    a.name = name;
    a.u64 = f;
    a.values = std::move(v);
    a.prefer_small = small;
    return a;
}

// RUN: double(const hz_run_config&), elapsed milliseconds of one run; throws
// if the configuration cannot run the workload.
template <typename RUN>
void coordinate_descent(hz_run_config start, const std::vector<Axis>& axes, RUN&& run, uint32_t rounds, uint32_t reps,
                        hz_tune_result* result, hz_tune_trial* log, uint32_t log_capacity) {
    const double INF = std::numeric_limits<double>::infinity();
    uint32_t trials = 0, failures = 0;
    auto measure = [&](const hz_run_config& c) {
        double best = INF;
        for (uint32_t r = 0; r < (reps ? reps : 1u); ++r) {
            double ms = INF;
            try {
                ms = run(c);
            } catch (const std::exception&) {
                failures++;
                (void)hipGetLastError();
                ms = INF;
            }
            if (log && trials < log_capacity) {
                log[trials].config = c;
                log[trials].ms = std::isfinite(ms) ? ms : -1.0;
            }
            trials++;
            if (!std::isfinite(ms)) break;
            best = ms < best ? ms : best;
        }
        return best;
    };
    hz_run_config cur = start;
    double cur_ms = measure(cur);
    const double start_ms = cur_ms;
    for (uint32_t round = 0; round < (rounds ? rounds : 1u); ++round) {
        bool changed = false;
        for (const Axis& ax : axes) {
            uint64_t keep = ax.get(cur);
            uint64_t best_v = keep;
            double best_ms = cur_ms;
            std::vector<std::pair<uint64_t, double>> scores{{keep, cur_ms}};
            for (uint64_t v : ax.values) {
                if (v == keep) continue;
                hz_run_config c = cur;
                ax.set(c, v);
                double ms = measure(c);
                scores.emplace_back(v, ms);
                if (ms < best_ms) {
                    best_ms = ms;
                    best_v = v;
                }
            }
            if (ax.prefer_small && std::isfinite(best_ms)) {
                for (auto& s : scores) {
                    if (s.first != 0 && s.first < best_v && s.second <= best_ms * 1.03) best_v = s.first;
                }
                for (auto& s : scores) {
                    if (s.first == best_v) best_ms = s.second;
                }
            }
            if (best_v != keep) {
                ax.set(cur, best_v);
                cur_ms = best_ms;
                changed = true;
            }
        }
        if (!changed) break;
    }
    if (!std::isfinite(cur_ms)) throw std::runtime_error("autotune: no configuration ran the workload");
    if (result) {
        result->best = cur;
        result->best_ms = cur_ms;
        result->start_ms = std::isfinite(start_ms) ? start_ms : -1.0;
        result->trials = trials;
        result->failures = failures;
    }
}

}  // namespace tune
}  // namespace hz
