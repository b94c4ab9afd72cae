// Autotuning entry points of the C ABI: per-program coordinate descent over
// workgroup size, stash, group count, queue count, coalescing threshold,
// arena size and the program's own knob, timing the program's C entry point;
// results are cached in a text file keyed by program, device and workload
// class.
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

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

// ---------------------------------------------------------------- the cache
std::mutex cache_mutex;
bool cache_path_set = false;
std::string cache_path;

std::string current_cache_path() {
    if (!cache_path_set) {
        const char* env = std::getenv("HZ_TUNE_CACHE");
        cache_path = env ? env : "";
        cache_path_set = true;
    }
    return cache_path;
}

// Fields of hz_run_config in file order.
constexpr int CONFIG_FIELDS = 10;

std::string format_entry(const std::string& key, const hz_run_config& c, double ms) {
    std::ostringstream o;
    o << key << '\t' << c.groups << ' ' << c.cycles << ' ' << c.arena << ' ' << c.queues << ' ' << c.coalesce << ' '
      << c.max_launches << ' ' << c.block_groups << ' ' << c.flags << ' ' << c.spec_param << ' ' << c.stash << '\t'
      << ms;
    return o.str();
}

bool parse_entry(const std::string& line, std::string& key, hz_run_config& c, double& ms) {
    const size_t t1 = line.find('\t');
    const size_t t2 = t1 == std::string::npos ? std::string::npos : line.find('\t', t1 + 1);
    if (t2 == std::string::npos) return false;
    key = line.substr(0, t1);
    std::istringstream f(line.substr(t1 + 1, t2 - t1 - 1));
    uint64_t v[CONFIG_FIELDS];
    for (int i = 0; i < CONFIG_FIELDS; ++i) {
        if (!(f >> v[i])) return false;
    }
    c = hz_run_config{};
    c.groups = (uint32_t)v[0];
    c.cycles = (uint32_t)v[1];
    c.arena = v[2];
    c.queues = (uint32_t)v[3];
    c.coalesce = (uint32_t)v[4];
    c.max_launches = (uint32_t)v[5];
    c.block_groups = (uint32_t)v[6];
    c.flags = (uint32_t)v[7];
    c.spec_param = (uint32_t)v[8];
    c.stash = (uint32_t)v[9];
    std::istringstream m(line.substr(t2 + 1));
    return (bool)(m >> ms);
}

std::map<std::string, std::pair<hz_run_config, double>> read_cache(const std::string& path) {
    std::map<std::string, std::pair<hz_run_config, double>> out;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        std::string key;
        hz_run_config c{};
        double ms = 0;
        if (parse_entry(line, key, c, ms)) out[key] = {c, ms};
    }
    return out;
}

bool cache_lookup(const std::string& key, hz_run_config& c, double& ms) {
    std::lock_guard<std::mutex> lock(cache_mutex);
    const std::string path = current_cache_path();
    if (path.empty()) return false;
    auto all = read_cache(path);
    auto it = all.find(key);
    if (it == all.end()) return false;
    c = it->second.first;
    ms = it->second.second;
    return true;
}

void cache_store(const std::string& key, const hz_run_config& c, double ms) {
    std::lock_guard<std::mutex> lock(cache_mutex);
    const std::string path = current_cache_path();
    if (path.empty()) return;
    auto all = read_cache(path);
    all[key] = {c, ms};
    const std::string tmp = path + ".tmp";
    {
        std::ofstream out(tmp, std::ios::trunc);
        if (!out) throw std::runtime_error("tune cache: cannot write " + tmp);
        for (auto& kv : all) out << format_entry(kv.first, kv.second.first, kv.second.second) << '\n';
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) throw std::runtime_error("tune cache: cannot replace " + path);
}

unsigned int log2_round(uint64_t n) {
    unsigned int k = 0;
    while (k < 63 && (1ull << (k + 1)) <= n + (n >> 1)) ++k;
    return k;
}

std::string device_key() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch = arch.substr(0, colon);
    return arch + "|" + std::to_string(prop.multiProcessorCount);
}

// Runs `tune` unless the cache holds the key; stores what it finds.
template <typename TUNE>
void tuned_or_cached(const std::string& workload, hz_tune_result* result, TUNE&& tune) {
    hz_tune_result local{};
    hz_tune_result* r = result ? result : &local;
    std::memset(r, 0, sizeof(*r));
    const std::string key = workload.substr(0, workload.find('|')) + "|" + device_key() + "|" +
                            workload.substr(workload.find('|') + 1);
    hz_run_config c{};
    double ms = 0;
    if (cache_lookup(key, c, ms)) {
        r->best = c;
        r->best_ms = ms;
        r->start_ms = -1.0;
        r->cached = 1;
        return;
    }
    tune(r);
    cache_store(key, r->best, r->best_ms);
}

}  // namespace

extern "C" int hz_tune_cache_set_path(const char* path) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(cache_mutex);
        cache_path = path ? path : "";
        cache_path_set = true;
    });
}

extern "C" int hz_tune_cache_lookup(const char* key, hz_run_config* config, double* best_ms, int* found) {
    return guarded([&] {
        if (!key || !config || !best_ms || !found) throw std::invalid_argument("hz_tune_cache_lookup: null argument");
        hz_run_config c{};
        double ms = 0;
        *found = cache_lookup(key, c, ms) ? 1 : 0;
        if (*found) {
            *config = c;
            *best_ms = ms;
        }
    });
}

extern "C" int hz_tune_cache_store(const char* key, const hz_run_config* config, double best_ms) {
    return guarded([&] {
        if (!key || !config) throw std::invalid_argument("hz_tune_cache_store: null argument");
        const std::string k(key);
        if (k.empty() || k.find('\t') != std::string::npos || k.find('\n') != std::string::npos) {
            throw std::invalid_argument("hz_tune_cache_store: a key is one line without tabs");
        }
        cache_store(k, *config, best_ms);
    });
}

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
        tuned_or_cached("collatz_async|starts=2^" + std::to_string(log2_round(count)), result, [&](hz_tune_result* r) {
            hz::tune::coordinate_descent(
                start, axes,
                [&](const hz_run_config& c) {
                    hz_run_stats st{};
                    ok(hz_collatz_run(HZ_ASYNC, d_starts, count, d_steps, &c, &st), "collatz run");
                    return st.elapsed_ms;
                },
                rounds, reps, r, log, log_capacity);
        });
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
        tuned_or_cached("photon_async|photons=2^" + std::to_string(log2_round(params->photons)), result,
                        [&](hz_tune_result* r) {
                            hz::tune::coordinate_descent(
                                start, axes,
                                [&](const hz_run_config& c) {
                                    hz_run_stats st{};
                                    ok(hz_photon_run(HZ_ASYNC, params, d_tally, d_events, &c, &st), "photon run");
                                    return st.elapsed_ms;
                                },
                                rounds, reps, r, log, log_capacity);
                        });
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
        const std::string workload = "bfs_batch_async|graphs=" + std::to_string(info.graphs) + ",vertices=2^" +
                                     std::to_string(log2_round(info.total_vertices)) + ",edges=2^" +
                                     std::to_string(log2_round(info.total_edges));
        tuned_or_cached(workload, result, [&](hz_tune_result* r) {
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
                rounds, reps, r, log, log_capacity);
        });
    });
}
