// ORACLE — test infrastructure only (see host_shim.h). Built into
// oracle/liboracle.so and called from tests, the smoke check and the
// bench's CPU-baseline leg through ctypes.
#include <omp.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "host_program.h"
#include "collatz.h"
#include "photon.h"
#include "barrier_demo.h"
#include "futures.h"
#include "bfs.h"
#include "tree.h"
#include "chase.h"
#include "philox.h"
#include "rmat.h"

extern "C" {

// Collatz step count of v with the runtime's overflow policy: the closed
// form of the reference's host checker (examples/cpp/collaz/collaz_hrm.cpp:
// 157-169; collaz_std.cpp:28-40), evaluated in 128-bit arithmetic so that a
// trajectory leaving 64 bits is detected exactly rather than wrapping.
// `evaluations` receives the number of promises the Collatz program
// evaluates for this start: one per step plus the final one, where a
// trajectory that overflows ends at the step whose 3v+1 would overflow.
static uint32_t collatz_closed_form(uint64_t start, uint64_t* evaluations) {
    unsigned __int128 v = start;
    const unsigned __int128 top = ((unsigned __int128)1 << 64);
    uint32_t steps = 0;
    while (v > 1) {
        if ((v & 1) == 0) {
            v >>= 1;
        } else {
            v = 3 * v + 1;
            if (v >= top) {
                *evaluations = (uint64_t)steps + 1u;
                return COLLATZ_OVERFLOW;
            }
        }
        steps++;
    }
    *evaluations = (uint64_t)steps + 1u;
    return steps;
}

uint32_t oracle_collatz_steps_one(uint64_t start) {
    uint64_t evaluations = 0;
    return collatz_closed_form(start, &evaluations);
}

// Closed form over many starts (starts may be NULL: start = index), on
// `threads` OpenMP threads (0 = all). Returns the number of promises the
// Collatz program evaluates over all starts.
uint64_t oracle_collatz_steps(const uint64_t* starts, uint32_t count, uint32_t* steps, int threads) {
    if (threads <= 0) threads = omp_get_max_threads();
    uint64_t promises = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096) reduction(+ : promises)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        uint64_t s = starts ? starts[i] : (uint64_t)i;
        uint64_t e = 0;
        steps[i] = collatz_closed_form(s, &e);
        promises += e;
    }
    return promises;
}

struct oracle_run_stats {
    uint64_t promises;
    uint64_t make_work_calls;
    uint64_t peak_queue;
    double elapsed_ms;
    int32_t threads;
    int32_t reserved;
};

static void fill_stats(oracle_run_stats* out, const oracle::RunStats& s, double ms, int threads) {
    if (!out) return;
    out->promises = s.promises;
    out->make_work_calls = s.make_work_calls;
    out->peak_queue = s.peak_queue;
    out->elapsed_ms = ms;
    out->threads = threads;
    out->reserved = 0;
}

// Runs the Collatz specification itself on the sequential host executor
// (threads == 1) or the OpenMP executor (threads != 1; 0 = all cores).
int oracle_collatz_program(const uint64_t* starts, uint32_t count, uint32_t* steps, int threads,
                           oracle_run_stats* stats) {
    util::iter::AtomicIter<unsigned int> iter(0u, count);
    CollatzState ds{(const unsigned long long*)starts, steps, &iter};
    auto t0 = std::chrono::steady_clock::now();
    oracle::RunStats s;
    int used = 1;
    if (threads == 1) {
        s = oracle::HostProgram<CollatzSpec>::run(ds);
    } else {
        used = threads <= 0 ? omp_get_max_threads() : threads;
        s = oracle::OmpProgram<CollatzSpec>::run(ds, used);
    }
    auto t1 = std::chrono::steady_clock::now();
    fill_stats(stats, s, std::chrono::duration<double, std::milli>(t1 - t0).count(), used);
    return 0;
}

// Philox4x32-10 of one counter under one key.
void oracle_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    hz::philox::Ctr r = hz::philox::philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    std::memcpy(out, r.v, sizeof(r.v));
}

// Same start sequence as hz_collatz_make_starts (declared in harmonize_amd.h).
void oracle_collatz_make_starts(uint64_t seed, uint32_t bits, uint32_t count, uint64_t* out) {
    const uint64_t modulus = (1ull << bits) - 1ull;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        hz::philox::Ctr r = hz::philox::philox4x32_10((uint32_t)i, 0u, 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
        uint64_t x = ((uint64_t)r.v[1] << 32) | r.v[0];
        out[i] = x % modulus + 1ull;
    }
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

// Photon-walk parameters, laid out as hz_photon_params in harmonize_amd.h.
struct oracle_photon_params {
    uint32_t photons;
    uint32_t bins;
    uint64_t seed;
    float thickness;
    float sigma_t;
    float albedo;
    float split_above;
    float roulette_below;
    float region_width;
    uint32_t regions;
    uint32_t max_split;
    uint32_t max_events;
    uint32_t replicas;
};

// Runs the photon specification on the host executors. tally receives the
// replica-summed tally (bins + PHOTON_EXTRA_SLOTS doubles), events one count
// per photon. Photons [first, first + count) are run (ids stay global).
int oracle_photon_program(const oracle_photon_params* prm, uint32_t first, uint32_t count, double* tally,
                          uint32_t* events, int threads, oracle_run_stats* stats) {
    int used = threads == 1 ? 1 : (threads <= 0 ? omp_get_max_threads() : threads);
    unsigned int replicas = 1;
    while (replicas < (unsigned int)used) replicas <<= 1;
    const unsigned int stride = prm->bins + PHOTON_EXTRA_SLOTS;
    std::vector<double> rep((size_t)replicas * stride, 0.0);
    util::iter::AtomicIter<unsigned int> iter(first, first + count);
    PhotonState st{};
    st.photon_count = prm->photons;
    st.seed_lo = (uint32_t)prm->seed;
    st.seed_hi = (uint32_t)(prm->seed >> 32);
    st.bins = prm->bins;
    st.replicas = replicas;
    st.max_split = prm->max_split;
    st.max_events = prm->max_events;
    st.regions = prm->regions;
    st.thickness = prm->thickness;
    st.mean_free_path = 1.0f / prm->sigma_t;
    st.albedo = prm->albedo;
    st.bin_scale = (float)prm->bins / prm->thickness;
    st.region_scale = 1.0f / prm->region_width;
    st.split_above = prm->split_above;
    st.roulette_below = prm->roulette_below;
    st.tally = rep.data();
    st.events = events;
    st.iterator = &iter;
    auto t0 = std::chrono::steady_clock::now();
    oracle::RunStats s = used == 1 ? oracle::HostProgram<PhotonSpec>::run(st)
                                   : oracle::OmpProgram<PhotonSpec>::run(st, used);
    auto t1 = std::chrono::steady_clock::now();
    for (unsigned int i = 0; i < stride; ++i) {
        double sum = 0.0;
        for (unsigned int k = 0; k < replicas; ++k) sum += rep[(size_t)k * stride + i];
        tally[i] = sum;
    }
    fill_stats(stats, s, std::chrono::duration<double, std::milli>(t1 - t0).count(), used);
    return 0;
}

// Runs the barrier specification on the host executors (same outputs as
// hz_barrier_run).
int oracle_barrier_program(uint32_t dep, uint32_t non_dep, uint32_t barrier_count, uint32_t* counter, uint32_t* hits,
                           uint32_t* violations, int threads, oracle_run_stats* stats) {
    using Barrier = RemappingBarrier<OpUnion<BarrierStep>>;
    std::vector<Barrier> barriers(barrier_count);
    for (uint32_t i = 0; i < barrier_count; ++i) {
        barriers[i] = Barrier::blank(dep / barrier_count + (i < dep % barrier_count ? 1u : 0u));
    }
    util::iter::AtomicIter<unsigned int> iter(0u, dep + non_dep);
    *counter = 0;
    *violations = 0;
    for (uint32_t i = 0; i < dep; ++i) hits[i] = 0;
    BarrierDemoState st{};
    st.dep_count = dep;
    st.non_dep_count = non_dep;
    st.barrier_count = barrier_count;
    st.source_id_iter = &iter;
    st.counter = counter;
    st.hits = hits;
    st.violations = violations;
    st.barrier = barriers.data();
    int used = threads == 1 ? 1 : (threads <= 0 ? omp_get_max_threads() : threads);
    auto t0 = std::chrono::steady_clock::now();
    oracle::RunStats s = used == 1 ? oracle::HostProgram<BarrierDemoSpec>::run(st)
                                   : oracle::OmpProgram<BarrierDemoSpec>::run(st, used);
    auto t1 = std::chrono::steady_clock::now();
    fill_stats(stats, s, std::chrono::duration<double, std::milli>(t1 - t0).count(), used);
    return 0;
}

// Runs the depth-level tree specification on the host executors (same
// outputs as hz_tree_run): order 0 = flat, 1 = 8 levels deepest first, 2 = 8
// levels shallowest first. The host executors evaluate in their own order;
// the levels they report follow the depth deltas of the calls.
int oracle_tree_program(int order, uint32_t roots, uint32_t fanout, uint32_t height, uint32_t* leaves,
                        uint32_t* violations, int threads, oracle_run_stats* stats) {
    util::iter::AtomicIter<unsigned int> iter(0u, roots);
    *violations = 0;
    for (uint32_t i = 0; i < roots; ++i) leaves[i] = 0;
    TreeState st{roots, fanout, height, order == 0 ? 1u : TREE_LEVELS, &iter, leaves, violations};
    int used = threads == 1 ? 1 : (threads <= 0 ? omp_get_max_threads() : threads);
    auto t0 = std::chrono::steady_clock::now();
    oracle::RunStats s;
    auto go = [&](auto* spec_tag) {
        using S = std::remove_pointer_t<decltype(spec_tag)>;
        s = used == 1 ? oracle::HostProgram<S>::run(st) : oracle::OmpProgram<S>::run(st, used);
    };
    if (order == 0) {
        go((TreeFlatSpec*)nullptr);
    } else if (order == 1) {
        go((TreeDeepSpec*)nullptr);
    } else {
        go((TreeShallowSpec*)nullptr);
    }
    auto t1 = std::chrono::steady_clock::now();
    fill_stats(stats, s, std::chrono::duration<double, std::milli>(t1 - t0).count(), used);
    return 0;
}

// Runs the pointer-chasing specification (lazy = LazyLoad parameter or
// plain node) on the host executors (same outputs as hz_chase_run).
int oracle_chase_program(int lazy, uint32_t lists, const uint32_t* heads, const uint32_t* nodes, uint64_t* sums,
                         uint32_t* lengths, int threads, oracle_run_stats* stats) {
    util::iter::AtomicIter<unsigned int> iter(0u, lists);
    ChaseState st{lists, 0u, heads, (const ChaseNode*)nodes, (unsigned long long*)sums, lengths, &iter};
    int used = threads == 1 ? 1 : (threads <= 0 ? omp_get_max_threads() : threads);
    auto t0 = std::chrono::steady_clock::now();
    oracle::RunStats s;
    if (lazy) {
        s = used == 1 ? oracle::HostProgram<ChaseLazySpec>::run(st) : oracle::OmpProgram<ChaseLazySpec>::run(st, used);
    } else {
        s = used == 1 ? oracle::HostProgram<ChaseEagerSpec>::run(st) : oracle::OmpProgram<ChaseEagerSpec>::run(st, used);
    }
    auto t1 = std::chrono::steady_clock::now();
    fill_stats(stats, s, std::chrono::duration<double, std::milli>(t1 - t0).count(), used);
    return 0;
}

// Runs the futures specification on the host executors (same outputs as
// hz_futures_run). Unresolved futures read as UINT64_MAX.
int oracle_futures_program(const uint64_t* starts, uint32_t leaves, uint32_t readers, uint64_t* echo,
                           uint64_t* forwarded, uint64_t* total, int threads, oracle_run_stats* stats) {
    if (leaves == 0 || (leaves & (leaves - 1u)) != 0) return -1;
    using TotalFuture = Future<unsigned long long, RemappingBarrier<FuturesOps>>;
    using UnitFuture = Future<unsigned long long, UnitBarrier<FuturesOps>>;
    std::vector<unsigned long long> slot((size_t)leaves * 2u, 0ull);
    std::vector<UnitBarrier<FuturesOps>> join(leaves);
    TotalFuture tot(1u);
    std::vector<UnitFuture> fwd(readers ? readers : 1u);
    util::iter::AtomicIter<unsigned int> iter(0u, leaves + readers);
    FuturesState st{};
    st.starts = (const unsigned long long*)starts;
    st.slot = slot.data();
    st.join = join.data();
    st.total = &tot;
    st.forwarded = fwd.data();
    st.echo = (unsigned long long*)echo;
    st.iter = &iter;
    st.leaves = leaves;
    st.readers = readers;
    uint32_t setup = leaves > readers ? leaves : readers;
    for (uint32_t k = 0; k < setup; ++k) futures_setup_node(st, k);
    int used = threads == 1 ? 1 : (threads <= 0 ? omp_get_max_threads() : threads);
    auto t0 = std::chrono::steady_clock::now();
    oracle::RunStats s = used == 1 ? oracle::HostProgram<FuturesSpec>::run(st)
                                   : oracle::OmpProgram<FuturesSpec>::run(st, used);
    auto t1 = std::chrono::steady_clock::now();
    *total = tot.barrier.semaphore.sem == 0u ? tot.data : FUTURES_EARLY;
    for (uint32_t r = 0; r < readers; ++r) {
        forwarded[r] = fwd[r].barrier.semaphore.sem == 0u ? fwd[r].data : FUTURES_EARLY;
    }
    fill_stats(stats, s, std::chrono::duration<double, std::milli>(t1 - t0).count(), used);
    return 0;
}

// One R-MAT edge (the generator shared with the device; its bit mapping is
// pinned separately by a numpy restatement in the tests).
void oracle_rmat_edge(uint64_t seed, uint32_t g, uint32_t e, uint32_t vertices, uint32_t* src, uint32_t* dst) {
    hz::rmat::edge(g, e, vertices, hz::rmat::scale_of(vertices), (uint32_t)seed, (uint32_t)(seed >> 32), *src, *dst);
}

// CSR of an R-MAT batch built on the host: edges generated in order and
// placed by a counting sort on the source, which keeps each source's
// neighbours in generation order. Layout as hz_bfs_batch_create's:
// row_ptr of graph g at vertex_base[g] + g, col at edge_base[g].
// Graphs are numbered from first_graph in the generator's counters, so any
// sub-range of a batch can be rebuilt on its own.
void oracle_bfs_csr(uint64_t seed, uint32_t first_graph, uint32_t graphs, const uint32_t* vertices,
                    uint32_t edge_factor, int32_t* row_ptr, int32_t* col) {
    std::vector<uint64_t> vbase(graphs), ebase(graphs);
    uint64_t vb = 0, eb = 0;
    for (uint32_t g = 0; g < graphs; ++g) {
        vbase[g] = vb;
        ebase[g] = eb;
        vb += vertices[g];
        eb += (uint64_t)edge_factor * vertices[g];
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t gi = 0; gi < (int64_t)graphs; ++gi) {
        const uint32_t g = (uint32_t)gi;
        const uint32_t V = vertices[g];
        const uint32_t E = edge_factor * V;
        const uint32_t scale = hz::rmat::scale_of(V);
        std::vector<uint32_t> src(E), dst(E);
        for (uint32_t e = 0; e < E; ++e) {
            hz::rmat::edge(first_graph + g, e, V, scale, (uint32_t)seed, (uint32_t)(seed >> 32), src[e], dst[e]);
        }
        int32_t* rows = row_ptr + vbase[g] + g;
        int32_t* cols = col + ebase[g];
        std::vector<int64_t> count((size_t)V + 1, 0);
        for (uint32_t e = 0; e < E; ++e) count[src[e] + 1]++;
        for (uint32_t v = 0; v < V; ++v) count[v + 1] += count[v];
        for (uint32_t v = 0; v <= V; ++v) rows[v] = (int32_t)count[v];
        for (uint32_t e = 0; e < E; ++e) cols[count[src[e]]++] = (int32_t)dst[e];
    }
}

// Textbook queue BFS of each graph of a batch (layout as above, sources may
// be NULL: vertex 0). Unreached vertices get BFS_UNREACHED.
void oracle_bfs_levels(uint32_t graphs, const uint32_t* vertices, const uint64_t* edges, const int32_t* row_ptr,
                       const int32_t* col, const uint32_t* sources, int32_t* levels, int threads) {
    std::vector<uint64_t> vbase(graphs), ebase(graphs);
    uint64_t vb = 0, eb = 0;
    for (uint32_t g = 0; g < graphs; ++g) {
        vbase[g] = vb;
        ebase[g] = eb;
        vb += vertices[g];
        eb += edges[g];
    }
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int64_t gi = 0; gi < (int64_t)graphs; ++gi) {
        const uint32_t g = (uint32_t)gi;
        const uint32_t V = vertices[g];
        const int32_t* rows = row_ptr + vbase[g] + g;
        const int32_t* cols = col + ebase[g];
        int32_t* lev = levels + vbase[g];
        for (uint32_t v = 0; v < V; ++v) lev[v] = BFS_UNREACHED;
        uint32_t s = sources ? sources[g] : 0u;
        if (s >= V) continue;
        std::vector<uint32_t> queue;
        queue.reserve(V);
        lev[s] = 0;
        queue.push_back(s);
        for (size_t head = 0; head < queue.size(); ++head) {
            uint32_t v = queue[head];
            for (int32_t i = rows[v]; i < rows[v + 1]; ++i) {
                uint32_t w = (uint32_t)cols[i];
                if (lev[w] == BFS_UNREACHED) {
                    lev[w] = lev[v] + 1;
                    queue.push_back(w);
                }
            }
        }
    }
}

// The BFS specification itself on the host executors (threads == 1:
// sequential; otherwise OpenMP, 0 = all cores). Same layout and outputs.
int oracle_bfs_program(uint32_t graphs, const uint32_t* vertices, const uint64_t* edges, const int32_t* row_ptr,
                       const int32_t* col, const uint32_t* sources, int32_t* levels, uint32_t split, int threads,
                       oracle_run_stats* stats) {
    std::vector<BfsJob> jobs(graphs);
    uint64_t vb = 0, eb = 0;
    for (uint32_t g = 0; g < graphs; ++g) {
        jobs[g].row_ptr = row_ptr + vb + g;
        jobs[g].col = col + eb;
        jobs[g].level = levels + vb;
        jobs[g].vertices = vertices[g];
        jobs[g].source = sources ? sources[g] : 0u;
        for (uint32_t v = 0; v < vertices[g]; ++v) levels[vb + v] = BFS_UNREACHED;
        vb += vertices[g];
        eb += edges[g];
    }
    util::iter::AtomicIter<unsigned int> iter(0u, graphs);
    BfsState st{};
    st.jobs = jobs.data();
    st.job_iter = &iter;
    st.job_count = graphs;
    st.split = split ? split : BFS_DEFAULT_SPLIT;
    int used = threads == 1 ? 1 : (threads <= 0 ? omp_get_max_threads() : threads);
    auto t0 = std::chrono::steady_clock::now();
    oracle::RunStats s = used == 1 ? oracle::HostProgram<BfsSpec>::run(st) : oracle::OmpProgram<BfsSpec>::run(st, used);
    auto t1 = std::chrono::steady_clock::now();
    fill_stats(stats, s, std::chrono::duration<double, std::milli>(t1 - t0).count(), used);
    return 0;
}

// Deterministic math of the photon specification, exposed for tests.
double oracle_det_log(double x) { return hz::det::det_log(x); }
double oracle_det_sqrt(double x) { return hz::det::det_sqrt(x); }

}  // extern "C"
