// Batched BFS of the C ABI: R-MAT graphs generated and turned into int32 CSR
// on the device, and one AsyncProgram (or EventProgram) that runs the BFS of
// every graph of the batch at once, optionally through a captured hipGraph.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "common.h"
#include "rmat.h"
#include "bfs.h"

using BfsAsync1 = AsyncProgram<BfsSpecT<1>>;
using BfsEvent = EventProgram<BfsSpecT<1>>;

static_assert(sizeof(BfsJob) == 32, "BfsJob layout");
static_assert(sizeof(hz_bfs_state) == sizeof(BfsState), "C and C++ BFS states must match");
static_assert(HZ_BFS_MAX_SPLIT == BFS_MAX_SPLIT, "C and C++ split limits must match");

HZ_DEFINE_PROGRAM_ABI(bfs_event, BfsEvent, BfsState)
HZ_DEFINE_PROGRAM_ABI(bfs_async, BfsAsync1, BfsState)

using hz::capi::check;
using hz::capi::guarded;

#include "graph.h"

using hz::graph::bfs_reset_kernel;
using hz::graph::grid_for;
using hz::graph::row_ptr_kernel;
using hz::graph::rmat_edges_kernel;

namespace {

struct DeviceBytes {
    void* p = nullptr;
    DeviceBytes() = default;
    explicit DeviceBytes(size_t bytes) { check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc(bfs)"); }
    ~DeviceBytes() {
        if (p) (void)hipFree(p);
    }
    DeviceBytes(const DeviceBytes&) = delete;
    DeviceBytes& operator=(const DeviceBytes&) = delete;
    DeviceBytes(DeviceBytes&& o) noexcept : p(o.p) { o.p = nullptr; }
    DeviceBytes& operator=(DeviceBytes&& o) noexcept {
        std::swap(p, o.p);
        return *this;
    }
    template <typename T>
    T* as() const {
        return (T*)p;
    }
};

// hipEvent timing on one stream.
struct StreamTimer {
    hipEvent_t a{}, b{};
    hipStream_t s;
    explicit StreamTimer(hipStream_t stream) : s(stream) {
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventCreate(&b), "hipEventCreate");
    }
    ~StreamTimer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    void start() { check(hipEventRecord(a, s), "hipEventRecord"); }
    double stop() {
        check(hipEventRecord(b, s), "hipEventRecord");
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        return ms;
    }
};

struct Runner {
    virtual ~Runner() = default;
    // audit: check the queues and arena on the host after every launch.
    virtual void run(hz_run_stats* stats, bool audit) = 0;
};

}  // namespace

struct hz_bfs_batch {
    uint32_t graphs = 0;
    std::vector<uint32_t> vertices, sources;
    std::vector<uint64_t> edges, vertex_base, edge_base, row_base;
    uint64_t total_vertices = 0, total_edges = 0;
    DeviceBytes row_ptr, col, level, jobs;
    std::unique_ptr<hz::capi::DeviceIter> iter;
    hipStream_t stream = nullptr;
    double build_ms = 0.0;
    std::map<std::vector<uint64_t>, std::unique_ptr<Runner>> runners;

    ~hz_bfs_batch() {
        runners.clear();
        if (stream) (void)hipStreamDestroy(stream);
    }

    void layout(const uint32_t* v, const uint64_t* e, const uint32_t* src) {
        vertices.assign(v, v + graphs);
        edges.assign(e, e + graphs);
        sources.resize(graphs);
        vertex_base.resize(graphs);
        edge_base.resize(graphs);
        row_base.resize(graphs);
        uint64_t vb = 0, eb = 0;
        for (uint32_t g = 0; g < graphs; ++g) {
            if (edges[g] > 0x7FFFFFFFull) throw std::invalid_argument("hz_bfs: a graph has 2^31 or more edges");
            if (vertices[g] > 0x7FFFFFFEu) throw std::invalid_argument("hz_bfs: a graph has too many vertices");
            sources[g] = src ? src[g] : 0u;
            vertex_base[g] = vb;
            row_base[g] = vb + g;
            edge_base[g] = eb;
            vb += vertices[g];
            eb += edges[g];
        }
        total_vertices = vb;
        total_edges = eb;
        row_ptr = DeviceBytes((total_vertices + graphs) * sizeof(int32_t));
        col = DeviceBytes(total_edges * sizeof(int32_t));
        level = DeviceBytes(total_vertices * sizeof(int32_t));
        jobs = DeviceBytes(graphs * sizeof(BfsJob));
        iter = std::make_unique<hz::capi::DeviceIter>(graphs);
        check(hipStreamCreate(&stream), "hipStreamCreate");
    }

    void write_jobs() {
        std::vector<BfsJob> host(graphs);
        for (uint32_t g = 0; g < graphs; ++g) {
            host[g].row_ptr = row_ptr.as<int32_t>() + row_base[g];
            host[g].col = col.as<int32_t>() + edge_base[g];
            host[g].level = level.as<int32_t>() + vertex_base[g];
            host[g].vertices = vertices[g];
            host[g].source = sources[g];
        }
        check(hipMemcpy(jobs.p, host.data(), graphs * sizeof(BfsJob), hipMemcpyHostToDevice), "jobs upload");
    }

    BfsState state(uint32_t split) const {
        BfsState s{};
        s.jobs = jobs.as<BfsJob>();
        s.job_iter = iter->ptr;
        s.job_count = graphs;
        s.split = split ? split : BFS_DEFAULT_SPLIT;
        return s;
    }

    void enqueue_reset() {
        hipLaunchKernelGGL(bfs_reset_kernel, dim3(grid_for(total_vertices)), dim3(256), 0, stream,
                           level.as<int32_t>(), total_vertices, iter->ptr, graphs);
    }
};

namespace {

uint32_t default_split(const hz_run_config& c) { return c.spec_param ? c.spec_param : BFS_DEFAULT_SPLIT; }

// Arena links for a batch: room for one visit per vertex and one edge-range
// promise per `split` edges at once, twice over, within [2^16, 2^26] links.
uint64_t default_arena(const hz_bfs_batch& b, uint32_t split) {
    uint64_t promises = b.total_vertices + b.total_edges / split + b.graphs;
    uint64_t links = 2 * ((promises + 63) / 64);
    if (links < (1ull << 16)) links = 1ull << 16;
    if (links > (1ull << 26)) links = 1ull << 26;
    return links;
}

template <typename P>
struct AsyncRunner : Runner {
    hz_bfs_batch& b;
    typename P::Instance inst;
    uint32_t groups, cycles, max_launches;
    hipGraphExec_t graph_exec = nullptr;

    AsyncRunner(hz_bfs_batch& batch, const hz_run_config& c)
        : b(batch),
          inst(c.arena ? c.arena : default_arena(batch, default_split(c)), batch.state(default_split(c)), c.queues,
               c.coalesce ? c.coalesce : 64u),
          groups(c.groups ? c.groups : (uint32_t)hz::capi::resident_groups_of<P>()),
          cycles(c.cycles ? c.cycles : (1u << 24)),
          max_launches(c.max_launches) {
        if (c.flags & HZ_RUN_GRAPH) {
            hipGraph_t graph = nullptr;
            check(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            enqueue_pass();
            check(hipStreamEndCapture(b.stream, &graph), "hipStreamEndCapture");
            hipError_t e = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            check(e, "hipGraphInstantiate");
        }
    }
    ~AsyncRunner() override {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    }

    // Reset of levels, jobs and queues, then the first exec launch.
    void enqueue_pass() {
        b.enqueue_reset();
        init<P>(inst, groups, b.stream);
        exec<P>(inst, groups, cycles, b.stream);
    }

    void run(hz_run_stats* stats, bool audit) override {
        StreamTimer t(b.stream);
        t.start();
        if (graph_exec) {
            check(hipGraphLaunch(graph_exec, b.stream), "hipGraphLaunch");
        } else {
            enqueue_pass();
        }
        check(hipGetLastError(), "bfs launch");
        check(hipStreamSynchronize(b.stream), "bfs exec");
        uint64_t launches = 1, last = ~0ull;
        uint32_t audits = 0;
        auto audit_now = [&] {
            if (!audit) return;
            const bool done = inst.complete();  // raises on a runtime error first
            hz::AsyncCensus c = inst.audit(false);
            audits++;
            if (done && c.queued_links != 0) {
                throw std::runtime_error("AsyncProgram audit: complete with links still queued");
            }
        };
        audit_now();
        // A pass that ends early (cycle budget) resumes with plain launches.
        while (!inst.complete()) {
            if (max_launches && launches >= max_launches) {
                throw std::runtime_error("AsyncProgram: launch bound reached before completion");
            }
            uint64_t progress = inst.read_status().promises;
            if (progress == last) throw std::runtime_error("AsyncProgram: a launch made no progress");
            last = progress;
            exec<P>(inst, groups, cycles, b.stream);
            check(hipGetLastError(), "bfs relaunch");
            check(hipStreamSynchronize(b.stream), "bfs exec");
            launches++;
            audit_now();
        }
        double ms = t.stop();
        if (stats) {
            hz::AsyncStatus s = inst.read_status();
            std::memset(stats, 0, sizeof(*stats));
            stats->promises = s.promises;
            stats->launches = launches;
            stats->spills = s.spills;
            stats->fills = s.fills;
            stats->make_work_calls = s.make_work_calls;
            stats->groups = (uint32_t)(blocks_for<P>(groups) * (P::WORK_GROUP_SIZE / 64));
            stats->elapsed_ms = ms;
            for (int i = 0; i < 8; ++i) stats->phase_cycles[i] = s.phase_cycles[i];
            stats->links_run = s.links_run;
            stats->forwarded = s.forwarded;
            stats->audits = audits;
            stats->arena_links = s.claim < inst.arena_size ? s.claim : inst.arena_size;
        }
    }
};

void run_event(hz_bfs_batch& b, const hz_run_config& c, hz_run_stats* stats) {
    uint64_t io = c.arena;
    if (io == 0) {
        uint64_t split = default_split(c);
        io = b.total_vertices + b.total_edges / split + b.graphs;
        if (io < (1ull << 20)) io = 1ull << 20;
        if (io > (1ull << 30)) io = 1ull << 30;
    }
    BfsEvent::Instance inst(io, b.state(default_split(c)));
    // 16 groups per CU: measured 22.2 ms on the 64-graph batch against 23.7
    // with 32 per CU and 23.9 with 8 per CU.
    uint32_t groups = c.groups ? c.groups : (uint32_t)(hz::capi::compute_units() * 16);
    uint32_t chunk = c.cycles ? c.cycles : 4u;
    uint32_t max_launches = c.max_launches ? c.max_launches : (1u << 24);
    StreamTimer t(b.stream);
    t.start();
    b.enqueue_reset();
    uint64_t launches = 0;
    do {
        if (launches >= max_launches) throw std::runtime_error("EventProgram: launch bound reached before completion");
        exec<BfsEvent>(inst, groups, chunk, b.stream);
        check(hipGetLastError(), "bfs event launch");
        check(hipStreamSynchronize(b.stream), "bfs event exec");
        launches++;
    } while (!inst.complete());
    double ms = t.stop();
    if (stats) {
        auto s = inst.read_status();
        std::memset(stats, 0, sizeof(*stats));
        stats->promises = s.promises;
        stats->launches = launches;
        stats->make_work_calls = s.make_work_calls;
        stats->groups = groups;
        stats->elapsed_ms = ms;
    }
}

std::vector<uint64_t> runner_key(const hz_run_config& c) {
    return {c.groups, c.cycles, c.arena, c.queues, c.coalesce, c.max_launches, c.block_groups, c.flags, c.spec_param,
            c.stash};
}

template <unsigned int STASH>
std::unique_ptr<Runner> make_async_runner_stash(hz_bfs_batch& b, const hz_run_config& c) {
    switch (c.block_groups) {
        case 0:
        case 1:
            return std::make_unique<AsyncRunner<AsyncProgram<BfsSpecT<1, STASH>>>>(b, c);
        case 2:
            return std::make_unique<AsyncRunner<AsyncProgram<BfsSpecT<2, STASH>>>>(b, c);
        case 4:
            return std::make_unique<AsyncRunner<AsyncProgram<BfsSpecT<4, STASH>>>>(b, c);
        default:
            throw std::invalid_argument("hz_bfs_batch_run: block_groups must be 1, 2 or 4");
    }
}

std::unique_ptr<Runner> make_async_runner(hz_bfs_batch& b, const hz_run_config& c) {
    switch (c.stash) {
        case 0:
        case 16:
            return make_async_runner_stash<16>(b, c);
        case 8:
            return make_async_runner_stash<8>(b, c);
        default:
            throw std::invalid_argument("hz_bfs_batch_run: stash must be 8 or 16 links");
    }
}

void check_batch(const hz_bfs_batch* b) {
    if (!b) throw std::invalid_argument("hz_bfs: null batch");
}

}  // namespace

extern "C" int hz_bfs_batch_sizes(uint64_t seed, uint32_t graphs, double min_log2, double max_log2,
                                  uint32_t* vertices) {
    return guarded([&] {
        if (graphs && !vertices) throw std::invalid_argument("hz_bfs_batch_sizes: null output");
        if (!(min_log2 >= 0.0 && max_log2 >= min_log2 && max_log2 <= 30.0)) {
            throw std::invalid_argument("hz_bfs_batch_sizes: need 0 <= min_log2 <= max_log2 <= 30");
        }
        for (uint32_t g = 0; g < graphs; ++g) {
            uint32_t bits = hz::rmat::size_bits(g, (uint32_t)seed, (uint32_t)(seed >> 32));
            double u = (double)bits * (1.0 / 4294967296.0);
            double v = std::nearbyint(std::exp2(min_log2 + (max_log2 - min_log2) * u));
            vertices[g] = v < 1.0 ? 1u : (uint32_t)v;
        }
    });
}

extern "C" int hz_bfs_batch_create(uint64_t seed, uint32_t graphs, const uint32_t* vertices, uint32_t edge_factor,
                                   hz_bfs_batch** out) {
    return guarded([&] {
        if (!out || (graphs && !vertices)) throw std::invalid_argument("hz_bfs_batch_create: null argument");
        *out = nullptr;
        if (graphs == 0) throw std::invalid_argument("hz_bfs_batch_create: no graphs");
        auto b = std::make_unique<hz_bfs_batch>();
        b->graphs = graphs;
        std::vector<uint64_t> edges(graphs);
        for (uint32_t g = 0; g < graphs; ++g) {
            if (vertices[g] == 0) throw std::invalid_argument("hz_bfs_batch_create: a graph has no vertices");
            edges[g] = (uint64_t)edge_factor * vertices[g];
        }
        b->layout(vertices, edges.data(), nullptr);

        uint64_t max_e = 0;
        for (uint64_t e : edges) max_e = e > max_e ? e : max_e;
        DeviceBytes src_in(max_e * 4), dst_in(max_e * 4), src_out(max_e * 4);
        size_t temp_bytes = 0;
        check(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, src_in.as<uint32_t>(), src_out.as<uint32_t>(),
                                                 dst_in.as<uint32_t>(), b->col.as<uint32_t>(), (int)max_e, 0, 32,
                                                 b->stream),
              "radix sort sizing");
        DeviceBytes temp(temp_bytes);
        StreamTimer t(b->stream);
        t.start();
        for (uint32_t g = 0; g < graphs; ++g) {
            uint32_t V = vertices[g];
            uint32_t E = (uint32_t)edges[g];
            uint32_t scale = hz::rmat::scale_of(V);
            int32_t* rows = b->row_ptr.as<int32_t>() + b->row_base[g];
            if (E > 0) {
                hipLaunchKernelGGL(rmat_edges_kernel, dim3(grid_for(E)), dim3(256), 0, b->stream, g, E, V, scale,
                                   (uint32_t)seed, (uint32_t)(seed >> 32), src_in.as<uint32_t>(),
                                   dst_in.as<uint32_t>());
                check(hipGetLastError(), "rmat launch");
                size_t tb = temp_bytes;
                check(hipcub::DeviceRadixSort::SortPairs(temp.p, tb, src_in.as<uint32_t>(), src_out.as<uint32_t>(),
                                                         dst_in.as<uint32_t>(),
                                                         b->col.as<uint32_t>() + b->edge_base[g], (int)E, 0,
                                                         scale > 0 ? (int)scale : 1, b->stream),
                      "radix sort");
            }
            hipLaunchKernelGGL(row_ptr_kernel, dim3(grid_for((uint64_t)V + 1)), dim3(256), 0, b->stream,
                               src_out.as<uint32_t>(), E, V, rows);
            check(hipGetLastError(), "row_ptr launch");
        }
        b->build_ms = t.stop();
        b->write_jobs();
        *out = b.release();
    });
}

extern "C" int hz_bfs_batch_from_csr(uint32_t graphs, const uint32_t* vertices, const uint64_t* edges,
                                     const int32_t* row_ptr, const int32_t* col, const uint32_t* sources,
                                     hz_bfs_batch** out) {
    return guarded([&] {
        if (!out || !vertices || !edges || !row_ptr) throw std::invalid_argument("hz_bfs_batch_from_csr: null argument");
        *out = nullptr;
        if (graphs == 0) throw std::invalid_argument("hz_bfs_batch_from_csr: no graphs");
        auto b = std::make_unique<hz_bfs_batch>();
        b->graphs = graphs;
        b->layout(vertices, edges, sources);
        if (b->total_edges && !col) throw std::invalid_argument("hz_bfs_batch_from_csr: null col");
        // Validate the CSR on the host before any kernel reads it.
        for (uint32_t g = 0; g < graphs; ++g) {
            const int32_t* rows = row_ptr + b->row_base[g];
            const int32_t* cols = col + b->edge_base[g];
            if (rows[0] != 0 || (uint64_t)rows[vertices[g]] != edges[g]) {
                throw std::invalid_argument("hz_bfs_batch_from_csr: row_ptr does not span the edges");
            }
            for (uint32_t v = 0; v < vertices[g]; ++v) {
                if (rows[v] > rows[v + 1]) throw std::invalid_argument("hz_bfs_batch_from_csr: row_ptr decreases");
            }
            for (uint64_t i = 0; i < edges[g]; ++i) {
                if (cols[i] < 0 || (uint32_t)cols[i] >= vertices[g]) {
                    throw std::invalid_argument("hz_bfs_batch_from_csr: column index out of range");
                }
            }
        }
        check(hipMemcpy(b->row_ptr.p, row_ptr, (b->total_vertices + graphs) * 4, hipMemcpyHostToDevice), "row_ptr");
        if (b->total_edges) check(hipMemcpy(b->col.p, col, b->total_edges * 4, hipMemcpyHostToDevice), "col");
        b->write_jobs();
        *out = b.release();
    });
}

extern "C" int hz_bfs_batch_destroy(hz_bfs_batch* batch) {
    return guarded([&] { delete batch; });
}

extern "C" int hz_bfs_batch_info(const hz_bfs_batch* batch, hz_bfs_info* info) {
    return guarded([&] {
        check_batch(batch);
        if (!info) throw std::invalid_argument("hz_bfs_batch_info: null output");
        std::memset(info, 0, sizeof(*info));
        info->graphs = batch->graphs;
        info->total_vertices = batch->total_vertices;
        info->total_edges = batch->total_edges;
        info->d_row_ptr = batch->row_ptr.as<int32_t>();
        info->d_col = batch->col.as<int32_t>();
        info->d_level = batch->level.as<int32_t>();
        info->build_ms = batch->build_ms;
    });
}

extern "C" int hz_bfs_batch_layout(const hz_bfs_batch* batch, uint32_t* vertices, uint64_t* edges,
                                   uint64_t* vertex_base, uint64_t* edge_base, uint64_t* row_base) {
    return guarded([&] {
        check_batch(batch);
        for (uint32_t g = 0; g < batch->graphs; ++g) {
            if (vertices) vertices[g] = batch->vertices[g];
            if (edges) edges[g] = batch->edges[g];
            if (vertex_base) vertex_base[g] = batch->vertex_base[g];
            if (edge_base) edge_base[g] = batch->edge_base[g];
            if (row_base) row_base[g] = batch->row_base[g];
        }
    });
}

extern "C" int hz_bfs_batch_run(hz_bfs_batch* batch, int kind, const hz_run_config* config, hz_run_stats* stats) {
    return guarded([&] {
        hz_run_config c{};
        if (config) c = *config;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (c.spec_param > BFS_MAX_SPLIT) {
            throw std::invalid_argument("hz_bfs_batch_run: the split threshold (spec_param) must be in [1, 2^24] "
                                        "(0 = default 64)");
        }
        check_batch(batch);
        if (kind == HZ_EVENT) {
            if (c.flags & HZ_RUN_GRAPH) throw std::invalid_argument("hz_bfs_batch_run: graph launch is for HZ_ASYNC");
            run_event(*batch, c, stats);
        } else if (kind == HZ_ASYNC) {
            auto key = runner_key(c);
            auto it = batch->runners.find(key);
            if (it == batch->runners.end()) it = batch->runners.emplace(key, make_async_runner(*batch, c)).first;
            it->second->run(stats, (c.flags & HZ_RUN_AUDIT) != 0);
        } else {
            throw std::invalid_argument("hz_bfs_batch_run: kind must be HZ_EVENT or HZ_ASYNC");
        }
    });
}

extern "C" int hz_bfs_batch_release(hz_bfs_batch* batch) {
    return guarded([&] {
        check_batch(batch);
        batch->runners.clear();
    });
}
