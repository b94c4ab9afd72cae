// Serving loop of the C ABI: a request server that admits Collatz, photon-walk
// and BFS jobs as they arrive and runs them concurrently on one GPU.
//
// Continuous batching, gfx950-first:
//  - the device is shared by `slots` concurrent units of work, each on its
//    own non-blocking HIP stream (one hardware queue each); a unit's
//    persistent AsyncProgram kernel is launched with the slot's share of the
//    resident wavefronts, so the units of different slots run side by side;
//  - a large Collatz or photon job is cut into up to `slots` chunks of
//    contiguous work items (ids stay global, so the results are those of the
//    whole job) that run on different slots; a BFS job is one unit and runs
//    under the level-synchronous EventProgram, which suits a single search;
//  - each slot keeps its program instances (arena, queues, status) resident
//    from server creation, so admitting a job allocates nothing on the device;
//  - job state (inputs, outputs, graph, sort scratch) lives in one device pool
//    carved by a host-side first-fit allocator; a pending job is admitted, in
//    arrival order, as soon as a slot is free and its state fits in the pool;
//  - completion is detected without blocking: each launch ends with an
//    asynchronous copy of the program status into pinned host memory and an
//    event; the server polls the event, re-launches a pass that ended early,
//    and otherwise records the job as done and frees its slot.
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "common.h"
#include "pool.h"
#include "collatz.h"
#include "collatz_kernels.h"
#include "graph.h"
#include "photon.h"
#include "photon_state.h"

using hz::capi::check;
using hz::capi::guarded;
using hz::capi::Pool;

namespace {

using SCollatz = AsyncProgram<CollatzSpec>;
using SPhoton = AsyncProgram<PhotonSpec>;
using SBfs = EventProgram<BfsSpecT<1>>;

constexpr uint32_t COLLATZ_BITS = 32;      // serving Collatz starts are uniform in [1, 2^32)
constexpr uint32_t PHOTON_REPLICAS = 64;   // tally copies per photon job

// Depth bins of a serving photon job: the workload default, the same value
// the launch zeroes and the job's tally layout reserves.
inline uint32_t photon_job_bins() {
    hz_photon_params p;
    hz_photon_defaults(&p);
    return p.bins;
}
constexpr uint32_t BFS_EDGE_FACTOR = 16;
constexpr uint64_t ALIGN = 256;
constexpr uint32_t CHUNK_MIN_ITEMS = 1u << 18;  // smallest chunk a job is cut into
constexpr uint64_t BFS_EVENT_IO = 1ull << 23;    // per-operation queue capacity of a slot's BFS program
// Generations enqueued between completion checks of a BFS job. R-MAT
// searches take 10-13 generations, so 12 finish most jobs in one pass; the
// empty generations after the last cost a launch each. Against 4: peak
// 1,049-1,075 -> 1,090-1,093 jobs/s, p99 at 80% of peak 20.4-21.8 ->
// 19.5-20.0 ms.
#ifndef HZ_SERVE_BFS_GENERATIONS
#define HZ_SERVE_BFS_GENERATIONS 12
#endif
constexpr uint32_t BFS_GENERATIONS_PER_PASS = HZ_SERVE_BFS_GENERATIONS;

uint64_t align_up(uint64_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

double now_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Device memory a job needs while it is resident, by type.
struct Layout {
    uint64_t bytes = 0;
    uint64_t a = 0, b = 0, c = 0, d = 0, e = 0, f = 0, g = 0;  // sub-buffer offsets
    uint64_t sort_temp = 0;                                   // offset of the BFS sort scratch
};

size_t bfs_sort_temp(uint32_t edges) {
    size_t temp = 0;
    check(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (uint32_t*)nullptr, (int)edges, 0, 32),
          "radix sort sizing");
    return temp;
}

Layout layout_for(const hz_job& j, uint32_t chunks) {
    Layout L;
    uint64_t n = j.size;
    auto take = [&](uint64_t bytes) {
        uint64_t off = L.bytes;
        L.bytes += align_up(bytes ? bytes : 1);
        return off;
    };
    switch (j.type) {
        case HZ_JOB_COLLATZ:
            L.a = take(8 * n);  // starts
            L.b = take(4 * n);  // steps
            break;
        case HZ_JOB_PHOTON:
            L.a = take(4 * n);  // events
            L.b = take((uint64_t)chunks * PHOTON_REPLICAS * (photon_job_bins() + HZ_PHOTON_EXTRA_SLOTS) * 8);  // tally per chunk
            break;
        case HZ_JOB_BFS: {
            uint64_t E = n * BFS_EDGE_FACTOR;
            L.a = take(4 * (n + 1));  // row_ptr
            L.b = take(4 * E);        // col
            L.c = take(4 * n);        // level
            L.d = take(4 * E);        // sources, generation order
            L.e = take(4 * E);        // targets, generation order
            L.f = take(4 * E);        // sources, sorted
            L.g = take(sizeof(BfsJob));
            L.sort_temp = take(bfs_sort_temp((uint32_t)E));  // radix-sort scratch
            break;
        }
        default:
            throw std::invalid_argument("hz_server: unknown job type");
    }
    return L;
}

__global__ void set_iter_kernel(util::iter::AtomicIter<unsigned int>* it, unsigned int first, unsigned int limit) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        it->value = first;
        it->limit = limit;
    }
}

__global__ void set_bfs_job_kernel(BfsJob* job, const int32_t* row_ptr, const int32_t* col, int32_t* level,
                                   uint32_t vertices) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        job->row_ptr = row_ptr;
        job->col = col;
        job->level = level;
        job->vertices = vertices;
        job->source = 0u;
    }
}

__global__ void __launch_bounds__(256) zero_kernel(uint32_t* p, uint64_t words) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0u;
}

struct Job {
    uint64_t id = 0;
    hz_job desc{};
    int32_t status = HZ_JOB_PENDING;
    uint32_t slot = ~0u;
    uint64_t off = Pool::NONE;
    Layout layout;
    double submit_ms = 0, start_ms = 0, end_ms = 0;
    uint64_t promises = 0;
    uint32_t launches = 0, groups = 0;
    uint32_t chunks = 1, chunks_done = 0, chunks_started = 0;
    std::string error;

    uint64_t first(uint32_t c) const { return (uint64_t)desc.size * c / chunks; }
    uint64_t end(uint32_t c) const { return (uint64_t)desc.size * (c + 1) / chunks; }
};

uint32_t event_share(uint32_t slots) {
    // 16 groups per CU in all: the BFS generations' fastest grid (measured).
    uint32_t g = (uint32_t)hz::capi::compute_units() * 16u / slots;
    return g < 4 ? 4 : g;
}

template <typename P>
uint32_t share_of(uint32_t slots) {
    uint32_t g = (uint32_t)hz::capi::resident_groups_of<P>() / slots;
    uint32_t per = (uint32_t)(P::WORK_GROUP_SIZE / 64);
    g = g / per * per;
    return g < per ? per : g;
}

}  // namespace

struct hz_server {
    hz_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    void* pool_base = nullptr;
    std::unique_ptr<Pool> pool;
    std::deque<Job> jobs;  // by id
    std::deque<std::pair<uint64_t, uint32_t>> pending;  // (job, chunk) in arrival order
    std::deque<uint64_t> finished;                      // jobs finished since the last poll
    uint32_t groups_collatz = 0, groups_photon = 0, groups_bfs = 0;

    struct Slot {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        hz::AsyncStatus* status = nullptr;      // pinned host copy (async programs)
        SBfs::Status* event_status = nullptr;   // pinned host copy (the BFS event program)
        hz::capi::DeviceIter iter{0};
        std::unique_ptr<SCollatz::Instance> collatz;
        std::unique_ptr<SPhoton::Instance> photon;
        std::unique_ptr<SBfs::Instance> bfs;
        int64_t job = -1;
        uint32_t chunk = 0;
    };
    std::vector<std::unique_ptr<Slot>> slots;

    ~hz_server() {
        for (auto& s : slots) {
            if (s->stream) (void)hipStreamSynchronize(s->stream);
        }
        for (auto& s : slots) {
            if (s->done) (void)hipEventDestroy(s->done);
            if (s->stream) (void)hipStreamDestroy(s->stream);
            if (s->status) (void)hipHostFree(s->status);
            if (s->event_status) (void)hipHostFree(s->event_status);
        }
        slots.clear();
        if (pool_base) (void)hipFree(pool_base);
    }

    char* mem(const Job& j, uint64_t sub) const { return (char*)pool_base + j.off + sub; }

    uint32_t job_groups(uint32_t share, uint64_t items) const {
        uint64_t want = (items + 255) / 256;  // a group per four links of initial work
        if (want < 4) want = 4;
        return (uint32_t)(want < share ? want : share);
    }

    // Enqueues one exec pass of job j's program on slot s, then the status copy and the event.
    template <typename P>
    void enqueue_exec(Slot& s, typename P::Instance& inst, uint32_t groups) {
        exec<P>(inst, groups, 1u << 24, s.stream);
        check(hipGetLastError(), "serve exec launch");
        check(hipMemcpyAsync(s.status, inst.status.data(), sizeof(hz::AsyncStatus), hipMemcpyDeviceToHost, s.stream),
              "serve status copy");
        check(hipEventRecord(s.done, s.stream), "serve event");
    }

    // BFS_GENERATIONS_PER_PASS generations of the slot's BFS program, then the status copy and the event.
    void enqueue_generations(Slot& s, uint32_t groups) {
        for (uint32_t g = 0; g < BFS_GENERATIONS_PER_PASS; ++g) {
            exec<SBfs>(*s.bfs, groups, 4u, s.stream);
            check(hipGetLastError(), "serve generation launch");
        }
        check(hipMemcpyAsync(s.event_status, s.bfs->status.data(), sizeof(SBfs::Status), hipMemcpyDeviceToHost,
                             s.stream),
              "serve status copy");
        check(hipEventRecord(s.done, s.stream), "serve event");
    }

    // Enqueues chunk c of job j on slot s.
    void launch(Slot& s, Job& j, uint32_t c) {
        const uint64_t first = j.first(c), end = j.end(c), count = end - first;
        switch (j.desc.type) {
            case HZ_JOB_COLLATZ: {
                auto* starts = (unsigned long long*)mem(j, j.layout.a);
                auto* steps = (unsigned int*)mem(j, j.layout.b);
                hipLaunchKernelGGL(hz::collatz::collatz_starts_range_kernel, dim3(hz::graph::grid_for(count)),
                                   dim3(256), 0, s.stream, (unsigned long long)j.desc.seed, COLLATZ_BITS,
                                   (unsigned int)first, (unsigned int)count, starts);
                hipLaunchKernelGGL(set_iter_kernel, dim3(1), dim3(64), 0, s.stream, s.iter.ptr, (unsigned int)first,
                                   (unsigned int)end);
                s.collatz->device_state = CollatzState{starts, steps, s.iter.ptr};
                uint32_t groups = job_groups(groups_collatz, count);
                j.groups = groups > j.groups ? groups : j.groups;
                init<SCollatz>(*s.collatz, groups, s.stream);
                enqueue_exec<SCollatz>(s, *s.collatz, groups);
                break;
            }
            case HZ_JOB_PHOTON: {
                hz_photon_params p{};
                hz_photon_defaults(&p);
                p.photons = j.desc.size;
                p.seed = j.desc.seed;
                p.replicas = PHOTON_REPLICAS;
                auto* events = (uint32_t*)mem(j, j.layout.a);
                const uint64_t tally_doubles = (uint64_t)PHOTON_REPLICAS * (p.bins + HZ_PHOTON_EXTRA_SLOTS);
                auto* tally = (double*)mem(j, j.layout.b) + tally_doubles * c;
                hipLaunchKernelGGL(zero_kernel, dim3(hz::graph::grid_for(count)), dim3(256), 0, s.stream,
                                   events + first, count);
                hipLaunchKernelGGL(zero_kernel, dim3(hz::graph::grid_for(2 * tally_doubles)), dim3(256), 0,
                                   s.stream, (uint32_t*)tally, 2 * tally_doubles);
                hipLaunchKernelGGL(set_iter_kernel, dim3(1), dim3(64), 0, s.stream, s.iter.ptr, (unsigned int)first,
                                   (unsigned int)end);
                s.photon->device_state = hz::photon_abi::make_state(p, tally, events, s.iter.ptr);
                uint32_t groups = job_groups(groups_photon, count);
                j.groups = groups > j.groups ? groups : j.groups;
                init<SPhoton>(*s.photon, groups, s.stream);
                enqueue_exec<SPhoton>(s, *s.photon, groups);
                break;
            }
            case HZ_JOB_BFS: {
                const uint32_t V = j.desc.size;
                const uint32_t E = V * BFS_EDGE_FACTOR;
                const uint32_t scale = hz::rmat::scale_of(V);
                auto* row_ptr = (int32_t*)mem(j, j.layout.a);
                auto* col = (int32_t*)mem(j, j.layout.b);
                auto* level = (int32_t*)mem(j, j.layout.c);
                auto* src_in = (uint32_t*)mem(j, j.layout.d);
                auto* dst_in = (uint32_t*)mem(j, j.layout.e);
                auto* src_out = (uint32_t*)mem(j, j.layout.f);
                auto* job = (BfsJob*)mem(j, j.layout.g);
                void* temp = mem(j, j.layout.sort_temp);
                size_t temp_bytes = bfs_sort_temp(E);
                hipLaunchKernelGGL(hz::graph::rmat_edges_kernel, dim3(hz::graph::grid_for(E)), dim3(256), 0, s.stream,
                                   0u, E, V, scale, (uint32_t)j.desc.seed, (uint32_t)(j.desc.seed >> 32), src_in,
                                   dst_in);
                check(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, src_in, src_out, dst_in, (uint32_t*)col,
                                                         (int)E, 0, scale > 0 ? (int)scale : 1, s.stream),
                      "serve radix sort");
                hipLaunchKernelGGL(hz::graph::row_ptr_kernel, dim3(hz::graph::grid_for((uint64_t)V + 1)), dim3(256),
                                   0, s.stream, src_out, E, V, row_ptr);
                hipLaunchKernelGGL(hz::graph::bfs_reset_kernel, dim3(hz::graph::grid_for(V)), dim3(256), 0, s.stream,
                                   level, (uint64_t)V, s.iter.ptr, 1u);
                hipLaunchKernelGGL(set_bfs_job_kernel, dim3(1), dim3(64), 0, s.stream, job, row_ptr, col, level, V);
                BfsState st{};
                st.jobs = job;
                st.job_iter = s.iter.ptr;
                st.job_count = 1;
                st.split = BFS_DEFAULT_SPLIT;
                s.bfs->device_state = st;
                s.bfs->reset(s.stream);
                // Generations are ordinary (non-persistent) launches: a job may
                // ask for the whole device and the hardware shares it between
                // the slots' streams.
                j.groups = job_groups(groups_bfs * cfg.slots, V);
                enqueue_generations(s, j.groups);
                break;
            }
        }
        j.launches++;
    }

    // The pass that just finished: 1 = chunk done, 0 = more work enqueued,
    // throws on a runtime error.
    int finish_pass(Slot& s, Job& j) {
        if (j.launches >= 1000000u) throw std::runtime_error("serve: launch bound reached");
        if (j.desc.type == HZ_JOB_BFS) {
            const SBfs::Status& st = *s.event_status;
            if (st.flags & SBfs::OVERFLOW_FLAG) throw std::runtime_error("serve: a BFS generation overflowed its queue");
            if (st.flags & SBfs::COMPLETION_FLAG) {
                j.promises += st.promises;
                return 1;
            }
            j.launches++;
            enqueue_generations(s, j.groups);
            return 0;
        }
        const hz::AsyncStatus& st = *s.status;
        if (st.flags & SCollatz::ARENA_EXHAUSTED_FLAG) {
            throw std::runtime_error("serve: the slot's link arena was exhausted (raise arena_links)");
        }
        if (st.flags & SCollatz::QUEUE_STALL_FLAG) {
            throw std::runtime_error(std::string("serve: a bounded wait gave up (") + hz::stall_name(st.stall_site) +
                                     ")");
        }
        if (st.flags & SCollatz::COMPLETION_FLAG) {
            j.promises += st.promises;
            return 1;
        }
        j.launches++;
        const uint32_t groups = j.groups;
        if (j.desc.type == HZ_JOB_COLLATZ) {
            enqueue_exec<SCollatz>(s, *s.collatz, groups);
        } else {
            enqueue_exec<SPhoton>(s, *s.photon, groups);
        }
        return 0;
    }

    void release_memory(Job& j) {
        if (j.off != Pool::NONE) {
            pool->free(j.off, j.layout.bytes);
            j.off = Pool::NONE;
        }
    }

    // The pending unit to admit next: the oldest (HZ_SERVE_FIFO), or the one
    // of the largest job, oldest first among equals (HZ_SERVE_LARGEST_FIRST);
    // a job's chunks stay in order either way.
    size_t next_unit() const {
        if (cfg.policy != HZ_SERVE_LARGEST_FIRST) return 0;
        size_t best = 0;
        for (size_t i = 1; i < pending.size(); ++i) {
            const Job& a = jobs[(size_t)pending[i].first];
            const Job& b = jobs[(size_t)pending[best].first];
            if (a.desc.size > b.desc.size) best = i;
        }
        return best;
    }

    // Reaps finished passes and admits pending chunks into free slots.
    void progress() {
        for (auto& sp : slots) {
            Slot& s = *sp;
            if (s.job < 0) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            Job& j = jobs[(size_t)s.job];
            check(q, "serve event query");
            int done = 0;
            try {
                done = finish_pass(s, j);
            } catch (const std::exception& e) {
                if (j.status != HZ_JOB_FAILED) {
                    j.status = HZ_JOB_FAILED;
                    j.error = e.what();
                    j.end_ms = now_ms(t0);
                }
                done = -1;  // the next unit's init clears the instance's flags
            }
            if (done == 0) continue;
            s.job = -1;
            j.chunks_done++;
            if (j.chunks_done == j.chunks_started && (j.chunks_done == j.chunks || j.status == HZ_JOB_FAILED)) {
                if (j.status != HZ_JOB_FAILED) {
                    j.status = HZ_JOB_DONE;
                    j.end_ms = now_ms(t0);
                }
                if (!cfg.keep_outputs || j.status == HZ_JOB_FAILED) release_memory(j);
                finished.push_back(j.id);
            }
        }
        for (auto& sp : slots) {
            Slot& s = *sp;
            if (s.job >= 0) continue;
            while (!pending.empty()) {
                size_t pick = next_unit();
                auto [id, c] = pending[pick];
                Job& j = jobs[(size_t)id];
                if (j.status == HZ_JOB_FAILED) {  // drop the rest of a failed job
                    pending.erase(pending.begin() + (std::ptrdiff_t)pick);
                    continue;
                }
                if (j.off == Pool::NONE) {
                    uint64_t off = pool->alloc(j.layout.bytes);
                    if (off == Pool::NONE) return;  // the chosen unit waits for memory; nothing overtakes it
                    j.off = off;
                    j.status = HZ_JOB_RUNNING;
                    j.start_ms = now_ms(t0);
                    j.slot = (uint32_t)(&sp - &slots[0]);
                }
                pending.erase(pending.begin() + (std::ptrdiff_t)pick);
                s.job = (int64_t)j.id;
                s.chunk = c;
                j.chunks_started++;
                try {
                    launch(s, j, c);
                } catch (...) {
                    j.status = HZ_JOB_FAILED;
                    j.end_ms = now_ms(t0);
                    s.job = -1;
                    j.chunks_done++;
                    if (j.chunks_done == j.chunks_started) {
                        release_memory(j);
                        finished.push_back(j.id);
                    }
                    throw;
                }
                break;
            }
        }
    }

    uint64_t submit(const hz_job& d, double submit_ms) {
        if (d.type > HZ_JOB_BFS) throw std::invalid_argument("hz_server_submit: unknown job type");
        if (d.size == 0 || d.size > (1u << 26)) throw std::invalid_argument("hz_server_submit: size must be in [1, 2^26]");
        if (d.type == HZ_JOB_BFS && d.size > (1u << 22)) {
            throw std::invalid_argument("hz_server_submit: BFS jobs have at most 2^22 vertices");
        }
        Job j;
        j.id = jobs.size();
        j.desc = d;
        if (d.type != HZ_JOB_BFS) {
            uint64_t k = (d.size + CHUNK_MIN_ITEMS - 1) / CHUNK_MIN_ITEMS;
            j.chunks = (uint32_t)(k < cfg.slots ? k : cfg.slots);
            if (j.chunks == 0) j.chunks = 1;
        }
        j.layout = layout_for(d, j.chunks);
        if (j.layout.bytes > pool->size()) throw std::invalid_argument("hz_server_submit: job larger than the pool");
        j.submit_ms = submit_ms;
        if (cfg.max_pending && pending.size() >= cfg.max_pending) {
            throw std::runtime_error("hz_server_submit: too many pending jobs");
        }
        jobs.push_back(j);
        for (uint32_t c = 0; c < j.chunks; ++c) pending.emplace_back(j.id, c);
        return j.id;
    }

    bool idle() const {
        if (!pending.empty()) return false;
        for (auto& s : slots) {
            if (s->job >= 0) return false;
        }
        return true;
    }

    void fill_result(const Job& j, hz_job_result* r) const {
        std::memset(r, 0, sizeof(*r));
        r->id = j.id;
        r->type = j.desc.type;
        r->size = j.desc.size;
        r->seed = j.desc.seed;
        r->status = j.status;
        r->slot = j.slot;
        r->promises = j.promises;
        r->launches = j.launches;
        r->groups = j.groups;
        r->submit_ms = j.submit_ms;
        r->start_ms = j.start_ms;
        r->end_ms = j.end_ms;
        r->state_bytes = j.layout.bytes;
        r->chunks = j.chunks;
    }

    Job& job(uint64_t id) {
        if (id >= jobs.size()) throw std::invalid_argument("hz_server: unknown job id");
        return jobs[(size_t)id];
    }
};

namespace {
void check_server(const hz_server* s) {
    if (!s) throw std::invalid_argument("hz_server: null server");
}
}  // namespace

extern "C" int hz_server_create(const hz_server_config* config, hz_server** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("hz_server_create: null output");
        *out = nullptr;
        auto s = std::make_unique<hz_server>();
        if (config) s->cfg = *config;
        hz_server_config& c = s->cfg;
        if (c.pool_bytes == 0) c.pool_bytes = 8ull << 30;
        // Three slots by default: a process has four hardware queues, and the
        // default stream takes one (measured: 3 slots gave ~1080 jobs/s at
        // peak and a p99 of 7.7 ms at 20% load, against ~975 and 9.7 with 4).
        if (c.slots == 0) c.slots = 3;
        if (c.slots > 8) throw std::invalid_argument("hz_server_create: at most 8 slots");
        if (c.arena_links == 0) c.arena_links = 1ull << 16;
        s->t0 = std::chrono::steady_clock::now();
        check(hipMalloc(&s->pool_base, c.pool_bytes), "hipMalloc(serve pool)");
        s->pool = std::make_unique<Pool>(c.pool_bytes);
        s->groups_collatz = share_of<SCollatz>(c.slots);
        s->groups_photon = share_of<SPhoton>(c.slots);
        s->groups_bfs = event_share(c.slots);
        for (uint32_t i = 0; i < c.slots; ++i) {
            auto sl = std::make_unique<hz_server::Slot>();
            check(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking), "hipStreamCreate");
            check(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming), "hipEventCreate");
            check(hipHostMalloc((void**)&sl->status, sizeof(hz::AsyncStatus), hipHostMallocDefault), "hipHostMalloc");
            std::memset(sl->status, 0, sizeof(hz::AsyncStatus));
            check(hipHostMalloc((void**)&sl->event_status, sizeof(SBfs::Status), hipHostMallocDefault),
                  "hipHostMalloc");
            std::memset(sl->event_status, 0, sizeof(SBfs::Status));
            sl->collatz = std::make_unique<SCollatz::Instance>(c.arena_links, CollatzState{}, 0, 64u);
            sl->photon = std::make_unique<SPhoton::Instance>(c.arena_links, PhotonState{}, 0, 64u);
            sl->bfs = std::make_unique<SBfs::Instance>(BFS_EVENT_IO, BfsState{});
            s->slots.push_back(std::move(sl));
        }
        check(hipDeviceSynchronize(), "serve setup");
        *out = s.release();
    });
}

extern "C" int hz_server_destroy(hz_server* server) {
    return guarded([&] { delete server; });
}

extern "C" int hz_server_submit(hz_server* server, const hz_job* job, uint64_t* id) {
    return guarded([&] {
        check_server(server);
        if (!job) throw std::invalid_argument("hz_server_submit: null job");
        uint64_t i = server->submit(*job, now_ms(server->t0));
        if (id) *id = i;
        server->progress();
    });
}

extern "C" int hz_server_progress(hz_server* server, uint32_t* running, uint32_t* pending) {
    return guarded([&] {
        check_server(server);
        server->progress();
        uint32_t r = 0;
        for (auto& s : server->slots) r += s->job >= 0 ? 1u : 0u;
        if (running) *running = r;
        if (pending) *pending = (uint32_t)server->pending.size();
    });
}

extern "C" int hz_server_drain(hz_server* server, double timeout_ms) {
    return guarded([&] {
        check_server(server);
        double t_end = now_ms(server->t0) + (timeout_ms > 0 ? timeout_ms : 600000.0);
        while (true) {
            server->progress();
            if (server->idle()) return;
            if (now_ms(server->t0) > t_end) throw std::runtime_error("hz_server_drain: timed out");
        }
    });
}

extern "C" int hz_server_job(hz_server* server, uint64_t id, hz_job_result* result) {
    return guarded([&] {
        check_server(server);
        if (!result) throw std::invalid_argument("hz_server_job: null output");
        const Job& j = server->job(id);
        server->fill_result(j, result);
        if (j.status == HZ_JOB_FAILED) throw std::runtime_error("job " + std::to_string(id) + " failed: " + j.error);
    });
}

extern "C" int hz_server_output(hz_server* server, uint64_t id, uint32_t which, void* host, uint64_t bytes) {
    return guarded([&] {
        check_server(server);
        Job& j = server->job(id);
        if (j.status != HZ_JOB_DONE) throw std::invalid_argument("hz_server_output: the job is not done");
        if (j.off == Pool::NONE) throw std::invalid_argument("hz_server_output: outputs were released");
        uint64_t n = j.desc.size, sub = 0, want = 0;
        switch (j.desc.type) {
            case HZ_JOB_COLLATZ:
                if (which == 0) sub = j.layout.b, want = 4 * n;          // steps
                else if (which == 1) sub = j.layout.a, want = 8 * n;     // starts
                break;
            case HZ_JOB_PHOTON:
                if (which == 0) sub = j.layout.a, want = 4 * n;          // events
                else if (which == 1) sub = j.layout.b, want = (uint64_t)j.chunks * PHOTON_REPLICAS * (photon_job_bins() + HZ_PHOTON_EXTRA_SLOTS) * 8;
                break;
            case HZ_JOB_BFS:
                if (which == 0) sub = j.layout.c, want = 4 * n;          // levels
                else if (which == 1) sub = j.layout.a, want = 4 * (n + 1);  // row_ptr
                else if (which == 2) sub = j.layout.b, want = 4 * n * BFS_EDGE_FACTOR;  // col
                break;
        }
        if (want == 0) throw std::invalid_argument("hz_server_output: no such output");
        if (!host || bytes != want) throw std::invalid_argument("hz_server_output: buffer size must be " + std::to_string(want));
        check(hipMemcpy(host, server->mem(j, sub), want, hipMemcpyDeviceToHost), "hz_server_output copy");
    });
}

extern "C" int hz_server_release(hz_server* server, uint64_t id) {
    return guarded([&] {
        check_server(server);
        Job& j = server->job(id);
        if (j.status == HZ_JOB_RUNNING || j.status == HZ_JOB_PENDING) {
            throw std::invalid_argument("hz_server_release: the job has not finished");
        }
        server->release_memory(j);
    });
}

extern "C" int hz_server_info(const hz_server* server, hz_server_info_t* info) {
    return guarded([&] {
        check_server(server);
        if (!info) throw std::invalid_argument("hz_server_info: null output");
        std::memset(info, 0, sizeof(*info));
        info->slots = server->cfg.slots;
        info->pool_bytes = server->pool->size();
        info->pool_used = server->pool->used();
        info->pool_largest_free = server->pool->largest_free();
        info->jobs = server->jobs.size();
        info->pending = (uint32_t)server->pending.size();
        info->groups_collatz = server->groups_collatz;
        info->groups_photon = server->groups_photon;
        info->groups_bfs = server->groups_bfs;
    });
}

extern "C" int hz_server_run_trace(hz_server* server, const hz_job* jobs, const double* arrival_ms, uint32_t count,
                                   double timeout_ms, hz_job_result* results) {
    return guarded([&] {
        check_server(server);
        if (count && (!jobs || !arrival_ms)) throw std::invalid_argument("hz_server_run_trace: null trace");
        for (uint32_t i = 1; i < count; ++i) {
            if (arrival_ms[i] < arrival_ms[i - 1]) throw std::invalid_argument("hz_server_run_trace: arrivals must be sorted");
        }
        std::vector<uint64_t> ids(count);
        const double base = now_ms(server->t0);
        const double t_end = base + (timeout_ms > 0 ? timeout_ms : 600000.0);
        uint32_t next = 0;
        while (true) {
            double t = now_ms(server->t0);
            while (next < count && base + arrival_ms[next] <= t) {
                // Latency is counted from the scheduled arrival, not from when
                // this loop got to it.
                ids[next] = server->submit(jobs[next], base + arrival_ms[next]);
                next++;
            }
            server->progress();
            if (next == count && server->idle()) break;
            if (t > t_end) throw std::runtime_error("hz_server_run_trace: timed out");
        }
        if (results) {
            for (uint32_t i = 0; i < count; ++i) server->fill_result(server->jobs[(size_t)ids[i]], &results[i]);
        }
    });
}

// Host-only access to the pool allocator, for tests of the admission logic.
extern "C" int hz_pool_test(uint64_t size, const int64_t* ops, uint32_t count, uint64_t* out) {
    return guarded([&] {
        Pool p(size);
        std::vector<std::pair<uint64_t, uint64_t>> live;
        for (uint32_t i = 0; i < count; ++i) {
            int64_t op = ops[i];
            if (op >= 0) {
                uint64_t off = p.alloc((uint64_t)op);
                out[i] = off;
                live.emplace_back(off, (uint64_t)op);
            } else {
                size_t k = (size_t)(-op - 1);
                if (k >= live.size() || live[k].first == Pool::NONE) throw std::invalid_argument("pool test: bad free");
                p.free(live[k].first, live[k].second);
                live[k].first = Pool::NONE;
                out[i] = p.largest_free();
            }
        }
    });
}

// The serving entry points under the names of a request/poll interface.
extern "C" int hz_serve_submit(hz_server* server, const hz_job* job, uint64_t* id) {
    return hz_server_submit(server, job, id);
}

extern "C" int hz_serve_poll(hz_server* server, uint64_t* finished_ids, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        check_server(server);
        if (!count || (capacity && !finished_ids)) throw std::invalid_argument("hz_serve_poll: null output");
        server->progress();
        uint32_t n = 0;
        while (n < capacity && !server->finished.empty()) {
            finished_ids[n++] = server->finished.front();
            server->finished.pop_front();
        }
        *count = n;
    });
}
