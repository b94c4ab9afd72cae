// Resident serving of the C ABI: Collatz or photon-walk jobs admitted into a
// running persistent kernel through a doorbell ring, no kernel launch per job.
//
// One AsyncProgram instance of the server's kind (ResidentCollatzSpec or
// ResidentPhotonSpec) holds the device. The host
// gives each admitted job a job-table entry (at most JOBS in flight) and a
// slice of the device pool for its outputs, writes its descriptor into the
// host-coherent ring and advances the ring head. If no launch is running, it
// launches one; a running launch installs the descriptor at its next
// housekeeping pass and its groups start claiming the job's items. The launch
// reports each completion in the ring's completion words, which the host
// polls without any HIP call, and ends once it has had no job for idle_us (or
// once idle after hz_resident_drain); the job table persists, so the next
// launch carries on.
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "common.h"
#include "photon_state.h"
#include "pool.h"
#include "resident.h"

using hz::capi::check;
using hz::capi::guarded;
using hz::capi::Pool;

namespace {

using RCollatz = AsyncProgram<ResidentCollatzSpec>;
using RPhoton = AsyncProgram<ResidentPhotonSpec>;
namespace R = hz::resident;

constexpr uint32_t PHOTON_REPLICAS = 64;  // tally copies per photon job

// The server's program, of either kind.
struct Program {
    virtual ~Program() = default;
    virtual void init(uint32_t groups, hipStream_t stream) = 0;
    virtual void exec(uint32_t groups, hipStream_t stream) = 0;
    virtual const hz::AsyncStatus* status() const = 0;
    virtual uint32_t resident_groups() const = 0;
};

template <typename P>
struct ProgramOf final : Program {
    typename P::Instance inst;
    ProgramOf(uint64_t arena_links, const ResidentState& st) : inst(arena_links, st, 0, 64u) {}
    void init(uint32_t groups, hipStream_t stream) override { ::init<P>(inst, groups, stream); }
    void exec(uint32_t groups, hipStream_t stream) override { ::exec<P>(inst, groups, 0xFFFFFFFFu, stream); }
    const hz::AsyncStatus* status() const override { return inst.status.data(); }
    uint32_t resident_groups() const override { return (uint32_t)hz::capi::resident_groups_of<P>(); }
};

double now_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Clears the per-launch fields of the shared core before a launch.
__global__ void resident_reset_kernel(ResidentCore* core) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        core->poll_lock = 0u;
        core->stopping = 0u;
        core->last_poll = 0ull;
        core->last_busy = 0ull;
        core->pass_start = 0ull;
    }
}

struct RJob {
    uint64_t id = 0;
    hz_job desc{};
    int32_t status = HZ_JOB_PENDING;
    uint32_t entry = 0;   // ring sequence number
    uint32_t table = 0;   // job-table entry
    uint64_t base = Pool::NONE;  // first item in the item space
    uint64_t items = 0;          // items reserved (the size, rounded up)
    bool joined = false;  // published while a launch was running
    double submit_ms = 0, start_ms = 0, end_ms = 0;
};

}  // namespace

struct hz_resident {
    hz_resident_config cfg{};
    std::chrono::steady_clock::time_point t0;
    uint32_t kind = HZ_JOB_COLLATZ;
    unsigned long long* starts = nullptr;  // item space: Collatz starts
    unsigned int* steps = nullptr;         // item space: Collatz step counts, or photon tallies and event counts
    uint64_t tally_stride = 0;             // photon: doubles per tally
    uint64_t capacity = 0;                 // items
    std::unique_ptr<Pool> pool;            // carves the item space, in items
    hipStream_t stream = nullptr;       // the resident launches
    hipStream_t copy_stream = nullptr;  // output copies, while a launch runs
    hipEvent_t pass_done = nullptr;
    ResidentDoorbell* bell = nullptr;   // host view
    ResidentDoorbell* bell_dev = nullptr;
    ResidentCore* core = nullptr;
    unsigned int* counters = nullptr;
    std::unique_ptr<Program> prog;
    uint32_t groups = 0;
    bool running = false;
    bool resume = false;     // the last launch ended before completion: exec without init
    bool draining = false;
    uint32_t head = 0;
    uint64_t table_free = ~0ull;
    std::deque<RJob> jobs;
    std::deque<uint64_t> pending;
    std::map<uint32_t, uint64_t> in_flight;  // ring entry -> job
    std::deque<uint64_t> finished;
    uint32_t launches = 0, joined = 0;
    hz::AsyncStatus totals{};  // runtime counters summed over ended launches

    ~hz_resident() {
        if (stream) {
            if (bell) __atomic_store_n(&bell->stop, 1u, __ATOMIC_RELEASE);
            (void)hipStreamSynchronize(stream);
        }
        if (copy_stream) (void)hipStreamSynchronize(copy_stream);
        prog.reset();
        if (pass_done) (void)hipEventDestroy(pass_done);
        if (stream) (void)hipStreamDestroy(stream);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        if (core) (void)hipFree(core);
        if (counters) (void)hipFree(counters);
        if (starts) (void)hipFree(starts);
        if (steps) (void)hipFree(steps);
        if (bell) (void)hipHostFree(bell);
    }

    // A photon job's allocation holds its tally (2 x tally_stride words),
    // then its event counts; a Collatz job's, its starts and step counts.
    uint64_t lead() const { return kind == HZ_JOB_PHOTON ? 2ull * tally_stride : 0ull; }
    unsigned long long* starts_of(const RJob& j) const { return starts + j.base; }
    unsigned int* steps_of(const RJob& j) const { return steps + j.base + lead(); }
    double* tally_of(const RJob& j) const { return reinterpret_cast<double*>(steps + j.base); }

    void release_memory(RJob& j) {
        if (j.base != Pool::NONE) {
            pool->free(j.base, j.items);
            j.base = Pool::NONE;
        }
    }

    // A launch that has ended: runtime errors raise; a launch that ended
    // before completing (its lifetime, or the cycle budget) is resumed.
    void reap_launch() {
        if (!running) return;
        hipError_t q = hipEventQuery(pass_done);
        if (q == hipErrorNotReady) return;
        check(q, "resident launch");
        running = false;
        hz::AsyncStatus st{};
        check(hipMemcpyAsync(&st, prog->status(), sizeof(st), hipMemcpyDeviceToHost, stream),
              "resident status copy");
        unsigned int error = 0;
        check(hipMemcpyAsync(&error, &core->error, sizeof(error), hipMemcpyDeviceToHost, stream),
              "resident error copy");
        check(hipStreamSynchronize(stream), "resident status");
        if (st.flags & RCollatz::ARENA_EXHAUSTED_FLAG) {
            throw std::runtime_error("resident: the link arena was exhausted (raise arena_links)");
        }
        if (st.flags & RCollatz::QUEUE_STALL_FLAG) {
            throw std::runtime_error(std::string("resident: a bounded wait gave up (") + hz::stall_name(st.stall_site) +
                                     ")");
        }
        totals.promises += st.promises;
        totals.links_run += st.links_run;
        totals.make_work_calls += st.make_work_calls;
        totals.idle_polls += st.idle_polls;
        for (int i = 0; i < 8; ++i) totals.phase_cycles[i] += st.phase_cycles[i];
        if (error) throw std::runtime_error("resident: the kernel rejected a ring entry");
        resume = (st.flags & RCollatz::COMPLETION_FLAG) == 0u;
    }

    void reap_jobs() {
        for (auto it = in_flight.begin(); it != in_flight.end();) {
            const uint32_t k = it->first;
            if (__atomic_load_n(&bell->done[k % R::RING], __ATOMIC_ACQUIRE) != k + 1u) {
                ++it;
                continue;
            }
            RJob& j = jobs[(size_t)it->second];
            j.status = HZ_JOB_DONE;
            j.end_ms = now_ms(t0);
            table_free |= 1ull << j.table;
            if (!cfg.keep_outputs) release_memory(j);
            finished.push_back(j.id);
            it = in_flight.erase(it);
        }
    }

    // Publishes pending jobs, in arrival order, while the table has room and
    // the pool has memory for the next one.
    void admit() {
        while (!pending.empty() && table_free != 0ull) {
            RJob& j = jobs[(size_t)pending.front()];
            if (j.base == Pool::NONE) {
                uint64_t base = pool->alloc(j.items);
                if (base == Pool::NONE) return;  // waits for item space; nothing overtakes it
                j.base = base;
            }
            pending.pop_front();
            j.table = (uint32_t)__builtin_ctzll(table_free);
            table_free &= ~(1ull << j.table);
            j.entry = head;
            ResidentEntry& e = bell->entries[head % R::RING];
            e.job = j.table;
            e.size = j.desc.size;
            e.seed = j.desc.seed;
            e.base = (unsigned int)(j.base + lead());
            __atomic_store_n(&bell->head, head + 1u, __ATOMIC_RELEASE);
            head += 1u;
            in_flight[j.entry] = j.id;
            j.status = HZ_JOB_RUNNING;
            j.start_ms = now_ms(t0);
            j.joined = running;
            if (running) joined++;
        }
    }

    void launch_if_needed() {
        if (running || in_flight.empty()) return;
        __atomic_store_n(&bell->stop, draining ? 1u : 0u, __ATOMIC_RELEASE);
        hipLaunchKernelGGL(resident_reset_kernel, dim3(1), dim3(64), 0, stream, core);
        check(hipGetLastError(), "resident reset launch");
        if (!resume) prog->init(groups, stream);
        prog->exec(groups, stream);
        check(hipGetLastError(), "resident launch");
        check(hipEventRecord(pass_done, stream), "resident event");
        running = true;
        launches++;
    }

    void progress() {
        reap_launch();
        reap_jobs();
        admit();
        launch_if_needed();
    }

    bool idle() const { return pending.empty() && in_flight.empty() && !running; }

    uint64_t submit(const hz_job& d, double submit_ms) {
        if (d.type != kind) {
            throw std::invalid_argument(kind == HZ_JOB_PHOTON ? "hz_resident_submit: this server runs photon-walk jobs"
                                                              : "hz_resident_submit: this server runs Collatz jobs");
        }
        if (d.size == 0 || d.size > (1u << 26)) throw std::invalid_argument("hz_resident_submit: size must be in [1, 2^26]");
        RJob j;
        j.id = jobs.size();
        j.desc = d;
        j.items = hz::capi::pool_align(lead() + d.size);
        if (j.items > pool->size()) throw std::invalid_argument("hz_resident_submit: job larger than the item space");
        j.submit_ms = submit_ms;
        jobs.push_back(j);
        pending.push_back(j.id);
        return j.id;
    }

    void fill_result(const RJob& j, hz_job_result* r) const {
        std::memset(r, 0, sizeof(*r));
        r->id = j.id;
        r->type = j.desc.type;
        r->size = j.desc.size;
        r->seed = j.desc.seed;
        r->status = j.status;
        r->slot = j.table;
        r->launches = j.joined ? 0u : 1u;
        r->groups = groups;
        r->submit_ms = j.submit_ms;
        r->start_ms = j.start_ms;
        r->end_ms = j.end_ms;
        r->state_bytes = (kind == HZ_JOB_PHOTON ? 4ull : 12ull) * j.items;
        r->chunks = 1;
    }

    RJob& job(uint64_t id) {
        if (id >= jobs.size()) throw std::invalid_argument("hz_resident: unknown job id");
        return jobs[(size_t)id];
    }

    void drain(double timeout_ms) {
        const double t_end = now_ms(t0) + (timeout_ms > 0 ? timeout_ms : 600000.0);
        draining = true;
        __atomic_store_n(&bell->stop, 1u, __ATOMIC_RELEASE);
        try {
            while (true) {
                progress();
                if (idle()) break;
                if (now_ms(t0) > t_end) throw std::runtime_error("hz_resident_drain: timed out");
            }
        } catch (...) {
            draining = false;
            throw;
        }
        draining = false;
    }
};

namespace {
void check_resident(const hz_resident* r) {
    if (!r) throw std::invalid_argument("hz_resident: null server");
}
}  // namespace

extern "C" int hz_resident_create(const hz_resident_config* config, hz_resident** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("hz_resident_create: null output");
        *out = nullptr;
        auto r = std::make_unique<hz_resident>();
        if (config) r->cfg = *config;
        hz_resident_config& c = r->cfg;
        if (c.pool_bytes == 0) c.pool_bytes = 4ull << 30;
        if (c.arena_links == 0) c.arena_links = 1ull << 18;
        if (c.idle_us == 0) c.idle_us = 1000;
        if (c.poll_us == 0) c.poll_us = 2;
        if (c.life_ms == 0) c.life_ms = 10000;
        if (c.bits == 0) c.bits = 32;
        if (c.bits > 40) throw std::invalid_argument("hz_resident_create: bits must be in [1, 40]");
        if (c.policy > R::SMALLEST_FIRST) throw std::invalid_argument("hz_resident_create: unknown policy");
        if (c.kind == HZ_JOB_PHOTON && c.policy == R::SMALLEST_FIRST) {
            // A photon server claiming smallest job first once stopped making
            // progress in a serving sweep, for a reason not yet found; the
            // combination stays refused until it is.
            throw std::invalid_argument("hz_resident_create: photon jobs are claimed in arrival order only");
        }
        if (c.kind != HZ_JOB_COLLATZ && c.kind != HZ_JOB_PHOTON) {
            throw std::invalid_argument("hz_resident_create: kind must be HZ_JOB_COLLATZ or HZ_JOB_PHOTON");
        }
        r->kind = c.kind;
        r->t0 = std::chrono::steady_clock::now();

        int device = 0, khz = 0;
        check(hipGetDevice(&device), "hipGetDevice");
        check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device), "wall clock rate");
        if (khz <= 0) throw std::runtime_error("hz_resident_create: no device wall clock");
        const unsigned long long per_us = (unsigned long long)khz / 1000ull > 0 ? (unsigned long long)khz / 1000ull : 1ull;

        check(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipStreamCreateWithFlags(&r->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreateWithFlags(&r->pass_done, hipEventDisableTiming), "hipEventCreate");
        check(hipHostMalloc((void**)&r->bell, sizeof(ResidentDoorbell), hipHostMallocCoherent | hipHostMallocMapped),
              "hipHostMalloc(doorbell)");
        std::memset(r->bell, 0, sizeof(ResidentDoorbell));
        check(hipHostGetDevicePointer((void**)&r->bell_dev, r->bell, 0), "hipHostGetDevicePointer");
        check(hipMalloc((void**)&r->core, sizeof(ResidentCore)), "hipMalloc(resident core)");
        check(hipMalloc((void**)&r->counters, sizeof(unsigned int) * R::COUNTER_WORDS), "hipMalloc(counters)");
        // The item space, indexed by u32: 12 bytes per Collatz item (a u64
        // start and a u32 count); for photon jobs 4 bytes per photon (its
        // event count) plus the job's tally, 8 x tally_stride bytes.
        const bool photon = r->kind == HZ_JOB_PHOTON;
        r->capacity = c.pool_bytes / (photon ? 4ull : 12ull);
        if (r->capacity > 0xFFFFFF00ull) r->capacity = 0xFFFFFF00ull;
        if (r->capacity < 256) throw std::invalid_argument("hz_resident_create: pool_bytes too small");
        if (!photon) check(hipMalloc((void**)&r->starts, 8ull * r->capacity), "hipMalloc(resident starts)");
        check(hipMalloc((void**)&r->steps, 4ull * r->capacity), "hipMalloc(resident outputs)");
        r->pool = std::make_unique<Pool>(r->capacity);
        hz_photon_params pp{};
        hz_photon_defaults(&pp);
        pp.replicas = PHOTON_REPLICAS;
        if (photon) r->tally_stride = (uint64_t)PHOTON_REPLICAS * (pp.bins + HZ_PHOTON_EXTRA_SLOTS);
        {
            auto host = std::make_unique<ResidentCore>();
            std::memset(host.get(), 0, sizeof(ResidentCore));
            for (unsigned int i = 0; i < R::JOBS; ++i) host->jobs[i].next = R::CLOSED;
            check(hipMemcpy(r->core, host.get(), sizeof(ResidentCore), hipMemcpyHostToDevice), "hipMemcpy(core)");
        }
        check(hipMemset(r->counters, 0, sizeof(unsigned int) * R::COUNTER_WORDS), "hipMemset(counters)");

        ResidentState st{};
        st.core = r->core;
        st.starts = r->starts;
        st.steps = r->steps;
        st.counters = r->counters;
        st.bell = r->bell_dev;
        st.poll_ticks = per_us * c.poll_us;
        st.idle_ticks = per_us * c.idle_us;
        st.life_ticks = per_us * 1000ull * c.life_ms;
        st.policy = c.policy;
        st.bits = c.bits;
        if (photon) {
            st.events = r->steps;
            st.tally_stride = (unsigned int)r->tally_stride;
            st.replicas = PHOTON_REPLICAS;
            st.photon = hz::photon_abi::make_state(pp, nullptr, nullptr, nullptr);
            r->prog = std::make_unique<ProgramOf<RPhoton>>(c.arena_links, st);
        } else {
            r->prog = std::make_unique<ProgramOf<RCollatz>>(c.arena_links, st);
        }
        r->groups = c.groups ? c.groups : r->prog->resident_groups();
        check(hipDeviceSynchronize(), "resident setup");
        *out = r.release();
    });
}

extern "C" int hz_resident_destroy(hz_resident* r) {
    return guarded([&] { delete r; });
}

extern "C" int hz_resident_submit(hz_resident* r, const hz_job* job, uint64_t* id) {
    return guarded([&] {
        check_resident(r);
        if (!job) throw std::invalid_argument("hz_resident_submit: null job");
        uint64_t i = r->submit(*job, now_ms(r->t0));
        if (id) *id = i;
        r->progress();
    });
}

extern "C" int hz_resident_poll(hz_resident* r, uint64_t* finished_ids, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        check_resident(r);
        if (!count || (capacity && !finished_ids)) throw std::invalid_argument("hz_resident_poll: null output");
        r->progress();
        uint32_t n = 0;
        while (n < capacity && !r->finished.empty()) {
            finished_ids[n++] = r->finished.front();
            r->finished.pop_front();
        }
        *count = n;
    });
}

extern "C" int hz_resident_drain(hz_resident* r, double timeout_ms) {
    return guarded([&] {
        check_resident(r);
        r->drain(timeout_ms);
    });
}

extern "C" int hz_resident_job(hz_resident* r, uint64_t id, hz_job_result* result) {
    return guarded([&] {
        check_resident(r);
        if (!result) throw std::invalid_argument("hz_resident_job: null output");
        r->fill_result(r->job(id), result);
    });
}

extern "C" int hz_resident_output(hz_resident* r, uint64_t id, uint32_t which, void* host, uint64_t bytes) {
    return guarded([&] {
        check_resident(r);
        RJob& j = r->job(id);
        if (j.status != HZ_JOB_DONE) throw std::invalid_argument("hz_resident_output: the job is not done");
        if (j.base == Pool::NONE) throw std::invalid_argument("hz_resident_output: outputs were released");
        const uint64_t n = j.desc.size;
        const void* src = nullptr;
        uint64_t want = 0;
        if (which == 0) {
            src = r->steps_of(j), want = 4 * n;  // step counts or event counts
        } else if (which == 1 && r->kind == HZ_JOB_COLLATZ) {
            src = r->starts_of(j), want = 8 * n;
        } else if (which == 1 && r->kind == HZ_JOB_PHOTON) {
            src = r->tally_of(j), want = 8ull * r->tally_stride;
        }
        if (!src) throw std::invalid_argument("hz_resident_output: no such output");
        if (!host || bytes != want) {
            throw std::invalid_argument("hz_resident_output: buffer size must be " + std::to_string(want));
        }
        check(hipMemcpyAsync(host, src, want, hipMemcpyDeviceToHost, r->copy_stream), "hz_resident_output copy");
        check(hipStreamSynchronize(r->copy_stream), "hz_resident_output copy");
    });
}

extern "C" int hz_resident_release(hz_resident* r, uint64_t id) {
    return guarded([&] {
        check_resident(r);
        RJob& j = r->job(id);
        if (j.status != HZ_JOB_DONE) throw std::invalid_argument("hz_resident_release: the job has not finished");
        r->release_memory(j);
    });
}

extern "C" int hz_resident_info(const hz_resident* r, hz_resident_info_t* info) {
    return guarded([&] {
        check_resident(r);
        if (!info) throw std::invalid_argument("hz_resident_info: null output");
        std::memset(info, 0, sizeof(*info));
        info->jobs = r->jobs.size();
        info->pending = (uint32_t)r->pending.size();
        info->in_flight = (uint32_t)r->in_flight.size();
        info->launches = r->launches;
        info->joined_running = r->joined;
        info->groups = r->groups;
        info->running = r->running ? 1u : 0u;
        const unsigned long long item_bytes = r->kind == HZ_JOB_PHOTON ? 4ull : 12ull;
        info->pool_bytes = item_bytes * r->pool->size();
        info->pool_used = item_bytes * r->pool->used();
        info->promises = r->totals.promises;
        info->links_run = r->totals.links_run;
        info->make_work_calls = r->totals.make_work_calls;
        info->idle_polls = r->totals.idle_polls;
        for (int i = 0; i < 8; ++i) info->phase_cycles[i] = r->totals.phase_cycles[i];
    });
}

extern "C" int hz_resident_run_trace(hz_resident* r, const hz_job* jobs, const double* arrival_ms, uint32_t count,
                                     double timeout_ms, hz_job_result* results) {
    return guarded([&] {
        check_resident(r);
        if (count && (!jobs || !arrival_ms)) throw std::invalid_argument("hz_resident_run_trace: null trace");
        for (uint32_t i = 1; i < count; ++i) {
            if (arrival_ms[i] < arrival_ms[i - 1]) {
                throw std::invalid_argument("hz_resident_run_trace: arrivals must be sorted");
            }
        }
        std::vector<uint64_t> ids(count);
        const double base = now_ms(r->t0);
        const double t_end = base + (timeout_ms > 0 ? timeout_ms : 600000.0);
        uint32_t next = 0;
        while (true) {
            const double t = now_ms(r->t0);
            while (next < count && base + arrival_ms[next] <= t) {
                // Latency counts from the scheduled arrival.
                ids[next] = r->submit(jobs[next], base + arrival_ms[next]);
                next++;
            }
            r->progress();
            if (next == count && r->pending.empty() && r->in_flight.empty()) break;
            if (t > t_end) throw std::runtime_error("hz_resident_run_trace: timed out");
        }
        if (results) {
            for (uint32_t i = 0; i < count; ++i) r->fill_result(r->jobs[(size_t)ids[i]], &results[i]);
        }
    });
}
