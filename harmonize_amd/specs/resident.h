// Resident multi-job programs: one persistent AsyncProgram launch admits jobs
// while it runs, so a served job needs no kernel launch of its own.
//
// GPU only (it reads a host-coherent doorbell and the device wall clock).
//
// The host publishes job descriptors in a doorbell ring in host-coherent
// memory (ResidentDoorbell). The running kernel installs them into a device
// job table (ResidentCore) and its groups claim work items from the installed
// jobs in priority order (admission order, or the smallest job first, so a
// small job overtakes a large one at the next claim). Each job completes on its
// own and the kernel reports it in the ring's completion words. A kernel with
// no job for `idle_ticks` of the device wall clock, or asked by the host to
// stop once idle, stops claiming and ends; the host relaunches it for later
// jobs, and the job table carries over between launches.
//
// Housekeeping (installing, completion checks, the idle decision) is done by
// whichever wavefront calls make_work first after `poll_ticks` have passed
// (one lock, taken with a compare-and-swap), so no wavefront is set aside for
// it.
//
// Job accounting, with no same-address hot spot:
//  - items are claimed from the job's counter `next` (one atomic per
//    wavefront claim); once `next` reaches `limit` every item has a claimant;
//  - whatever ends (a Collatz chain writing its step count) adds 1 to one of
//    STRIPES `ended` counters after its own outputs have reached memory;
//    entities created beyond the items (none for Collatz) would add to the
//    `created` counters before they are queued;
//  - the poller reads every `ended` counter, then every `created` counter.
//    With every item claimed, limit + created == ended means nothing of the
//    job is alive: each counted end was preceded by the creation it ends, and
//    anything alive at the first scan was created before the second.
// Outputs are stored with device-scope stores, written through past the XCD's
// L2, so they are in memory when the job is reported and no dirty line of a
// finished job can later overwrite memory the host has reused.
#pragma once

#include <cstdint>

#include "collatz.h"
#include "philox.h"
#include "photon.h"

namespace hz {
namespace resident {

constexpr unsigned int JOBS = 64;           // job table entries
constexpr unsigned int STRIPES = 16;        // accounting counters per job and kind
constexpr unsigned int STRIPE_WORDS = 32;   // 128 bytes between counters
constexpr unsigned int RING = 256;          // doorbell entries (a power of two)
constexpr unsigned int CLOSED = 0x80000000u;
constexpr unsigned int NO_JOB = 0xFFFFFFFFu;
constexpr unsigned int COUNTER_WORDS = JOBS * 2u * STRIPES * STRIPE_WORDS;
// Groups that take part in housekeeping (the first POLLERS of the launch,
// spread over the XCDs by the round-robin dispatch of workgroups).
constexpr unsigned int POLLERS = 64;
// Scheduling cycles a group skips make_work after a claim found nothing.
constexpr unsigned int CLAIM_PAUSE = 32;

// Policies for the order in which jobs are claimed from.
constexpr unsigned int BY_ADMISSION = 0u;
constexpr unsigned int SMALLEST_FIRST = 1u;

// Error bits the poller reports in ResidentCore::error.
constexpr unsigned int ERR_BAD_ENTRY = 1u;

}  // namespace resident
}  // namespace hz

// A job descriptor as the host publishes it (32 bytes). A job's items
// occupy [base, base + size) of the server's item space, so its outputs are
// slices of the program's output arrays and a promise that carries its item's
// index in that space needs nothing else to write its result.
struct ResidentEntry {
    unsigned int job;          // job table entry it installs into
    unsigned int size;         // work items
    unsigned long long seed;
    unsigned int base;         // first item in the item space
    unsigned int pad;
    unsigned long long reserved;
};

// Host-coherent memory shared by the host and the running kernel.
struct ResidentDoorbell {
    unsigned int head;         // entries published (host)
    unsigned int stop;         // 1: stop once idle (host)
    unsigned int consumed;     // entries installed (device)
    unsigned int pad;
    unsigned int done[hz::resident::RING];      // entry k: k + 1 once its job has completed (device)
    ResidentEntry entries[hz::resident::RING];  // entry k at k % RING (host)
};

// One installed job (device memory).
struct ResidentJob {
    unsigned int next;         // items claimed; CLOSED while the entry takes no claims
    unsigned int limit;        // items
    unsigned int entry;        // ring sequence number of its descriptor
    unsigned int pad;
    unsigned long long seed;
    unsigned int base;
    unsigned int pad2;
};

// Device memory shared by every group of one resident program.
struct ResidentCore {
    unsigned int poll_lock;
    unsigned int stopping;     // no more claims: make_work returns false
    unsigned int consumed;     // entries installed
    unsigned int active_count;
    unsigned long long open_jobs;   // bit j: job entry j may have items left to claim
    unsigned long long last_poll;
    unsigned long long last_busy;
    unsigned long long pass_start;
    unsigned int error;
    unsigned int installs;     // jobs installed, over all launches
    unsigned int completions;  // jobs completed, over all launches
    unsigned int pad;
    unsigned int active[hz::resident::JOBS];  // claim order
    unsigned int active_size[hz::resident::JOBS];
    ResidentJob jobs[hz::resident::JOBS];
};

struct ResidentState {
    ResidentCore* core;
    unsigned long long* starts;  // Collatz starts by item (the item space)
    unsigned int* steps;         // Collatz step counts by item
    unsigned int* counters;    // [JOBS][2 (ended, created)][STRIPES][STRIPE_WORDS]
    ResidentDoorbell* bell;    // device view of the host-coherent doorbell
    unsigned long long poll_ticks;   // wall-clock ticks between housekeeping passes
    unsigned long long idle_ticks;   // idle ticks before the launch ends
    unsigned long long life_ticks;   // ticks after which a launch stops claiming (0: none)
    unsigned int policy;       // hz::resident::BY_ADMISSION or SMALLEST_FIRST
    unsigned int bits;         // Collatz starts are uniform in [1, 2^bits)
    // Photon jobs: the item space as u32 words, where a job's allocation
    // holds its tally of `replicas` x (bins + PHOTON_EXTRA_SLOTS) doubles
    // (tally_stride doubles, 2 x tally_stride words) followed by its event
    // counts, which start at the job's base; and the walk's parameters (seed,
    // tally, events and iterator of `photon` unused; each job has its own
    // seed).
    unsigned int* events;
    unsigned int tally_stride;
    unsigned int replicas;
    PhotonState photon;
};

// Per lane: items claimed and not yet started, ends counted but not yet
// added to the job's counters, the job whose results the wavefront holds in
// LDS (the same in every lane; NO_JOB if none) and when this lane last saw a
// housekeeping pass (device wall clock).
struct ResidentClaim {
    unsigned int job;
    unsigned int next;
    unsigned int end;
    unsigned int pend_job;  // job of the pending ends
    unsigned int pend;      // ends not yet added to the `ended` counters
    unsigned int held;
    unsigned long long poll_seen;
};

namespace hz {
namespace resident {

// Memory ordering. Every field that several wavefronts share (the job table,
// the active list, the counters, the doorbell) is read and written only with
// relaxed device- or system-scope atomics, which on gfx950 are coherent
// loads and write-through stores (sc1): a load reads the value at the
// device's coherence point and a completed store is there. A writer that
// publishes (opens a job, releases the lock, reports a completion) first
// waits until its earlier stores have completed (s_waitcnt), so whoever sees
// the publishing store sees them too. Acquire and release orderings would add
// an L2 invalidate or write-back to every claim and poll; none of the shared
// data needs them. The signal fences keep the compiler from moving memory
// operations across the wait.
__device__ HZ_INLINE void stores_done() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <typename T>
__device__ HZ_INLINE T ld(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE void st(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ HZ_INLINE T sys_ld(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <typename T>
__device__ HZ_INLINE void sys_st(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Outputs are stored the same way: written through to memory.
template <typename T>
__device__ HZ_INLINE void store_out(T* p, T v) {
    st(p, v);
}

__device__ HZ_INLINE unsigned int* counter(const ResidentState& s, unsigned int job, unsigned int kind,
                                           unsigned int stripe) {
    return s.counters + (((size_t)job * 2u + kind) * STRIPES + stripe) * STRIPE_WORDS;
}

__device__ HZ_INLINE unsigned long long now() { return (unsigned long long)wall_clock64(); }

// Sum over the wavefront (every lane active).
__device__ HZ_INLINE unsigned int wave_sum(unsigned int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += (unsigned int)__shfl_xor((int)v, d, 64);
    return v;
}

// Ends are counted per lane and added to the job's `ended` counters later,
// after a wait for the wavefront's stores (so that a job is never reported
// before its outputs are in memory): at the lane's next end for another job,
// at every make_work call, and in finalize. A deferred end can only delay a
// completion, never report one early.
__device__ HZ_INLINE void flush_lane(const ResidentState& s, ResidentClaim& t) {
    stores_done();
    __hip_atomic_fetch_add(counter(s, t.pend_job, 0u, hz_group_index() % STRIPES), t.pend, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    t.pend = 0u;
}

__device__ HZ_INLINE void count_end(const ResidentState& s, ResidentClaim& t, unsigned int job) {
    if (t.pend != 0u && t.pend_job != job) flush_lane(s, t);
    t.pend_job = job;
    t.pend += 1u;
}

// Adds every lane's pending ends (the whole wavefront, every lane active):
// one wait, then one atomic per distinct job.
__device__ HZ_INLINE void flush_ends(const ResidentState& s, ResidentClaim& t) {
    uint64_t todo = hz::wave::ballot(t.pend != 0u);
    if (todo == 0ull) return;
    stores_done();
    const unsigned int stripe = hz_group_index() % STRIPES;
    while (todo != 0ull) {
        const uint32_t lead = (uint32_t)__builtin_ctzll(todo);
        const unsigned int j = hz::wave::read_lane(t.pend_job, lead);
        const bool mine = t.pend != 0u && t.pend_job == j;
        const unsigned int sum = wave_sum(mine ? t.pend : 0u);
        if (hz::wave::lane() == lead) {
            __hip_atomic_fetch_add(counter(s, j, 0u, stripe), sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        todo &= ~hz::wave::ballot(mine);
    }
    t.pend = 0u;
}

// State of an installed job (the whole wavefront calls it with the same
// job): items left to claim, every item claimed, or finished (every item and
// everything created from them has ended). The `ended` scan completes before
// the `created` scan is issued.
constexpr unsigned int JOB_OPEN = 0u, JOB_CLAIMED = 1u, JOB_FINISHED = 2u;

__device__ HZ_INLINE unsigned int job_state(const ResidentState& s, unsigned int job) {
    ResidentJob& J = s.core->jobs[job];
    const unsigned int limit = hz::wave::first(ld(&J.limit));
    const unsigned int next = hz::wave::first(ld(&J.next));
    if (next < limit) return JOB_OPEN;
    const uint32_t lane = hz::wave::lane();
    unsigned int e = lane < STRIPES ? ld(counter(s, job, 0u, lane)) : 0u;
    e = hz::wave::first(wave_sum(e));
    unsigned int c = lane < STRIPES ? ld(counter(s, job, 1u, lane)) : 0u;
    c = hz::wave::first(wave_sum(c));
    return e == limit + c ? JOB_FINISHED : JOB_CLAIMED;
}

// Installs ring entry k into its job table entry (the whole wavefront, under
// the housekeeping lock). `n` is the active list's length.
template <typename KIND>
__device__ HZ_INLINE bool install(const ResidentState& s, unsigned int k, unsigned int n) {
    ResidentCore* core = s.core;
    const ResidentEntry* e = &s.bell->entries[k % RING];
    const uint32_t lane = hz::wave::lane();
    // The 32-byte descriptor in one round trip over the bus: lane w reads word w.
    static_assert(sizeof(ResidentEntry) == 32, "descriptor layout");
    const unsigned int word = lane < 8u ? sys_ld(&((const unsigned int*)e)[lane]) : 0u;
    auto word_at = [&](unsigned int w) { return hz::wave::read_lane(word, w); };
    auto pair_at = [&](unsigned int w) {
        return ((unsigned long long)word_at(w + 1u) << 32) | (unsigned long long)word_at(w);
    };
    const unsigned int job = word_at(0);
    if (job >= JOBS) {
        if (lane == 0) __hip_atomic_fetch_or(&core->error, ERR_BAD_ENTRY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
    ResidentJob& J = core->jobs[job];
    const unsigned int size = word_at(1);
    if (lane == 0) st(&J.next, CLOSED);
    stores_done();
    if (lane < STRIPES) {
        st(counter(s, job, 0u, lane), 0u);
        st(counter(s, job, 1u, lane), 0u);
    }
    if (lane == 0) {
        st(&J.limit, size);
        st(&J.entry, k);
        st(&J.seed, pair_at(2));
        st(&J.base, word_at(4));
    }
    KIND::install(s, word_at(4));
    // Claims that see the job open see its fields and zeroed counters.
    stores_done();
    if (lane == 0) {
        st(&J.next, 0u);
        __hip_atomic_fetch_or(&core->open_jobs, 1ull << job, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // Claim order: append, or insert before the first larger job.
        unsigned int at = n;
        if (s.policy == SMALLEST_FIRST) {
            while (at > 0 && ld(&core->active_size[at - 1]) > size) --at;
        }
        for (unsigned int i = n; i > at; --i) {
            st(&core->active[i], ld(&core->active[i - 1]));
            st(&core->active_size[i], ld(&core->active_size[i - 1]));
        }
        st(&core->active[at], job);
        st(&core->active_size[at], size);
        st(&core->installs, ld(&core->installs) + 1u);
    }
    stores_done();
    if (lane == 0) st(&core->active_count, n + 1u);
    hz::wave::sync();
    return true;
}

// One housekeeping pass by the calling wavefront (every lane active) if it
// is one of the POLLERS, `poll_ticks` have passed since the last pass it saw
// and no other wavefront holds the lock: completions, new entries, the idle
// decision.
template <typename KIND>
__device__ HZ_INLINE void poll(const ResidentState& s, ResidentClaim& t) {
    if (hz_group_index() >= POLLERS) return;
    ResidentCore* core = s.core;
    const uint32_t lane = hz::wave::lane();
    const unsigned long long clock = hz::wave::first((uint64_t)now());
    if (clock - t.poll_seen < s.poll_ticks) return;
    unsigned int go = 0u;
    if (lane == 0) {
        const unsigned long long last = ld(&core->last_poll);
        if (clock - last >= s.poll_ticks &&
            __hip_atomic_compare_exchange_strong(&core->poll_lock, &go, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            go = 1u;
        } else {
            t.poll_seen = last;
            go = 0u;
        }
    }
    go = hz::wave::first(go);
    t.poll_seen = hz::wave::first((uint64_t)t.poll_seen);
    if (!go) return;
    t.poll_seen = clock;
    if (lane == 0) {
        st(&core->last_poll, clock);
        if (ld(&core->pass_start) == 0ull) st(&core->pass_start, clock);
    }

    // Completions: a finished job leaves the active list, then is reported;
    // a job with every item claimed leaves the open set.
    unsigned int n = hz::wave::first(ld(&core->active_count));
    if (n > JOBS) n = JOBS;
    const unsigned long long open = hz::wave::first((uint64_t)ld(&core->open_jobs));
    for (unsigned int i = 0; i < n;) {
        const unsigned int job = hz::wave::first(ld(&core->active[i]));
        const unsigned int state = job < JOBS ? job_state(s, job) : JOB_OPEN;
        if (state != JOB_OPEN && (open >> job & 1ull) && lane == 0) {
            __hip_atomic_fetch_and(&core->open_jobs, ~(1ull << job), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (state != JOB_FINISHED) {
            ++i;
            continue;
        }
        if (lane == 0) {
            for (unsigned int m = i; m + 1u < n; ++m) {
                st(&core->active[m], ld(&core->active[m + 1u]));
                st(&core->active_size[m], ld(&core->active_size[m + 1u]));
            }
            st(&core->active_count, n - 1u);
            st(&core->completions, ld(&core->completions) + 1u);
            const unsigned int entry = ld(&core->jobs[job].entry);
            stores_done();
            sys_st(&s.bell->done[entry % RING], entry + 1u);
        }
        hz::wave::sync();
        n -= 1u;
    }

    // New entries, while the table has room.
    const unsigned int head = hz::wave::first(sys_ld(&s.bell->head));
    unsigned int consumed = hz::wave::first(ld(&core->consumed));
    while (consumed != head && n < JOBS) {
        if (install<KIND>(s, consumed, n)) n += 1u;
        consumed += 1u;
        if (lane == 0) {
            st(&core->consumed, consumed);
            sys_st(&s.bell->consumed, consumed);
        }
    }

    if (lane == 0) {
        bool stop = false;
        if (n == 0u && consumed == head) {
            stop = sys_ld(&s.bell->stop) != 0u || clock - ld(&core->last_busy) >= s.idle_ticks;
        } else {
            st(&core->last_busy, clock);
        }
        if (s.life_ticks != 0ull && clock - ld(&core->pass_start) >= s.life_ticks) stop = true;
        if (stop) st(&core->stopping, 1u);
        stores_done();
        st(&core->poll_lock, 0u);
    }
    hz::wave::sync();
}

// Gives each lane without items a share of one job's items, from the first
// job in claim order with items left: the active list, the open set and the
// jobs' claim counters are read by the lanes in parallel (one round trip when
// no job is open), then one atomic claims for the whole wavefront, of
// KIND::per_lane(left, lanes) items per lane (`left` items unclaimed, `lanes`
// lanes in the launch). Returns false if nothing could be claimed;
// `stopping` receives the launch's stop flag, `claimed` the job claimed from
// (NO_JOB if none) and `bulk` whether lanes took more than one item each.
template <typename KIND>
__device__ HZ_INLINE bool claim(const ResidentState& s, ResidentClaim& t, bool& stopping, unsigned int& claimed,
                                bool& bulk) {
    stopping = false;
    claimed = NO_JOB;
    bulk = false;
    const uint64_t need = hz::wave::ballot(t.next >= t.end);
    if (need == 0ull) return true;
    ResidentCore* core = s.core;
    const uint32_t lane = hz::wave::lane();
    const unsigned int slot = ld(&core->active[lane]);
    const unsigned int n = hz::wave::first(ld(&core->active_count));
    const unsigned long long open = hz::wave::first((uint64_t)ld(&core->open_jobs));
    stopping = hz::wave::first(ld(&core->stopping)) != 0u;
    if (stopping || open == 0ull) return false;
    const bool listed = lane < n && slot < JOBS && (open >> slot & 1ull);
    unsigned int seen = 0u, limit = 0u;
    if (listed) {
        seen = ld(&core->jobs[slot].next);
        limit = ld(&core->jobs[slot].limit);
    }
    uint64_t avail = hz::wave::ballot(listed && seen < limit);
    const unsigned int want_lanes = hz::wave::popc(need);
    while (avail != 0ull) {
        const uint32_t lead = (uint32_t)__builtin_ctzll(avail);
        avail &= avail - 1ull;
        unsigned int base = 0u, got = 0u, per = 1u;
        if (lane == lead) {
            per = KIND::per_lane(limit - seen, gridDim.x * blockDim.x);
            ResidentJob& J = core->jobs[slot];
            const unsigned int want = want_lanes * per;
            const unsigned int old = __hip_atomic_fetch_add(&J.next, want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // The job entry may have been replaced since it was read: the
            // limit that counts is the one read after the claim.
            const unsigned int lim = ld(&J.limit);
            base = old;
            got = old < lim ? (lim - old < want ? lim - old : want) : 0u;
        }
        got = hz::wave::read_lane(got, lead);
        if (got == 0u) continue;
        base = hz::wave::read_lane(base, lead);
        per = hz::wave::read_lane(per, lead);
        const unsigned int job = hz::wave::read_lane(slot, lead);
        claimed = job;
        bulk = per > 1u;
        if (t.next >= t.end) {
            const unsigned int r = hz::wave::rank(need) * per;
            t.job = job;
            t.next = base + (r < got ? r : got);
            t.end = base + (r + per < got ? r + per : got);
        }
        return true;
    }
    return false;
}

// The spec hooks shared by the resident kinds. KIND supplies the operations
// (OpSet), STASH_SIZE, and hooks that the whole wavefront calls:
//  - per_lane(left, lanes): items per lane of a claim from a job with `left`
//    items unclaimed, in a launch of `lanes` lanes (one lane calls it);
//  - init_group(prog) when a group starts;
//  - claimed(prog, job, bulk) after a claim for lanes without items: `job`
//    is the job claimed from (NO_JOB if nothing could be claimed), `bulk`
//    whether lanes took more than one item each;
//  - install(state, base) when a job whose items start at `base` is
//    installed (under the housekeeping lock, before the job opens);
//  - launch(prog, job, item) for one claimed item (per lane);
//  - finalize(prog) when the group's launch ends.
template <typename KIND>
struct ResidentSpecT {
    typedef typename KIND::OpSet OpSet;
    typedef ResidentState DeviceState;
    typedef ResidentClaim ThreadState;

    static const size_t STASH_SIZE = KIND::STASH_SIZE;
    static const size_t FRAME_SIZE = 8;
    static const size_t POOL_SIZE = 8;

    template <typename PROGRAM>
    __device__ static void initialize(PROGRAM prog) {
        prog.thread.job = NO_JOB;
        prog.thread.next = 0u;
        prog.thread.end = 0u;
        prog.thread.pend_job = NO_JOB;
        prog.thread.pend = 0u;
        prog.thread.held = NO_JOB;
        prog.thread.poll_seen = 0ull;
        KIND::init_group(prog);
    }

    // Items claimed and not started are started now (a launch that ends at
    // its cycle budget loses none of them).
    template <typename PROGRAM>
    __device__ static void finalize(PROGRAM prog) {
        while (prog.thread.next < prog.thread.end) {
            KIND::launch(prog, prog.thread.job, prog.thread.next);
            prog.thread.next += 1u;
        }
        flush_ends(prog.device, prog.thread);
        KIND::finalize(prog);
    }

    // Housekeeping when due, a claim for lanes without items, then one item
    // per lane. More work may come until the launch stops claiming; after a
    // claim that found nothing the group runs its own links for CLAIM_PAUSE
    // cycles before asking again (idle groups sleep meanwhile), so that
    // groups without claimable work do not keep reading the job table.
    template <typename PROGRAM>
    __device__ static bool make_work(PROGRAM prog) {
        const ResidentState& s = prog.device;
        flush_ends(s, prog.thread);
        poll<KIND>(s, prog.thread);
        bool stopping = false, bulk = false;
        unsigned int claimed = NO_JOB;
        const bool found = claim<KIND>(s, prog.thread, stopping, claimed, bulk);
        if (claimed != NO_JOB || !found) KIND::claimed(prog, claimed, bulk);
        if (!found && !stopping) prog.pause_make_work(CLAIM_PAUSE);
        if (prog.thread.next < prog.thread.end) {
            KIND::launch(prog, prog.thread.job, prog.thread.next);
            prog.thread.next += 1u;
            return true;
        }
        return !stopping;
    }
};

}  // namespace resident
}  // namespace hz

// ------------------------------------------------------------------ Collatz
// The serving Collatz job: start i = the seeded Philox start in [1, 2^bits)
// (the generator of the Collatz entry points), step counts as CollatzEven /
// CollatzOdd compute them. A promise carries (value, job << 20 | step, item),
// the item's index in the server's item space.

constexpr unsigned int RESIDENT_STEP_BITS = 20;
constexpr unsigned int RESIDENT_STEP_MASK = (1u << RESIDENT_STEP_BITS) - 1u;

struct RCollatzEven;
struct RCollatzOdd;

template <typename PROGRAM>
__device__ HZ_INLINE void rcollatz_dispatch(PROGRAM& prog, unsigned long long val, unsigned int tag,
                                            unsigned int index) {
    if ((val & 1ull) == 0) {
        prog.template async<RCollatzEven>(val, tag, index);
    } else {
        prog.template async<RCollatzOdd>(val, tag, index);
    }
}

template <typename PROGRAM>
__device__ HZ_INLINE void rcollatz_finish(PROGRAM& prog, unsigned int tag, unsigned int index, unsigned int steps) {
    const unsigned int job = tag >> RESIDENT_STEP_BITS;
    hz::resident::store_out(&prog.device.steps[index], steps);
    hz::resident::count_end(prog.device, prog.thread, job);
}

struct RCollatzEven {
    using Type = void (*)(unsigned long long val, unsigned int tag, unsigned int index);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned long long val, unsigned int tag, unsigned int index) {
        if (val <= 1) {
            rcollatz_finish(prog, tag, index, tag & RESIDENT_STEP_MASK);
            return;
        }
        rcollatz_dispatch(prog, val >> 1, tag + 1u, index);
    }
};

struct RCollatzOdd {
    using Type = void (*)(unsigned long long val, unsigned int tag, unsigned int index);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, unsigned long long val, unsigned int tag, unsigned int index) {
        if (val <= 1) {
            rcollatz_finish(prog, tag, index, tag & RESIDENT_STEP_MASK);
            return;
        }
        if (val > 0x5555555555555554ull) {
            rcollatz_finish(prog, tag, index, COLLATZ_OVERFLOW);
            return;
        }
        rcollatz_dispatch(prog, val * 3ull + 1ull, tag + 1u, index);
    }
};

struct ResidentCollatzKind {
    typedef OpUnion<RCollatzEven, RCollatzOdd> OpSet;
    static const size_t STASH_SIZE = 7;

    // One share of what is left per lane of the launch, 1 to 8 starts: a
    // large job is claimed in bulk, so claims (one atomic on the job's claim
    // counter per wavefront) are few, and every lane still gets starts. A
    // fixed 8 per lane above 2^17 starts left ran 2^18-2^20-start jobs 40%
    // slower, since their starts then fill only part of the device's lanes.
    __device__ static HZ_INLINE unsigned int per_lane(unsigned int left, unsigned int lanes) {
        const unsigned int per = left / lanes;
        return per < 1u ? 1u : (per > 8u ? 8u : per);
    }

    template <typename PROGRAM>
    __device__ static HZ_INLINE void init_group(PROGRAM&) {}
    template <typename PROGRAM>
    __device__ static HZ_INLINE void claimed(PROGRAM&, unsigned int, bool) {}
    template <typename PROGRAM>
    __device__ static HZ_INLINE void finalize(PROGRAM&) {}
    __device__ static HZ_INLINE void install(const ResidentState&, unsigned int) {}

    template <typename PROGRAM>
    __device__ static HZ_INLINE void launch(PROGRAM& prog, unsigned int job, unsigned int index) {
        ResidentJob& J = prog.device.core->jobs[job];
        const unsigned long long seed = hz::resident::ld(&J.seed);
        const unsigned int item = hz::resident::ld(&J.base) + index;
        const unsigned long long modulus = (1ull << prog.device.bits) - 1ull;
        hz::philox::Ctr r = hz::philox::philox4x32_10(index, 0u, 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
        const unsigned long long start = ((((unsigned long long)r.v[1] << 32) | r.v[0]) % modulus) + 1ull;
        hz::resident::store_out(&prog.device.starts[item], start);
        rcollatz_dispatch(prog, start, job << RESIDENT_STEP_BITS, item);
    }
};

using ResidentCollatzSpec = hz::resident::ResidentSpecT<ResidentCollatzKind>;

// ------------------------------------------------------------------ photon walk
// The serving photon job: the photon walk with the workload's default
// parameters, `size` photons and the job's seed, computed by the same event
// body as PhotonStep (photon_detail::event), so every event, tally and count
// is the oracle's. A promise is a Photon whose id holds the job-table entry
// in its top bits and the photon's index within its job below them.
// Results go to:
//  - events[base + photon] (the item space), zeroed when the photon starts;
//  - the job's tally, just before its event counts in the item space
//    (replica by group), zeroed at install.
// A split into c lineages adds c - 1 to the job's `created` counters; every
// lineage's end counts one end.
//
// Tally deposits are FP64 atomics. A wavefront that claims a bulk of a job's
// photons holds that job's tally in LDS: deposits of the job's events that
// the wavefront runs go there, the others to the job's device replica. While
// it holds the tally the wavefront counts as one more lineage of the job (one
// added to `created` when it takes the tally, one end after it has added the
// tally to the device replica and waited for those writes), so the job cannot
// be reported before its LDS deposits are in memory. It gives the tally back
// when a claim finds nothing, when it claims from another job, and when its
// launch ends.
//
// Per event the job's seed and base are needed. They are cached per group
// in LDS, tagged with links_pulled(): an entry is used only while the group
// has pulled no link since it was filled, and the entry of a job the group
// claims from is dropped. A group holds a job's promises only through its own
// claims or pulled links, so a cached entry always belongs to the incarnation
// of the job whose promises the group holds.

constexpr unsigned int RESIDENT_PHOTON_ID_BITS = 26;
constexpr unsigned int RESIDENT_PHOTON_ID_MASK = (1u << RESIDENT_PHOTON_ID_BITS) - 1u;
constexpr unsigned int RESIDENT_CACHE_NONE = 0xFFFFFFFFu;

struct ResidentJobCache {
    unsigned long long seed;
    unsigned int base;
    unsigned int tag;  // links_pulled() when filled, or RESIDENT_CACHE_NONE
};

template <typename PROGRAM>
__device__ HZ_INLINE ResidentJobCache* resident_job_cache() {
    constexpr unsigned int WAVES = (unsigned int)(PROGRAM::WORK_GROUP_SIZE / 64);
    __shared__ ResidentJobCache cache[WAVES][hz::resident::JOBS];
    return cache[hz_block_thread() / 64u];
}

template <typename PROGRAM>
__device__ HZ_INLINE ResidentJobCache resident_job_params(PROGRAM& prog, unsigned int job) {
    ResidentJobCache* cache = resident_job_cache<PROGRAM>();
    const unsigned int tag = (unsigned int)prog.links_pulled();
    ResidentJobCache e = cache[job];
    if (e.tag != tag) {
        ResidentJob& J = prog.device.core->jobs[job];
        e.seed = hz::resident::ld(&J.seed);
        e.base = hz::resident::ld(&J.base);
        e.tag = tag;
        cache[job] = e;
    }
    return e;
}

template <typename PROGRAM>
__device__ HZ_INLINE double* resident_photon_tally() {
    constexpr unsigned int WAVES = (unsigned int)(PROGRAM::WORK_GROUP_SIZE / 64);
    __shared__ double tally[WAVES][photon_detail::PHOTON_LDS_SLOTS];
    return tally[hz_block_thread() / 64u];
}

// The tally of the job whose event counts start at item `base`.
__device__ HZ_INLINE double* resident_photon_job_tally(const ResidentState& s, unsigned int base) {
    return reinterpret_cast<double*>(s.events + (base - 2u * s.tally_stride));
}

__device__ HZ_INLINE double* resident_photon_replica(const ResidentState& s, unsigned int base) {
    return resident_photon_job_tally(s, base) +
           (size_t)(hz_group_index() & (s.replicas - 1u)) * (s.photon.bins + PHOTON_EXTRA_SLOTS);
}

struct RPhotonStep {
    using Type = void (*)(Photon p);

    template <typename PROGRAM>
    __device__ static void eval(PROGRAM prog, Photon p) {
        const ResidentState& s = prog.device;
        const unsigned int job = p.id >> RESIDENT_PHOTON_ID_BITS;
        const ResidentJobCache e = resident_job_params(prog, job);
        PhotonState st = s.photon;
        st.seed_lo = (unsigned int)e.seed;
        st.seed_hi = (unsigned int)(e.seed >> 32);
        Photon q = p;
        q.id = p.id & RESIDENT_PHOTON_ID_MASK;
        double* tally = job == prog.thread.held ? resident_photon_tally<PROGRAM>() : resident_photon_replica(s, e.base);
        photon_detail::event(
            st, q, [&](unsigned int slot, double value) { atomicAdd(&tally[slot], value); },
            [&](const Photon& r) {
                atomicAdd(&s.events[e.base + r.id], (unsigned int)r.pending);
                hz::resident::count_end(s, prog.thread, job);
            },
            [&](unsigned int copies) {
                // No wait is needed before the copies are queued: an end of
                // any of them is added to the `ended` counters only after a
                // wait for this wavefront's earlier memory operations (the
                // deferred flush), or after the release fence of the spill
                // that hands them to another group.
                __hip_atomic_fetch_add(hz::resident::counter(s, job, 1u, hz_group_index() % hz::resident::STRIPES),
                                       copies - 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            },
            [&](const Photon& c) {
                Photon d = c;
                d.id = c.id | (job << RESIDENT_PHOTON_ID_BITS);
                prog.template async<RPhotonStep>(d);
            });
    }
};

struct ResidentPhotonKind {
    typedef OpUnion<RPhotonStep> OpSet;
    static const size_t STASH_SIZE = HZ_PHOTON_STASH;

    // Eight photons per lane while HZ_PHOTON_SPREAD_BELOW or more are
    // left, else one. Bulk claims also take the job's tally into LDS. (One
    // share of what is left per lane, as for Collatz, lowers the photon
    // serving peak by a quarter; four per lane, the photon program's own
    // claim, by 6%.)
    __device__ static HZ_INLINE unsigned int per_lane(unsigned int left, unsigned int) {
        return left >= HZ_PHOTON_SPREAD_BELOW ? 8u : 1u;
    }

    template <typename PROGRAM>
    __device__ static HZ_INLINE void init_group(PROGRAM&) {
        ResidentJobCache* cache = resident_job_cache<PROGRAM>();
        cache[hz::wave::lane()].tag = RESIDENT_CACHE_NONE;  // JOBS == 64 lanes
        hz::wave::sync();
    }

    template <typename PROGRAM>
    __device__ static HZ_INLINE void claimed(PROGRAM& prog, unsigned int job, bool bulk) {
        if (job != hz::resident::NO_JOB) {
            if (hz::wave::lane() == 0) resident_job_cache<PROGRAM>()[job].tag = RESIDENT_CACHE_NONE;
            hz::wave::sync();
        }
        const unsigned int held = prog.thread.held;
        if (held == job) return;
        if (held != hz::resident::NO_JOB) give_back(prog);
        const bool fits = prog.device.photon.bins + PHOTON_EXTRA_SLOTS <= photon_detail::PHOTON_LDS_SLOTS;
        if (job != hz::resident::NO_JOB && bulk && fits) take(prog, job);
    }

    template <typename PROGRAM>
    __device__ static HZ_INLINE void finalize(PROGRAM& prog) {
        if (prog.thread.held != hz::resident::NO_JOB) give_back(prog);
    }

    // Takes `job`'s tally into LDS (zeroed) and counts the holder as a lineage.
    template <typename PROGRAM>
    __device__ static HZ_INLINE void take(PROGRAM& prog, unsigned int job) {
        const ResidentState& s = prog.device;
        double* t = resident_photon_tally<PROGRAM>();
        const unsigned int slots = s.photon.bins + PHOTON_EXTRA_SLOTS;
        for (unsigned int i = hz::wave::lane(); i < slots; i += 64u) t[i] = 0.0;
        if (hz::wave::lane() == 0) {
            __hip_atomic_fetch_add(hz::resident::counter(s, job, 1u, hz_group_index() % hz::resident::STRIPES), 1u,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        hz::wave::sync();
        prog.thread.held = job;
    }

    // Adds the held tally to the job's device replica, waits for those
    // writes, then ends the holder's lineage.
    template <typename PROGRAM>
    __device__ static HZ_INLINE void give_back(PROGRAM& prog) {
        const ResidentState& s = prog.device;
        const unsigned int job = prog.thread.held;
        hz::wave::sync();
        const double* t = resident_photon_tally<PROGRAM>();
        double* rep = resident_photon_replica(s, hz::wave::first(hz::resident::ld(&s.core->jobs[job].base)));
        const unsigned int slots = s.photon.bins + PHOTON_EXTRA_SLOTS;
        for (unsigned int i = hz::wave::lane(); i < slots; i += 64u) {
            const double v = t[i];
            if (v != 0.0) atomicAdd(&rep[i], v);
        }
        hz::resident::stores_done();
        if (hz::wave::lane() == 0) {
            __hip_atomic_fetch_add(hz::resident::counter(s, job, 0u, hz_group_index() % hz::resident::STRIPES), 1u,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        prog.thread.held = hz::resident::NO_JOB;
    }

    // Zeroes the job's tally (the whole wavefront).
    __device__ static HZ_INLINE void install(const ResidentState& s, unsigned int base) {
        double* t = resident_photon_job_tally(s, base);
        for (unsigned int i = hz::wave::lane(); i < s.tally_stride; i += 64u) hz::resident::st(&t[i], 0.0);
    }

    template <typename PROGRAM>
    __device__ static HZ_INLINE void launch(PROGRAM& prog, unsigned int job, unsigned int index) {
        ResidentJob& J = prog.device.core->jobs[job];
        const unsigned int item = hz::resident::ld(&J.base) + index;
        hz::resident::store_out(&prog.device.events[item], 0u);
        Photon p;
        p.x = 0.0f;
        p.y = 0.0f;
        p.z = 0.0f;
        p.ux = 0.0f;
        p.uy = 0.0f;
        p.uz = 1.0f;
        p.w = 1.0f;
        p.id = (job << RESIDENT_PHOTON_ID_BITS) | index;
        p.event = 0u;
        p.lineage = 0u;
        p.pending = 0u;
        prog.template async<RPhotonStep>(p);
    }
};

using ResidentPhotonSpec = hz::resident::ResidentSpecT<ResidentPhotonKind>;
