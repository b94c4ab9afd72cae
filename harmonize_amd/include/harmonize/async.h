// AsyncProgram<Spec>: the persistent, promise-coalescing program type.
//
// Semantics follow the reference's program/async.h: a persistent kernel in
// which every group repeatedly picks a link of same-operation promises and
// evaluates it one promise per lane (exec_cycle :2155-2319), calls the
// spec's make_work when it has no full link (:2201-2222), spills surplus
// links to device memory and refills from it (:1246-1308, :1800-2010), and
// the last group to leave decides completion (cleanup_runtime :2323-2393).
// A launch that ends early (cycle budget, halt_early) leaves every promise in
// device memory so that the next exec resumes it.
//
// The design is gfx950-native rather than a translation:
//  - a group is one 64-lane wavefront and a link holds 64 promises;
//  - coalescing uses the EXEC mask: one ballot, mbcnt lane ranks and
//    readfirstlane broadcasts per call site; the stash bookkeeping is
//    private to the wavefront, so it needs no atomics and no barriers;
//  - stash links live in LDS in structure-of-arrays form, [link][word][lane],
//    so appends, reads, spills and refills are bank-conflict free and spills
//    and refills move whole 256-byte rows with all 64 lanes (the reference
//    copies links with a single lane and re-enqueues refilled promises one at
//    a time, async.h:1207, :1767-1771);
//  - the HBM arena holds links in the same [link][word][lane] layout;
//    full links travel through bounded lock-free ring queues of link ids
//    (one queue per XCD by default, selected from blockIdx % 8, since
//    workgroups are dealt to the 8 XCDs round-robin), with
//    stealing from other queues when a group's own queue is empty;
//  - termination is counter-based: `units` counts groups that hold promises
//    plus links parked in queues, `makers` counts groups whose make_work may
//    still produce work; the run is quiescent when both are zero.
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.h"
#include "iter.h"
#include "spec.h"

namespace hz {

// One bounded multi-producer multi-consumer ring of link ids. The three
// counters sit on separate 128-byte lines.
struct LinkQueue {
    unsigned int tail;
    unsigned int pad0[31];
    unsigned int head;
    unsigned int pad1[31];
    int count;
    unsigned int pad2[31];
};

// Run status and statistics of an AsyncProgram instance, kept in HBM.
struct AsyncStatus {
    unsigned int flags;
    unsigned int stall_site;          // first bounded wait that gave up (AsyncStall)
    unsigned long long stall_detail;  // what it waited on
    unsigned int pad0[28];
    int units;
    unsigned int pad1[31];
    int makers;
    unsigned int pad2[31];
    int hungry;  // groups with no work of their own, waiting on the queues
    unsigned int pad4[31];
    unsigned int checkout;
    unsigned int claim;
    unsigned int pad3[30];
    unsigned long long int promises;
    unsigned long long int spills;
    unsigned long long int fills;
    unsigned long long int make_work_calls;
    unsigned long long int launches;
    unsigned long long int idle_polls;
    unsigned long long int links_run;  // links evaluated (promises / links_run = mean lanes busy)
    // Promises written into links: by async calls into a stash link and by
    // barrier awaits into a barrier link. Every promise written is evaluated
    // exactly once, so appended = promises + those still queued or held.
    unsigned long long int appended;
    // Promises run straight from registers by link forwarding (they are
    // counted in `appended` and `promises` but never pass through a stash).
    unsigned long long int forwarded;
    // Shader-clock cycles per scheduler phase, summed over groups; filled only
    // when the runtime is built with HZ_PHASE_TIMING (the counterpart of the
    // reference's HRM_TIME phase timers, preamble/mod.h:59-67).
    unsigned long long int phase_cycles[8];
};

// Scheduler phases timed under HZ_PHASE_TIMING.
enum AsyncPhase : unsigned int {
    PHASE_CYCLE = 0,   // a whole scheduling cycle
    PHASE_SPILL = 1,   // keeping free stash links
    PHASE_PULL = 2,    // pulling queued links
    PHASE_MAKE = 3,    // make_work
    PHASE_PICK = 4,    // choosing the next link
    PHASE_READ = 5,    // reading a link into registers
    PHASE_EVAL = 6,    // evaluating promises (spec code + async calls)
    PHASE_IDLE = 7,    // idle polling
};

constexpr unsigned int LINK_NONE = 0xFFFFFFFFu;

// What a host-side audit of an AsyncProgram instance found between launches
// (AsyncProgram::Instance::audit).
struct AsyncCensus {
    unsigned long long claimed;          // arena links ever handed out
    unsigned long long queued_links;     // links parked in the work queues
    unsigned long long queued_promises;  // promises in those links
    unsigned long long free_links;       // links in the free ring
    unsigned long long held_links;       // links held by barriers (neither queued nor free)
    unsigned long long held_promises;    // promises written into held links
    unsigned long long evaluated;        // promises evaluated so far
    unsigned long long appended;         // promises written into links so far
};

// A host copy of an AsyncProgram instance's status, rings and link metadata,
// taken between launches.
struct AsyncSnapshot {
    AsyncStatus status{};
    std::vector<LinkQueue> queues;    // the work queues, then the free ring
    std::vector<unsigned int> slots;  // queues.size() rings of `capacity` slots
    unsigned int capacity = 0;        // power of two
    std::vector<unsigned int> meta;   // per claimed link: op | level << 8 | count << 16
    std::vector<unsigned int> aux;    // per claimed link: next, (target << 16) | filled
    unsigned int arena_links = 0;
    unsigned int ops = 0;
    unsigned int levels = 1;
};

// Checks a snapshot: every group has left (checkout, makers and hungry are
// zero), each ring's count equals tail - head and exactly its slots in
// [head, tail) hold link ids, no link id is queued twice or both queued and
// free, units equals the queued links, every queued link names a valid
// operation and holds 1..64 promises, and every promise written into a link
// is accounted for: appended = evaluated + queued + held. Claimed links that
// are neither queued nor free are held by barriers (allowed only if
// `allow_held`). Throws std::runtime_error naming the first violation.
inline AsyncCensus audit_snapshot(const AsyncSnapshot& s, bool allow_held) {
    auto fail = [](const std::string& what) { throw std::runtime_error("AsyncProgram audit: " + what); };
    const AsyncStatus& st = s.status;
    if (st.checkout != 0 || st.makers != 0 || st.hungry != 0) {
        fail("groups still checked in (checkout " + std::to_string(st.checkout) + ", makers " +
             std::to_string(st.makers) + ", hungry " + std::to_string(st.hungry) + ")");
    }
    const unsigned int cap = s.capacity;
    if (cap == 0 || (cap & (cap - 1u)) != 0 || s.slots.size() != (size_t)s.queues.size() * cap) {
        fail("malformed snapshot");
    }
    // The bump counter may overshoot the arena (claims that lost the race).
    const unsigned int claimed = st.claim < s.arena_links ? st.claim : s.arena_links;
    if (s.meta.size() < claimed || s.aux.size() < 2 * (size_t)claimed) fail("malformed snapshot");
    std::vector<unsigned char> seen(claimed, 0);  // 1 = queued for work, 2 = free
    AsyncCensus c{};
    c.claimed = claimed;
    c.evaluated = st.promises;
    c.appended = st.appended;
    const size_t nq = s.queues.size();
    for (size_t q = 0; q < nq; ++q) {
        const bool is_free = q + 1 == nq;
        const std::string name = is_free ? std::string("free ring") : "queue " + std::to_string(q);
        const LinkQueue& Q = s.queues[q];
        const unsigned int span = Q.tail - Q.head;
        if (Q.count < 0 || (unsigned int)Q.count != span || span > cap) {
            fail(name + ": count " + std::to_string(Q.count) + " but tail - head = " + std::to_string(span));
        }
        const unsigned int* ring = s.slots.data() + q * cap;
        for (unsigned int k = 0; k < cap; ++k) {
            const unsigned int pos = Q.head + k;
            const unsigned int id = ring[pos & (cap - 1u)];
            if (k >= span) {
                if (id != LINK_NONE) fail(name + ": occupied slot outside [head, tail)");
                continue;
            }
            if (id == LINK_NONE || id >= claimed) fail(name + ": slot " + std::to_string(pos) + " holds no valid link");
            if (seen[id]) fail("link " + std::to_string(id) + " appears twice in the rings");
            seen[id] = is_free ? 2 : 1;
            if (is_free) {
                c.free_links++;
                continue;
            }
            const unsigned int op = s.meta[id] & 0xFFu, level = (s.meta[id] >> 8) & 0xFFu, count = s.meta[id] >> 16;
            if (op >= s.ops) fail("queued link " + std::to_string(id) + " names operation " + std::to_string(op));
            if (level >= s.levels) fail("queued link " + std::to_string(id) + " names level " + std::to_string(level));
            if (count < 1 || count > 64) {
                fail("queued link " + std::to_string(id) + " holds " + std::to_string(count) + " promises");
            }
            c.queued_links++;
            c.queued_promises += count;
        }
    }
    for (unsigned int id = 0; id < claimed; ++id) {
        if (seen[id]) continue;
        c.held_links++;
        const unsigned int filled = s.aux[2u * id + 1u] & 0xFFFFu;
        if (filled > 64) fail("held link " + std::to_string(id) + " counts " + std::to_string(filled) + " promises");
        c.held_promises += filled;
    }
    if (!allow_held && c.held_links != 0) fail(std::to_string(c.held_links) + " arena links are neither queued nor free");
    if (st.units < 0 || (unsigned long long)st.units != c.queued_links) {
        fail("units " + std::to_string(st.units) + " but " + std::to_string(c.queued_links) + " queued links");
    }
    if (c.appended != c.evaluated + c.queued_promises + c.held_promises) {
        fail("promises unaccounted for: appended " + std::to_string(c.appended) + ", evaluated " +
             std::to_string(c.evaluated) + ", queued " + std::to_string(c.queued_promises) + ", held " +
             std::to_string(c.held_promises));
    }
    return c;
}

// Bounded waits of the runtime; the first one to give up is recorded.
enum AsyncStall : unsigned int {
    STALL_NONE = 0,
    STALL_QUEUE_PUSH = 1,     // a ring slot stayed occupied (detail: queue << 32 | slot)
    STALL_QUEUE_POP = 2,      // a reserved ring slot stayed empty (detail: queue << 32 | slot)
    STALL_BARRIER_WAIT = 3,   // a barrier block's link did not appear (detail: its ring entry)
    STALL_BARRIER_INSTALL = 4,  // a released barrier's last block link did not appear (detail: ring entry)
    STALL_BARRIER_LIST = 5,   // a full barrier link could not be listed (detail: list head)
};

inline const char* stall_name(unsigned int site) {
    switch (site) {
        case STALL_QUEUE_PUSH: return "link queue push";
        case STALL_QUEUE_POP: return "link queue pop";
        case STALL_BARRIER_WAIT: return "barrier await";
        case STALL_BARRIER_INSTALL: return "barrier link install";
        case STALL_BARRIER_LIST: return "barrier full list";
        default: return "unknown";
    }
}

}  // namespace hz

template <typename PROGRAM_SPEC>
class AsyncProgram {
   public:
    typedef AsyncProgram<PROGRAM_SPEC> ProgramType;

    using AdrType = typename hz::spec::AdrTypeOr<PROGRAM_SPEC, unsigned int>::type;
    using OpSet = hz::spec::ProgramOpSet<PROGRAM_SPEC>;
    using DeviceState = typename hz::spec::DeviceStateOr<PROGRAM_SPEC, VoidState>::type;
    using GroupState = typename hz::spec::GroupStateOr<PROGRAM_SPEC, VoidState>::type;
    using ThreadState = typename hz::spec::ThreadStateOr<PROGRAM_SPEC, VoidState>::type;

    template <typename OP_SET, typename ADR_TYPE>
    friend struct RemappingBarrier;

    typedef PromiseUnion<OpSet> PromiseUnionType;

    template <typename TYPE>
    struct Lookup {
        typedef typename PromiseUnionType::template Lookup<TYPE>::type type;
    };

    static_assert(!hz::spec::GROUP_SIZEOr<PROGRAM_SPEC, 64>::given ||
                      hz::spec::GROUP_SIZEOr<PROGRAM_SPEC, 64>::value == 64,
                  "On gfx950 a group is one 64-lane wavefront; GROUP_SIZE must be 64 if given.");

    static const size_t STASH_SIZE = hz::spec::STASH_SIZEOr<PROGRAM_SPEC, 16>::value;
    static const size_t FRAME_SIZE = hz::spec::FRAME_SIZEOr<PROGRAM_SPEC, 32>::value;
    static const size_t POOL_SIZE = hz::spec::POOL_SIZEOr<PROGRAM_SPEC, 32>::value;
    static const size_t STACK_SIZE = hz::spec::STACK_SIZEOr<PROGRAM_SPEC, 0>::value;
    static const size_t GROUP_SIZE = 64;
    // Groups per workgroup: each group keeps its own stash and scheduler
    // state; the groups of a workgroup share its __shared__ memory, so a
    // specification's block-level LDS data is paid for once per workgroup.
    static constexpr uint32_t GROUPS_PER_BLOCK =
        (uint32_t)hz::spec::GROUPS_PER_BLOCKOr<PROGRAM_SPEC, 1>::value;
    static_assert(GROUPS_PER_BLOCK >= 1 && GROUPS_PER_BLOCK <= 16, "GROUPS_PER_BLOCK must be in [1, 16]");
    static const size_t WORK_GROUP_SIZE = 64 * GROUPS_PER_BLOCK;
    static constexpr bool SHARE_WORK = hz::spec::SHARE_WORKOr<PROGRAM_SPEC, 0>::value != 0;
    static constexpr bool OLDEST_FIRST = hz::spec::OLDEST_FIRSTOr<PROGRAM_SPEC, 0>::value != 0;
    // Groups live for the whole launch (specifications may keep group-level
    // partial results, e.g. LDS tallies, and flush them in finalize).
    static constexpr bool PERSISTENT = true;

    // Depth levels of the work stack (reference WorkStack, mem/work.h:133-164:
    // STACK_SIZE frames, 0 = one flat frame). Every promise carries a level in
    // [0, LEVELS): make_work's promises start at level 0, and an async call
    // places its promise at the caller's level plus a depth delta
    // (async_call_cast(depth_delta, promise)), clamped to the stack. Each
    // level has its own open links in the stash and its own link queues in
    // device memory, and groups run and pull the deepest level first, as the
    // reference's pull_promises_any_level does (program/async.h:1164-1181),
    // which keeps the population of a recursive program bounded; a
    // specification with SHALLOW_FIRST runs the shallowest level first.
    static constexpr uint32_t LEVELS = STACK_SIZE == 0 ? 1u : (uint32_t)STACK_SIZE;
    static_assert(LEVELS <= 16, "at most 16 depth levels (STACK_SIZE)");
    static constexpr bool SHALLOW_FIRST = hz::spec::SHALLOW_FIRSTOr<PROGRAM_SPEC, 0>::value != 0;

    static const unsigned char FN_ID_COUNT = PromiseUnionType::Info::COUNT;
    static constexpr uint32_t OPS = PromiseUnionType::Info::COUNT;
    static_assert(OPS <= 255, "at most 255 operations");
    static constexpr uint32_t WORDS = hz::PromiseWords<OpSet>::value;

    static_assert(std::is_trivially_copyable<PromiseUnionType>::value,
                  "promise arguments must be trivially copyable");

    // LDS stash: STASH_SIZE links of 64 promises, bounded to 32 links and to
    // 48 KiB per wavefront, and never fewer than one open link per operation
    // plus headroom.
    static constexpr uint32_t LINK_BYTES = WORDS * 64u * 4u;
    static constexpr uint32_t LDS_LINK_CAP = (48u * 1024u) / LINK_BYTES;
    static constexpr uint32_t MIN_LINKS = OPS * LEVELS + 3u;
    static constexpr uint32_t WANT_LINKS =
        (STASH_SIZE < 32 ? (uint32_t)STASH_SIZE : 32u) < LDS_LINK_CAP ? (STASH_SIZE < 32 ? (uint32_t)STASH_SIZE : 32u)
                                                                      : LDS_LINK_CAP;
    static constexpr uint32_t LINKS = WANT_LINKS > MIN_LINKS ? WANT_LINKS : MIN_LINKS;
    static_assert(LINKS <= 32, "operation set too large for one wavefront's stash");
    static constexpr uint32_t ALL_LINKS = (LINKS == 32) ? 0xFFFFFFFFu : ((1u << LINKS) - 1u);
    // Free links kept before a link is evaluated: one per operation for the
    // links its children may open, plus one.
    static constexpr uint32_t MARGIN = OPS + 1u;
    // Ready links a group may run per scheduling cycle. Running up to four
    // took Collatz 2^26 from 70.7 to 67.0 ms and the photon walk from 5.78 to
    // 5.66 ms, eight from 60.7 to 59.4 ms (with the group counters in
    // registers), sixteen from 52.6-53.0 to 52.3-52.5 ms (with link
    // forwarding; the photon walk, futures and barrier unchanged); a program
    // that shares work checks its hungry count once per cycle, so it runs one.
#ifndef HZ_BURST
#define HZ_BURST 16
#endif
    static constexpr uint32_t BURST = SHARE_WORK ? 1u : (uint32_t)HZ_BURST;

    // Asynchronous argument loads (reference ASYNC_LOADS, async/promise.h:
    // 284-303, program/async.h:2082-2144): an operation whose parameters
    // include a LazyLoad<T> has the T of each promise loaded before it runs.
    // When a group starts a link, it issues direct-to-LDS loads
    // (global_load_lds) of the lazy values of the ready link it will most
    // likely run next, lane i of that link into column i of a per-group
    // staging block, and then evaluates the current link while they are in
    // flight; the next link reads its values from the block. A link that was
    // not prefetched reads its values when it runs.
    template <unsigned I>
    struct LazyFacts {
        using Op = typename hz::OpOf<I, OpSet>::type;
        using First = hz::FirstLazy<typename Promise<Op>::Args>;
        static constexpr bool HAS = First::HAS;
        static constexpr uint32_t VALUE_WORDS = HAS ? (uint32_t)(sizeof(typename First::Value) / 4u) : 0u;
        static_assert(!HAS || sizeof(typename First::Value) % 4u == 0, "a LazyLoad value must be whole 32-bit words");
    };
    template <unsigned I = 0>
    static constexpr uint32_t lazy_words() {
        if constexpr (I >= OPS) {
            return 0u;
        } else {
            constexpr uint32_t here = LazyFacts<I>::VALUE_WORDS;
            constexpr uint32_t rest = lazy_words<I + 1>();
            return here > rest ? here : rest;
        }
    }
    static constexpr uint32_t LAZY_WORDS = lazy_words();
    static constexpr bool HAS_LAZY = LAZY_WORDS > 0;
    static_assert(LAZY_WORDS <= 16, "a LazyLoad value may span at most 16 words");
    static constexpr uint32_t SPARES = 4;

    // Link forwarding. While a link is evaluated, each lane's first async
    // call (at the caller's depth level) is held in the lane's registers
    // instead of being appended to the stash. If afterwards every held
    // promise names the same operation and at least FORWARD_MIN lanes hold
    // one, those promises run at once as a link of their own, straight from
    // registers, and so on for up to FORWARD_RUNS links; otherwise, and
    // after the last forwarded run, the held promises are appended to the
    // stash as usual. A full link of one operation whose promises each queue
    // one promise of another (a Collatz odd step is always followed by an
    // even one) thus skips the stash's append, bookkeeping and read for its
    // successor link (Collatz 2^26: 58.9 -> 52.1 ms per launch). Order of
    // evaluation is all that changes. Only flat programs (one depth level)
    // without asynchronous argument loads forward, and a specification can
    // turn it off (FORWARD_LINKS = 0).
#ifndef HZ_FORWARD
#define HZ_FORWARD 1
#endif
#ifndef HZ_FORWARD_MIN
#define HZ_FORWARD_MIN 48
#endif
#ifndef HZ_FORWARD_RUNS
#define HZ_FORWARD_RUNS 4
#endif
    // Layout of a link's promise words, the same in the stash and in the
    // arena (so moving a link is a flat copy): promise-major when a promise
    // is a whole number of 16-byte units, so that a lane writes and reads its
    // promise with 16-byte accesses, and word-major otherwise.
#ifndef HZ_PROMISE_MAJOR
#define HZ_PROMISE_MAJOR 1
#endif
    static constexpr bool PROMISE_MAJOR = HZ_PROMISE_MAJOR != 0 && WORDS % 4u == 0u;
    __device__ static HZ_INLINE uint32_t word_index(uint32_t promise, uint32_t word) {
        return PROMISE_MAJOR ? promise * WORDS + word : word * 64u + promise;
    }

    static constexpr bool FORWARD = HZ_FORWARD != 0 && hz::spec::FORWARD_LINKSOr<PROGRAM_SPEC, 1>::value != 0 &&
                                    LEVELS == 1 && !HAS_LAZY;
    static constexpr uint32_t FORWARD_MIN = HZ_FORWARD_MIN;
    static constexpr uint32_t FORWARD_RUNS = HZ_FORWARD_RUNS;

    static const unsigned int BAD_FUNC_ID_FLAG = 0x00000001;
    static const unsigned int STASH_FAIL_FLAG = 0x00000002;
    static const unsigned int ARENA_EXHAUSTED_FLAG = 0x00000008;
    static const unsigned int QUEUE_STALL_FLAG = 0x00000010;
    static const unsigned int USAGE_FLAG = 0x00000020;
    static const unsigned int EARLY_HALT_FLAG = 0x40000000;
    static const unsigned int COMPLETION_FLAG = 0x80000000;

    // Bound on spins while waiting for a queue slot that another running
    // wavefront has claimed but not yet written or cleared.
    static constexpr uint32_t SLOT_SPIN_LIMIT = 1u << 22;
    // A wavefront with an empty stash leaves the launch after this many
    // consecutive idle polls. It holds no promises, so leaving is always
    // safe; the wavefronts that still hold work finish it, and completion is
    // decided by the last one out.
    static constexpr uint32_t IDLE_EXIT_POLLS = 1u << 16;
#ifndef HZ_IDLE_BACKOFF_MAX
#define HZ_IDLE_BACKOFF_MAX 1
#endif
    // Longest idle pause, in units of 256 shader cycles. Growing pauses
    // (e.g. 32) cut idle polling traffic but cost Collatz 3% (groups pick up
    // freshly queued links later), so the default pause stays fixed.
    static constexpr uint32_t IDLE_BACKOFF_MAX = HZ_IDLE_BACKOFF_MAX;

    struct ThreadContext {
        unsigned int thread_id;
        unsigned int rand_state;
        // Promises this lane wrote into links. Counted per lane, in a
        // register: appends run in divergent code, where a shared counter
        // would need a cross-lane ordering on every call.
        unsigned long long appended;
        // Promises evaluated and links run by this lane's group: the same
        // value in every lane, kept in registers rather than read and
        // written back in LDS on every link.
        unsigned long long executed;
        unsigned long long runs;
        // Scheduler counters of the group, likewise uniform and in registers.
        // They are written only where the whole wavefront is active.
        unsigned long long made;   // make_work calls
        unsigned long long idle;   // idle polls
        uint32_t idle_streak;      // consecutive idle cycles
        uint32_t queue_check;      // cycles until the device queues are consulted again
        uint32_t make_pause;       // cycles before make_work is called again (pause_make_work)
        // Link forwarding: whether async calls are held (uniform; set while
        // a link is evaluated), and this lane's held promise (its operation,
        // LINK_NONE if none, and its words).
        uint32_t fwd_on;
        uint32_t fwd_op;
        uint32_t fwd_words[FORWARD ? WORDS : 1];
        unsigned long long forwarded;  // promises the group ran by forwarding (uniform)
    };

    // Per-wavefront scheduler state, in LDS.
    struct GroupContext {
        alignas(16) uint32_t words[LINKS][WORDS * 64];  // word_index(promise, word)
        uint32_t link_op[LINKS];
        uint32_t link_count[LINKS];
        uint32_t link_seq[OLDEST_FIRST ? LINKS : 1];  // fill order of ready links (OLDEST_FIRST)
        uint32_t link_level[LEVELS > 1 ? LINKS : 1];  // depth level of each link's promises
        uint32_t seq;
        uint32_t level;                                // level of the promises being evaluated
        // Open link of (operation, level) in bits 0-7 and its promise count
        // in bits 8-15 (one LDS read gives an append both), or LINK_NONE.
        uint32_t open[OPS > 0 ? OPS * LEVELS : 1];
        uint32_t free_mask;
        uint32_t ready_mask;
        uint32_t spare[SPARES];
        uint32_t spare_count;
        uint32_t holding;
        uint32_t can_make_work;
        uint32_t maker;
        uint32_t home;
        unsigned long long int spills;
        unsigned long long int fills;
        uint32_t share_check;  // cycles until the hungry count is consulted again
        uint32_t hungry;       // this group is counted in status->hungry
        uint32_t no_makers;    // this group has seen (and acquired) makers == 0
        uint32_t pf_link;      // stash link whose lazy values are being loaded (HAS_LAZY)
#ifdef HZ_PHASE_TIMING
        unsigned long long int phase[8];
#endif
    };

#ifdef HZ_PHASE_TIMING
    struct PhaseTimer {
        GroupContext& g;
        unsigned int phase;
        unsigned long long int start;
        __device__ PhaseTimer(GroupContext& gc, unsigned int p) : g(gc), phase(p), start(clock64()) {}
        __device__ ~PhaseTimer() {
            unsigned long long int now = clock64();
            if (hz::wave::leader()) g.phase[phase] += now - start;
        }
    };
#define HZ_PHASE(p) PhaseTimer _hz_phase_timer_##p(_grp_ctx, hz::p)
#else
#define HZ_PHASE(p) \
    do {            \
    } while (0)
#endif

    struct DeviceContext {
        typedef ProgramType ParentProgramType;
        hz::AsyncStatus* status;
        hz::LinkQueue* queues;      // [LEVELS][queue_count]
        unsigned int* queue_slots;  // [LEVELS][queue_count][queue_capacity]
        hz::LinkQueue* free_queue;
        unsigned int* free_slots;   // [queue_capacity]
        unsigned int* arena;        // [links][WORDS][64]
        unsigned int* link_meta;    // [links]: op | level << 8 | count << 16
        unsigned int* link_aux;     // [links]: barrier bookkeeping (next link, fill)
        unsigned int links;
        unsigned int queue_count;   // link queues per depth level
        unsigned int queue_mask;    // queue_capacity - 1
        unsigned int coalesce;      // minimum fill before an open link runs while work can still be made
    };

    // Owns the arena, queues and status of one program instance (reference
    // program/async.h:302-384). `arena_links` is the number of 64-promise
    // links of device memory available for spilled work.
    struct Instance {
        unsigned int arena_size;
        unsigned int queue_count;
        unsigned int queue_capacity;
        unsigned int coalesce;
        util::host::DevBuf<hz::AsyncStatus> status;
        util::host::DevBuf<hz::LinkQueue> queues;
        util::host::DevBuf<unsigned int> queue_slots;
        util::host::DevBuf<hz::LinkQueue> free_queue;
        util::host::DevBuf<unsigned int> free_slots;
        util::host::DevBuf<unsigned int> arena;
        util::host::DevBuf<unsigned int> link_meta;
        util::host::DevBuf<unsigned int> link_aux;
        DeviceState device_state;
        hz::AsyncStatus host_status{};

        static unsigned int pow2_at_least(size_t n) {
            unsigned int p = 64;
            while (p < n) p <<= 1;
            return p;
        }

        static unsigned int default_queue_count() {
            size_t f = FRAME_SIZE;
            if (f >= 8) return 8;
            return f == 0 ? 1u : (unsigned int)f;
        }

        Instance(size_t arena_links, DeviceState gs, unsigned int queues_wanted = 0, unsigned int coalesce_min = 64)
            : arena_size((unsigned int)(arena_links < 1 ? 1 : arena_links)),
              queue_count(queues_wanted ? queues_wanted : default_queue_count()),
              queue_capacity(pow2_at_least(arena_links < 1 ? 1 : arena_links)),
              coalesce(coalesce_min > 64 ? 64 : coalesce_min),
              status((size_t)1),
              queues((size_t)queue_count * LEVELS),
              queue_slots((size_t)queue_count * LEVELS * pow2_at_least(arena_links < 1 ? 1 : arena_links)),
              free_queue((size_t)1),
              free_slots((size_t)pow2_at_least(arena_links < 1 ? 1 : arena_links)),
              arena((size_t)(arena_links < 1 ? 1 : arena_links) * WORDS * 64u),
              link_meta((size_t)(arena_links < 1 ? 1 : arena_links)),
              link_aux((size_t)2 * (arena_links < 1 ? 1 : arena_links)),
              device_state(gs) {
            if (arena_links > 0x7FFFFFFFull) throw std::runtime_error("AsyncProgram::Instance: arena too large");
            reset(nullptr);
            util::host::auto_throw(hipDeviceSynchronize());
        }

        DeviceContext to_context() {
            DeviceContext c;
            c.status = status.data();
            c.queues = queues.data();
            c.queue_slots = queue_slots.data();
            c.free_queue = free_queue.data();
            c.free_slots = free_slots.data();
            c.arena = arena.data();
            c.link_meta = link_meta.data();
            c.link_aux = link_aux.data();
            c.links = arena_size;
            c.queue_count = queue_count;
            c.queue_mask = queue_capacity - 1u;
            c.coalesce = coalesce;
            return c;
        }

        // Empties every queue and clears the status (graph-capturable: a kernel, no memset).
        void reset(hipStream_t stream);

        hz::AsyncStatus read_status() {
            util::host::auto_throw(
                hipMemcpy(&host_status, status.data(), sizeof(hz::AsyncStatus), hipMemcpyDeviceToHost));
            return host_status;
        }

        // True once every promise has been evaluated and make_work is exhausted.
        bool complete() { return status_complete(read_status()); }

        // Host-side consistency check of the queues and the arena between
        // launches (the role of the reference's runtime_overview,
        // program/async.h:2766-2883); see hz::audit_snapshot. Call only while
        // no launch of this instance runs. `allow_held` says whether the
        // specification awaits on barriers at all.
        hz::AsyncCensus audit(bool allow_held = false) {
            hz::AsyncSnapshot snap;
            snap.status = read_status();
            snap.capacity = queue_capacity;
            snap.arena_links = arena_size;
            snap.ops = OPS;
            snap.levels = LEVELS;
            const size_t rings = (size_t)queue_count * LEVELS;
            snap.queues.resize(rings + 1u);
            util::host::auto_throw(hipMemcpy(snap.queues.data(), queues.data(), sizeof(hz::LinkQueue) * rings,
                                             hipMemcpyDeviceToHost));
            util::host::auto_throw(hipMemcpy(&snap.queues[rings], free_queue.data(), sizeof(hz::LinkQueue),
                                             hipMemcpyDeviceToHost));
            snap.slots.resize((rings + 1u) * queue_capacity);
            util::host::auto_throw(hipMemcpy(snap.slots.data(), queue_slots.data(),
                                             sizeof(unsigned int) * rings * queue_capacity, hipMemcpyDeviceToHost));
            util::host::auto_throw(hipMemcpy(snap.slots.data() + rings * queue_capacity, free_slots.data(),
                                             sizeof(unsigned int) * queue_capacity, hipMemcpyDeviceToHost));
            const unsigned int claimed = snap.status.claim < arena_size ? snap.status.claim : arena_size;
            snap.meta.resize(claimed);
            snap.aux.resize(2 * (size_t)claimed);
            if (claimed) {
                util::host::auto_throw(hipMemcpy(snap.meta.data(), link_meta.data(), sizeof(unsigned int) * claimed,
                                                 hipMemcpyDeviceToHost));
                util::host::auto_throw(hipMemcpy(snap.aux.data(), link_aux.data(), 2 * sizeof(unsigned int) * claimed,
                                                 hipMemcpyDeviceToHost));
            }
            return hz::audit_snapshot(snap, allow_held);
        }

        void clear_flags() {
            unsigned int zero = 0;
            util::host::auto_throw(hipMemcpy(&status.data()->flags, &zero, sizeof(zero), hipMemcpyHostToDevice));
        }

        unsigned long long int promises() { return read_status().promises; }
    };

    // Reads a status copied back from the device: raises if the run lost
    // promises or a bounded wait gave up, and returns true once every promise
    // has been evaluated and make_work is exhausted.
    static bool status_complete(const hz::AsyncStatus& st) {
        if (st.flags & ARENA_EXHAUSTED_FLAG) {
            throw std::runtime_error(
                "AsyncProgram: the link arena was exhausted and promises were lost; "
                "construct the Instance with a larger arena");
        }
        if (st.flags & USAGE_FLAG) {
            throw std::runtime_error(
                "AsyncProgram: a Return forwarded a future that cannot hold promises "
                "(a unit future that was not ready)");
        }
        if (st.flags & QUEUE_STALL_FLAG) {
            char detail[64];
            std::snprintf(detail, sizeof(detail), "%016llx", (unsigned long long)st.stall_detail);
            throw std::runtime_error(std::string("AsyncProgram: a bounded wait gave up (") +
                                     hz::stall_name(st.stall_site) + ", detail 0x" + detail + ")");
        }
        return (st.flags & COMPLETION_FLAG) != 0;
    }

    DeviceContext& _dev_ctx;
    GroupContext& _grp_ctx;
    ThreadContext& _thd_ctx;

    DeviceState& device;
    GroupState& group;
    ThreadState& thread;

    // The group's staging block for prefetched lazy values, [word][lane]
    // (a __shared__ object of its own, so that waiting for the loads into it
    // does not hold up the stash's LDS traffic).
    uint32_t (*_lazy_stage)[64] = nullptr;

    __device__ AsyncProgram(DeviceContext& d_c, GroupContext& g_c, ThreadContext& t_c, DeviceState& d,
                            GroupState& g, ThreadState& t)
        : _dev_ctx(d_c), _grp_ctx(g_c), _thd_ctx(t_c), device(d), group(g), thread(t) {}

   protected:
    // ---------------------------------------------------------------- status
    __device__ void set_flags(unsigned int bits) { hz::wave::atomic_or(&_dev_ctx.status->flags, bits); }

    // Records the first bounded wait that gave up and flags the run.
    __device__ HZ_INLINE void stall(unsigned int site, unsigned long long detail) {
        if (atomicCAS(&_dev_ctx.status->stall_site, 0u, site) == 0u) {
            hz::wave::atomic_store(&_dev_ctx.status->stall_detail, detail);
        }
        set_flags(QUEUE_STALL_FLAG);
    }

    __device__ HZ_INLINE unsigned long long queue_detail(hz::LinkQueue* q, unsigned int pos) {
        unsigned long long qi = q == _dev_ctx.free_queue ? 0xFFFFFFFFull : (unsigned long long)(q - _dev_ctx.queues);
        return (qi << 32) | pos;
    }

    // Memory ordering. On gfx950 an agent-scope release is an L2 write-back
    // (buffer_wbl2) and an acquire an invalidation (buffer_inv) that other
    // wavefronts on the XCD then pay for in cache misses, so the runtime
    // orders only what carries data: a link is published by one release
    // after its words and metadata are written, and consumed by one acquire
    // before they are read. Counters (units, queue counts, heads, tails) are
    // relaxed; the termination check synchronises with the makers count once
    // it reads zero (see exec_cycle).
    __device__ void add_units(int delta) { hz::wave::atomic_add(&_dev_ctx.status->units, delta); }

    // ------------------------------------------------------------ link queues
    // Both functions are called by one lane. `publish`: store the id with
    // release semantics (the caller has not fenced the link's data itself).
    __device__ HZ_INLINE bool queue_push(hz::LinkQueue* q, unsigned int* slots, unsigned int id,
                                         bool publish = true) {
        unsigned int t = hz::wave::atomic_add(&q->tail, 1u);
        unsigned int* slot = slots + (t & _dev_ctx.queue_mask);
        uint32_t spins = 0;
        while (hz::wave::atomic_load(slot) != hz::LINK_NONE) {
            if (++spins > SLOT_SPIN_LIMIT) {
                stall(hz::STALL_QUEUE_PUSH, queue_detail(q, t));
                return false;
            }
            hz::wave::sleep_short();
        }
        if (publish) {
            hz::wave::atomic_store_release(slot, id);
        } else {
            hz::wave::atomic_store(slot, id);
        }
        hz::wave::atomic_add(&q->count, 1);
        return true;
    }

    // `consume`: acquire the link's data (the caller reads it afterwards and
    // does not fence itself).
    __device__ HZ_INLINE unsigned int queue_pop(hz::LinkQueue* q, unsigned int* slots, bool consume = false) {
        if (hz::wave::atomic_load(&q->count) <= 0) return hz::LINK_NONE;
        int c = hz::wave::atomic_add(&q->count, -1);
        if (c <= 0) {
            hz::wave::atomic_add(&q->count, 1);
            return hz::LINK_NONE;
        }
        unsigned int h = hz::wave::atomic_add(&q->head, 1u);
        unsigned int* slot = slots + (h & _dev_ctx.queue_mask);
        unsigned int id;
        uint32_t spins = 0;
        while ((id = hz::wave::atomic_load(slot)) == hz::LINK_NONE) {
            if (++spins > SLOT_SPIN_LIMIT) {
                stall(hz::STALL_QUEUE_POP, queue_detail(q, h));
                return hz::LINK_NONE;
            }
            hz::wave::sleep_short();
        }
        if (consume) hz::wave::fence_acquire();
        hz::wave::atomic_store(slot, hz::LINK_NONE);
        return id;
    }

    // Ring r = level * queue_count + queue (level 0 rings are 0..queue_count-1).
    __device__ hz::LinkQueue* work_queue(unsigned int r) { return _dev_ctx.queues + r; }
    __device__ HZ_INLINE unsigned int* work_slots(unsigned int r) {
        return _dev_ctx.queue_slots + (size_t)r * (_dev_ctx.queue_mask + 1u);
    }
    __device__ HZ_INLINE unsigned int ring_of(unsigned int level, unsigned int q) {
        return LEVELS > 1 ? level * _dev_ctx.queue_count + q : q;
    }
    // The i-th level in scheduling order (deepest first unless SHALLOW_FIRST).
    __device__ static HZ_INLINE uint32_t level_in_order(uint32_t i) { return SHALLOW_FIRST ? i : LEVELS - 1u - i; }
    __device__ HZ_INLINE uint32_t link_level_of(uint32_t L) {
        if constexpr (LEVELS > 1) return hz::wave::first(_grp_ctx.link_level[L]);
        return 0u;
    }
    // Index of stash link L's (operation, level) in the open-link table.
    __device__ HZ_INLINE uint32_t open_index(uint32_t L) {
        return hz::wave::first(_grp_ctx.link_op[L]) * LEVELS + link_level_of(L);
    }

    // ------------------------------------------------------------ arena links
    // Called by one lane. Returns LINK_NONE when the arena is exhausted.
    __device__ HZ_INLINE unsigned int arena_alloc_lane() {
        GroupContext& g = _grp_ctx;
        if (g.spare_count > 0) {
            g.spare_count -= 1;
            return g.spare[g.spare_count];
        }
        unsigned int id = queue_pop(_dev_ctx.free_queue, _dev_ctx.free_slots);
        if (id != hz::LINK_NONE) return id;
        unsigned int claim = hz::wave::atomic_load(&_dev_ctx.status->claim);
        if (claim < _dev_ctx.links) {
            claim = hz::wave::atomic_add(&_dev_ctx.status->claim, 1u);
            if (claim < _dev_ctx.links) return claim;
        }
        // Another group may have just recycled one.
        return queue_pop(_dev_ctx.free_queue, _dev_ctx.free_slots);
    }

    __device__ HZ_INLINE void arena_free_lane(unsigned int id) {
        GroupContext& g = _grp_ctx;
        if (g.spare_count < SPARES) {
            g.spare[g.spare_count] = id;
            g.spare_count += 1;
        } else {
            queue_push(_dev_ctx.free_queue, _dev_ctx.free_slots, id, false);
        }
    }

    __device__ unsigned int* arena_link(unsigned int id) { return _dev_ctx.arena + (size_t)id * (WORDS * 64u); }

    // -------------------------------------------------------------- the stash
    __device__ static uint32_t lowest(uint32_t mask) { return (uint32_t)__builtin_ctz(mask); }

    // The ready link of `mask` (non-empty) that filled first (OLDEST_FIRST).
    __device__ HZ_INLINE uint32_t oldest(uint32_t mask) {
        uint32_t best_link = lowest(mask), best = 0xFFFFFFFFu;
        for (uint32_t m = mask; m != 0u; m &= m - 1u) {
            uint32_t i = lowest(m);
            uint32_t q = hz::wave::first(_grp_ctx.link_seq[i]);
            if (q < best) {
                best = q;
                best_link = i;
            }
        }
        return best_link;
    }

    // Waits until this wavefront's earlier memory operations have completed:
    // write-through stores are then at the device's coherence point, so a
    // later relaxed store that publishes them needs no release fence (which
    // would write back the XCD's whole L2). The signal fences keep the
    // compiler from moving memory operations across the wait.
    __device__ static HZ_INLINE void stores_done() {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }

    // Moves stash link L to the arena and queues it. Cooperative over the
    // active lanes (all 64 in the scheduler loop, a subset inside an async
    // call). Returns false if the arena is exhausted.
    __device__ HZ_INLINE bool spill_link(uint32_t L) {
        GroupContext& g = _grp_ctx;
        uint64_t act = hz::wave::active();
        uint32_t n = hz::wave::popc(act);
        uint32_t r = hz::wave::rank(act);
        unsigned int id = 0;
        if (r == 0) id = arena_alloc_lane();
        id = hz::wave::first(id);
        if (id == hz::LINK_NONE) {
            if (r == 0) set_flags(ARENA_EXHAUSTED_FLAG);
            return false;
        }
        unsigned int* dst = arena_link(id);
        const uint32_t* src = &g.words[L][0];
        // The link's words and metadata go out as write-through stores
        // (agent-scope relaxed), and the wavefront waits for them before the
        // id is published; the taker acquires before it reads them. A release
        // fence here wrote back the XCD's whole L2 on every spill, which cost
        // the 64-graph BFS batch a quarter of its time (38.4-38.7 against
        // 27.9-28.9 ms).
        if (n == 64) {
#pragma unroll
            for (uint32_t w = 0; w < WORDS; ++w)
                __hip_atomic_store(&dst[w * 64u + r], src[w * 64u + r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            for (uint32_t i = r; i < WORDS * 64u; i += n)
                __hip_atomic_store(&dst[i], src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        uint32_t op = hz::wave::first(g.link_op[L]);
        uint32_t cnt = hz::wave::first(g.link_count[L]);
        uint32_t lv = link_level_of(L);
        if constexpr (HAS_LAZY) {
            if (hz::wave::first(g.pf_link) == L) g.pf_link = hz::LINK_NONE;
        }
        if (r == 0) {
            __hip_atomic_store(&_dev_ctx.link_meta[id], op | (lv << 8) | (cnt << 16), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        stores_done();
        if (r == 0) {
            add_units(1);
            const uint32_t ring = ring_of(lv, g.home);
            queue_push(work_queue(ring), work_slots(ring), id, false);
            g.free_mask |= (1u << L);
            g.ready_mask &= ~(1u << L);
            g.spills += 1;
        }
        hz::wave::sync();
        return true;
    }

    // Spills ready links until at least `want` stash links are free.
    __device__ HZ_INLINE void spill_until_free(uint32_t want) {
        GroupContext& g = _grp_ctx;
        while (true) {
            uint32_t fm = hz::wave::first(g.free_mask);
            if ((uint32_t)__builtin_popcount(fm) >= want) return;
            uint32_t rm = hz::wave::first(g.ready_mask);
            if (rm == 0) return;
            if (!spill_link(lowest(rm))) return;
        }
    }

    // Takes a free stash link, spilling a ready one if none is free. With
    // at least OPS+3 links and at most one open link per operation, a full
    // stash always has a ready link. Returns LINK_NONE only if the arena is
    // exhausted (the flag is then set and the caller drops its promises).
    __device__ HZ_INLINE uint32_t take_link() {
        GroupContext& g = _grp_ctx;
        uint32_t fm = hz::wave::first(g.free_mask);
        if (fm == 0) {
            uint32_t rm = hz::wave::first(g.ready_mask);
            if (rm == 0 || !spill_link(lowest(rm))) return hz::LINK_NONE;
            fm = hz::wave::first(g.free_mask);
        }
        uint32_t L = lowest(fm);
        g.free_mask = fm & ~(1u << L);
        hz::wave::sync();
        return L;
    }

    // Pulls up to `max_links` queued links into the stash with all 64 lanes.
    __device__ HZ_INLINE uint32_t pull_links(uint32_t max_links) {
        GroupContext& g = _grp_ctx;
        const uint32_t lane = hz::wave::lane();
        uint32_t pulled = 0;
        for (uint32_t k = 0; k < max_links; ++k) {
            uint32_t fm = hz::wave::first(g.free_mask);
            if (fm == 0) break;
            unsigned int id = hz::LINK_NONE;
            if (lane == 0) {
                uint32_t qn = _dev_ctx.queue_count;
                for (uint32_t li = 0; li < LEVELS && id == hz::LINK_NONE; ++li) {
                    const uint32_t lv = level_in_order(li);
                    for (uint32_t i = 0; i < qn && id == hz::LINK_NONE; ++i) {
                        uint32_t q = g.home + i;
                        if (q >= qn) q -= qn;
                        const uint32_t ring = ring_of(lv, q);
                        id = queue_pop(work_queue(ring), work_slots(ring));
                    }
                }
            }
            id = hz::wave::first(id);
            if (id == hz::LINK_NONE) break;
            hz::wave::fence_acquire();
            uint32_t L = lowest(fm);
            const unsigned int* src = arena_link(id);
#pragma unroll
            for (uint32_t w = 0; w < WORDS; ++w) g.words[L][w * 64u + lane] = src[w * 64u + lane];
            uint32_t meta = hz::wave::first(_dev_ctx.link_meta[id]);
            if (lane == 0) {
                g.link_op[L] = meta & 0xFFu;
                if constexpr (LEVELS > 1) g.link_level[L] = (meta >> 8) & 0xFFu;
                g.link_count[L] = meta >> 16;
                g.free_mask = fm & ~(1u << L);
                g.ready_mask |= (1u << L);
                if constexpr (OLDEST_FIRST) g.link_seq[L] = g.seq++;
                g.fills += 1;
                arena_free_lane(id);
                if (g.holding) {
                    add_units(-1);
                } else {
                    g.holding = 1;
                }
            }
            hz::wave::sync();
            pulled++;
        }
        return pulled;
    }

    // Entries of GroupContext::open.
    __device__ static HZ_INLINE uint32_t open_entry(uint32_t link, uint32_t count) { return link | (count << 8); }
    __device__ static HZ_INLINE uint32_t open_link(uint32_t entry) {
        return entry == hz::LINK_NONE ? hz::LINK_NONE : (entry & 0xFFu);
    }
    __device__ static HZ_INLINE uint32_t open_count(uint32_t entry) { return entry >> 8; }

    // Appends one promise per active lane to the open link of operation D at
    // depth level lv (the same for every active lane).
    __device__ HZ_INLINE void stash_append(uint32_t D, uint32_t lv, const uint32_t* words) {
        GroupContext& g = _grp_ctx;
        uint64_t act = hz::wave::active();
        uint32_t n = hz::wave::popc(act);
        uint32_t r = hz::wave::rank(act);
        const uint32_t slot = D * LEVELS + lv;
        const uint32_t entry = hz::wave::first(g.open[slot]);
        uint32_t L, c = 0;
        _thd_ctx.appended += 1u;
        if (entry == hz::LINK_NONE) {
            L = take_link();
            if (L == hz::LINK_NONE) return;
            g.link_op[L] = D;
            if constexpr (LEVELS > 1) g.link_level[L] = lv;
            g.link_count[L] = 0;
            g.open[slot] = open_entry(L, 0u);
            hz::wave::sync();
        } else {
            L = open_link(entry);
            c = open_count(entry);
        }
        uint32_t room = 64u - c;
        if (n < room) {
            // The common case, decided by a wave-uniform test: every active
            // lane writes its promise and the (uniform) new count, with no
            // per-lane predication for the scalar unit to set up.
#pragma unroll
            for (uint32_t w = 0; w < WORDS; ++w) g.words[L][word_index(c + r, w)] = words[w];
            g.link_count[L] = c + n;
            g.open[slot] = open_entry(L, c + n);
            hz::wave::sync();
            return;
        }
        if (r < room) {
#pragma unroll
            for (uint32_t w = 0; w < WORDS; ++w) g.words[L][word_index(c + r, w)] = words[w];
        }
        // L is now full: close it, then open another for the overflow.
        uint32_t rmask = hz::wave::first(g.ready_mask);
        g.link_count[L] = 64u;
        g.ready_mask = rmask | (1u << L);
        if constexpr (OLDEST_FIRST) {
            uint32_t q = hz::wave::first(g.seq);
            g.link_seq[L] = q;
            g.seq = q + 1u;
        }
        g.open[slot] = hz::LINK_NONE;
        hz::wave::sync();
        uint32_t rest = n - room;
        if (rest == 0) return;
        uint32_t L2 = take_link();
        if (L2 == hz::LINK_NONE) return;
        if (r >= room) {
#pragma unroll
            for (uint32_t w = 0; w < WORDS; ++w) g.words[L2][word_index(r - room, w)] = words[w];
        }
        g.link_op[L2] = D;
        if constexpr (LEVELS > 1) g.link_level[L2] = lv;
        g.link_count[L2] = rest;
        g.open[slot] = open_entry(L2, rest);
        hz::wave::sync();
    }

    // Depth level of a promise queued with `depth_delta` by the promise being
    // evaluated (or by make_work, at level 0), clamped to the stack.
    __device__ HZ_INLINE uint32_t child_level(int depth_delta) {
        if constexpr (LEVELS == 1) {
            return 0u;
        } else {
            int lv = (int)hz::wave::first(_grp_ctx.level) + depth_delta;
            lv = lv < 0 ? 0 : lv;
            return lv >= (int)LEVELS ? LEVELS - 1u : (uint32_t)lv;
        }
    }

   public:
    // Queues a promise at the current depth level plus depth_delta (reference
    // async_call_cast, program/async.h:1627-1660); async<Op>(...) is the
    // depth_delta = 0 form. A delta is meaningful only with STACK_SIZE levels.
    template <typename TYPE>
    __device__ HZ_INLINE void async_call_cast(int depth_delta, Promise<TYPE> param_value) {
        static_assert(Lookup<TYPE>::type::CONTAINED, "operation is not part of this program's OpSet");
        constexpr OpDisc D = Lookup<TYPE>::type::DISC;
        PromiseUnionType pu;
        pu.template cast<TYPE>() = param_value;
        uint32_t words[WORDS];
        __builtin_memcpy(words, &pu, sizeof(PromiseUnionType));
        if constexpr (FORWARD) {
            if (depth_delta == 0 && hz::wave::first(_thd_ctx.fwd_on) != 0u && _thd_ctx.fwd_op == hz::LINK_NONE) {
                _thd_ctx.fwd_op = D;
#pragma unroll
                for (uint32_t w = 0; w < WORDS; ++w) _thd_ctx.fwd_words[w] = words[w];
                return;
            }
        }
        stash_append(D, child_level(depth_delta), words);
    }

    // Depth level of the promise being evaluated (0 in make_work and without
    // STACK_SIZE levels).
    __device__ HZ_INLINE unsigned int depth() {
        if constexpr (LEVELS == 1) return 0u;
        return hz::wave::first(_grp_ctx.level);
    }

    // Links this group has pulled from the device queues in this launch. A
    // specification that caches per-group data about the promises it holds
    // can tell from it whether promises have arrived from elsewhere.
    __device__ HZ_INLINE unsigned long long links_pulled() { return hz::wave::first((uint64_t)_grp_ctx.fills); }

    // Called from make_work (by the whole wavefront) when it found nothing
    // to start but more may come later (a resident program waiting for
    // jobs): the group skips make_work for the next `cycles` scheduling
    // cycles, running its own links meanwhile, and pauses between cycles if
    // it has none. It stays a maker, so the launch does not end.
    __device__ HZ_INLINE void pause_make_work(unsigned int cycles) {
        _thd_ctx.make_pause = cycles;
    }

   protected:

    template <typename TYPE>
    __device__ HZ_INLINE void immediate_call_cast(Promise<TYPE> param_value) {
        param_value(*this);
    }

    // Stash link with the most promises among the open ones, or LINK_NONE.
    // With depth levels, the open links of the first level (in scheduling
    // order) that has any are considered.
    __device__ HZ_INLINE uint32_t fullest_open(uint32_t& count) {
        GroupContext& g = _grp_ctx;
        uint32_t best = hz::LINK_NONE, best_count = 0;
        for (uint32_t li = 0; li < LEVELS && best == hz::LINK_NONE; ++li) {
            const uint32_t lv = level_in_order(li);
            for (uint32_t d = 0; d < OPS; ++d) {
                const uint32_t entry = hz::wave::first(g.open[d * LEVELS + lv]);
                if (entry == hz::LINK_NONE) continue;
                const uint32_t L = open_link(entry);
                const uint32_t c = open_count(entry);
                if (c > best_count) {
                    best = L;
                    best_count = c;
                }
            }
        }
        count = best_count;
        return best;
    }

    // The ready link of `mask` (non-empty) whose level comes first in
    // scheduling order (depth levels), lowest stash position among equals.
    __device__ HZ_INLINE uint32_t first_by_level(uint32_t mask) {
        uint32_t best_link = lowest(mask), best = 0xFFFFFFFFu;
        for (uint32_t m = mask; m != 0u; m &= m - 1u) {
            uint32_t i = lowest(m);
            uint32_t lv = link_level_of(i);
            uint32_t rank = SHALLOW_FIRST ? lv : LEVELS - 1u - lv;
            if (rank < best) {
                best = rank;
                best_link = i;
            }
        }
        return best_link;
    }

    __device__ HZ_INLINE uint32_t pick_ready(uint32_t mask) {
        if constexpr (LEVELS > 1) return first_by_level(mask);
        return OLDEST_FIRST ? oldest(mask) : lowest(mask);
    }

    __device__ bool stash_empty() { return hz::wave::first(_grp_ctx.free_mask) == ALL_LINKS; }

    __device__ HZ_INLINE void drop_holding_if_empty() {
        GroupContext& g = _grp_ctx;
        if (stash_empty() && hz::wave::first(g.holding)) {
            if (hz::wave::lane() == 0) {
                g.holding = 0;
                add_units(-1);
            }
            hz::wave::sync();
        }
    }

    // The group joins `units` only if make_work left promises in its stash,
    // so a call that made nothing (an idle resident program polling for
    // jobs) costs no atomic. While make_work runs the group is still a maker,
    // which blocks termination by itself; the units increment precedes the
    // release that gives up the maker role.
    __device__ HZ_INLINE void make_work_step() {
        GroupContext& g = _grp_ctx;
        const uint32_t lane = hz::wave::lane();
        _thd_ctx.made += 1u;
        if constexpr (LEVELS > 1) g.level = 0u;
        bool mine;
        [[clang::always_inline]] mine = PROGRAM_SPEC::make_work(*this);
        bool more = hz::wave::ballot(mine) != 0;
        const bool holds = !stash_empty();
        if (lane == 0) {
            if (holds && !g.holding) {
                g.holding = 1;
                add_units(1);
            }
            g.can_make_work = more ? 1u : 0u;
            if (!more && g.maker) {
                g.maker = 0;
                hz::wave::atomic_add_acqrel(&_dev_ctx.status->makers, -1);
            }
        }
        hz::wave::sync();
        drop_holding_if_empty();
    }

    // ------------------------------------------------- lazy argument loads
    // Word offset of operation I's first LazyLoad member (its source pointer
    // and its value) inside a promise.
    template <unsigned I>
    __device__ static HZ_INLINE uint32_t lazy_source_word() {
        PromiseUnionType pu;
        auto& z = LazyFacts<I>::First::of(pu.template cast<typename LazyFacts<I>::Op>().args);
        return (uint32_t)(((const char*)&z.source - (const char*)&pu) / 4);
    }

    // Issues the direct-to-LDS loads of stash link N's lazy values (operation
    // I), lane i's into column i of the staging block.
    template <unsigned I = 0>
    __device__ HZ_INLINE void prefetch_lazy(uint32_t N, uint32_t op, uint32_t cnt) {
        if constexpr (I < OPS) {
            if constexpr (LazyFacts<I>::HAS) {
                if (op == I) {
                    const uint32_t lane = hz::wave::lane();
                    const uint32_t sw = lazy_source_word<I>();
                    if (lane < cnt) {
                        const uint64_t lo = _grp_ctx.words[N][word_index(lane, sw)];
                        const uint64_t hi = _grp_ctx.words[N][word_index(lane, sw + 1u)];
                        const uint32_t* src = (const uint32_t*)(uintptr_t)((hi << 32) | lo);
#pragma unroll
                        for (uint32_t k = 0; k < LazyFacts<I>::VALUE_WORDS; ++k) {
                            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + k),
                                                             (__attribute__((address_space(3))) void*)&_lazy_stage[k][0],
                                                             4, 0, 0);
                        }
                    }
                    return;
                }
            }
            prefetch_lazy<I + 1>(N, op, cnt);
        }
    }

    // Hands this lane's prefetched lazy value to its promise (operation I).
    template <unsigned I = 0>
    __device__ HZ_INLINE void take_lazy(PromiseUnionType& pu, uint32_t op, const uint32_t* staged) {
        if constexpr (I < OPS) {
            if constexpr (LazyFacts<I>::HAS) {
                if (op == I) {
                    auto& z = LazyFacts<I>::First::of(pu.template cast<typename LazyFacts<I>::Op>().args);
                    typename LazyFacts<I>::First::Value v;
                    __builtin_memcpy(&v, staged, sizeof(v));
                    z.fill(v);
                    return;
                }
            }
            take_lazy<I + 1>(pu, op, staged);
        }
    }

    template <unsigned I = 0>
    __device__ static HZ_INLINE bool op_has_lazy(uint32_t op) {
        if constexpr (I < OPS) {
            return (op == I && LazyFacts<I>::HAS) || op_has_lazy<I + 1>(op);
        }
        return false;
    }

    __device__ HZ_INLINE void run_link(uint32_t L) {
        GroupContext& g = _grp_ctx;
        const uint32_t lane = hz::wave::lane();
        uint32_t op = hz::wave::first(g.link_op[L]);
        uint32_t cnt = hz::wave::first(g.link_count[L]);
        uint32_t words[WORDS];
        uint32_t staged[HAS_LAZY ? LAZY_WORDS : 1];
        bool prefetched = false;
        {
            HZ_PHASE(PHASE_READ);
            if (lane < cnt) {
#pragma unroll
                for (uint32_t w = 0; w < WORDS; ++w) words[w] = g.words[L][word_index(lane, w)];
            }
            if constexpr (HAS_LAZY) {
                prefetched = hz::wave::first(g.pf_link) == L;
                if (prefetched && lane < cnt) {
#pragma unroll
                    for (uint32_t k = 0; k < LAZY_WORDS; ++k) staged[k] = _lazy_stage[k][lane];
                }
                // The staging block is read before the next link's loads refill it.
                __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            }
            hz::wave::sync();
            // All lanes store the same values (uniform read-modify-writes of
            // one wavefront): no lane-0 predication on the hot path.
            uint32_t fm = hz::wave::first(g.free_mask);
            uint32_t rmask = hz::wave::first(g.ready_mask);
            g.free_mask = fm | (1u << L);
            g.ready_mask = rmask & ~(1u << L);
            _thd_ctx.executed += cnt;
            _thd_ctx.runs += 1u;
            if constexpr (LEVELS > 1) g.level = link_level_of(L);
            if constexpr (HAS_LAZY) {
                // Start the loads of the link most likely to run next.
                uint32_t next = hz::LINK_NONE;
                const uint32_t rest = rmask & ~(1u << L);
                if (rest != 0u) {
                    const uint32_t N = pick_ready(rest);
                    const uint32_t nop = hz::wave::first(g.link_op[N]);
                    if (op_has_lazy(nop)) {
                        prefetch_lazy(N, nop, hz::wave::first(g.link_count[N]));
                        next = N;
                    }
                }
                g.pf_link = next;
            }
            hz::wave::sync();
        }
        {
            HZ_PHASE(PHASE_EVAL);
            if constexpr (FORWARD) _thd_ctx.fwd_on = 1u;
            if (lane < cnt) {
                PromiseUnionType pu;
                __builtin_memcpy(&pu, words, sizeof(PromiseUnionType));
                if constexpr (HAS_LAZY) {
                    if (prefetched) take_lazy(pu, op, staged);
                }
                pu.template loose_eval<ProgramType>(*this, op);
            }
            hz::wave::sync();
            if constexpr (FORWARD) run_forwarded();
        }
    }

    // Runs the promises the lanes hold in registers while they form a link
    // of one operation of at least FORWARD_MIN lanes (at most FORWARD_RUNS
    // such links), then appends what is still held to the stash.
    __device__ HZ_INLINE void run_forwarded() {
        for (uint32_t k = 0; k < FORWARD_RUNS; ++k) {
            const uint64_t held = hz::wave::ballot(_thd_ctx.fwd_op != hz::LINK_NONE);
            if (held == 0ull) break;
            const uint32_t D = hz::wave::read_lane(_thd_ctx.fwd_op, (uint32_t)__builtin_ctzll(held));
            const uint64_t same = hz::wave::ballot(_thd_ctx.fwd_op == D);
            const uint32_t n = (uint32_t)__builtin_popcountll(same);
            if (same != held || n < FORWARD_MIN) break;
            uint32_t words[WORDS];
#pragma unroll
            for (uint32_t w = 0; w < WORDS; ++w) words[w] = _thd_ctx.fwd_words[w];
            const bool mine = _thd_ctx.fwd_op != hz::LINK_NONE;
            _thd_ctx.fwd_op = hz::LINK_NONE;
            _thd_ctx.executed += n;
            _thd_ctx.runs += 1u;
            _thd_ctx.forwarded += n;
            if (mine) {
                _thd_ctx.appended += 1u;  // counted as written into a link, as an append would
                PromiseUnionType pu;
                __builtin_memcpy(&pu, words, sizeof(PromiseUnionType));
                pu.template loose_eval<ProgramType>(*this, D);
            }
            hz::wave::sync();
        }
        _thd_ctx.fwd_on = 0u;
        uint64_t held = hz::wave::ballot(_thd_ctx.fwd_op != hz::LINK_NONE);
        while (held != 0ull) {
            const uint32_t D = hz::wave::read_lane(_thd_ctx.fwd_op, (uint32_t)__builtin_ctzll(held));
            const bool mine = _thd_ctx.fwd_op == D;
            held &= ~hz::wave::ballot(mine);
            if (mine) {
                uint32_t words[WORDS];
#pragma unroll
                for (uint32_t w = 0; w < WORDS; ++w) words[w] = _thd_ctx.fwd_words[w];
                stash_append(D, 0u, words);
            }
            hz::wave::sync();
        }
        _thd_ctx.fwd_op = hz::LINK_NONE;
    }

    // True if the group's home queue holds links (at some depth level).
    __device__ HZ_INLINE bool queues_hint() {
        const uint32_t home = hz::wave::first(_grp_ctx.home);
        int32_t any = 0;
        for (uint32_t lv = 0; lv < LEVELS && any <= 0; ++lv) {
            hz::LinkQueue* q = work_queue(ring_of(lv, home));
            any = hz::wave::first((int32_t)hz::wave::atomic_load(&q->count));
        }
        return any > 0;
    }

    enum CycleResult { RUN = 0, DONE = 1 };

    // Work sharing (specifications with SHARE_WORK). A group's promises stay
    // in its LDS stash until it runs short of free links, so work that grows
    // from a few seeds (a BFS frontier) would stay with the few groups that
    // started it. Groups without work count themselves in status->hungry; a
    // group about to run link `keep` that holds other links consults that
    // count every SHARE_CHECK_PERIOD cycles and spills up to that many of
    // them (full links first, then its fullest open link) to the queues.
    static constexpr uint32_t SHARE_CHECK_PERIOD = 4;

    __device__ HZ_INLINE void set_hungry(bool h) {
        if constexpr (!SHARE_WORK) return;
        GroupContext& g = _grp_ctx;
        if ((hz::wave::first(g.hungry) != 0) == h) return;
        if (hz::wave::lane() == 0) {
            g.hungry = h ? 1u : 0u;
            hz::wave::atomic_add(&_dev_ctx.status->hungry, h ? 1 : -1);
        }
        hz::wave::sync();
    }

    __device__ HZ_INLINE void share_links(uint32_t keep) {
        if constexpr (!SHARE_WORK) return;
        GroupContext& g = _grp_ctx;
        const uint32_t others = ALL_LINKS & ~hz::wave::first(g.free_mask) & ~(1u << keep);
        if (others == 0) return;
        uint32_t countdown = hz::wave::first(g.share_check);
        if (hz::wave::lane() == 0) g.share_check = countdown == 0 ? SHARE_CHECK_PERIOD - 1 : countdown - 1;
        hz::wave::sync();
        if (countdown != 0) return;
        int hungry = 0;
        if (hz::wave::lane() == 0) hungry = hz::wave::atomic_load(&_dev_ctx.status->hungry);
        hungry = hz::wave::first(hungry);
        for (int n = 0; n < hungry; ++n) {
            uint32_t L = hz::LINK_NONE;
            uint32_t rm = hz::wave::first(g.ready_mask) & ~(1u << keep);
            if (rm != 0) {
                L = lowest(rm);
            } else {
                uint32_t best = 0;
                for (uint32_t d = 0; d < OPS * LEVELS; ++d) {
                    const uint32_t entry = hz::wave::first(g.open[d]);
                    if (entry == hz::LINK_NONE) continue;
                    const uint32_t o = open_link(entry);
                    if (o == keep) continue;
                    const uint32_t c = open_count(entry);
                    if (c > best) {
                        best = c;
                        L = o;
                    }
                }
                if (L != hz::LINK_NONE) {
                    g.open[open_index(L)] = hz::LINK_NONE;
                    hz::wave::sync();
                }
            }
            if (L == hz::LINK_NONE || !spill_link(L)) break;
        }
    }

    // One scheduling cycle (reference exec_cycle, program/async.h:2155-2319).
    // Without a ready link, queued links take priority over making new work
    // (that bounds the population parked in the arena), but a group that can
    // still make work consults the device queues (a device-scope load) only on
    // its first such cycle of a launch and then every QUEUE_CHECK_PERIOD-th.
    static constexpr uint32_t QUEUE_CHECK_PERIOD = 16;

    __device__ HZ_INLINE CycleResult exec_cycle() {
        GroupContext& g = _grp_ctx;
        HZ_PHASE(PHASE_CYCLE);
        {
            HZ_PHASE(PHASE_SPILL);
            spill_until_free(MARGIN);
        }

        uint32_t rm = hz::wave::first(g.ready_mask);
        const uint32_t pause = hz::wave::first(_thd_ctx.make_pause);
        if (pause != 0u) _thd_ctx.make_pause = pause - 1u;
        if (rm == 0) {
            bool check = !hz::wave::first(g.can_make_work);
            if (!check) {
                uint32_t countdown = hz::wave::first(_thd_ctx.queue_check);
                check = countdown == 0;
                _thd_ctx.queue_check = check ? QUEUE_CHECK_PERIOD - 1 : countdown - 1;
            }
            if (check) {
                HZ_PHASE(PHASE_PULL);
                if (queues_hint()) {
                    pull_links(1);
                    rm = hz::wave::first(g.ready_mask);
                    _thd_ctx.queue_check = 0;  // keep draining while links remain
                }
            }
            // Make work while no link is full, bounded so open links still run.
            HZ_PHASE(PHASE_MAKE);
            for (uint32_t tries = 0; rm == 0 && tries < 4 && pause == 0u && hz::wave::first(g.can_make_work) &&
                                     hz::wave::first(_thd_ctx.make_pause) == 0u;
                 ++tries) {
                uint32_t c = 0;
                fullest_open(c);
                if (tries > 0 && c >= _dev_ctx.coalesce) break;
                make_work_step();
                rm = hz::wave::first(g.ready_mask);
            }
        }

        uint32_t L = hz::LINK_NONE;
        {
            HZ_PHASE(PHASE_PICK);
            if (rm != 0) {
                L = pick_ready(rm);
            } else {
                uint32_t c = 0;
                L = fullest_open(c);
                if (L != hz::LINK_NONE) {
                    g.open[open_index(L)] = hz::LINK_NONE;
                    hz::wave::sync();
                }
            }
        }

        if (L != hz::LINK_NONE) {
            _thd_ctx.idle_streak = 0;
            set_hungry(false);
            share_links(L);
            run_link(L);
            // Further ready links run in the same cycle while the margin of
            // free links holds, skipping the cycle's other checks.
            for (uint32_t b = 1; b < BURST; ++b) {
                uint32_t rm2 = hz::wave::first(g.ready_mask);
                if (rm2 == 0) break;
                uint32_t fm2 = hz::wave::first(g.free_mask);
                if ((uint32_t)__builtin_popcount(fm2) < MARGIN) break;
                run_link(pick_ready(rm2));
            }
            // A group whose stash emptied stops counting as holding work once
            // per cycle, not after every link (Collatz 66.8 -> 64.5 ms).
            drop_holding_if_empty();
            return RUN;
        }

        // Nothing local: look for queued links anywhere, then test for quiescence.
        HZ_PHASE(PHASE_IDLE);
        if (pull_links(1) > 0) return RUN;
        if (hz::wave::first(g.can_make_work)) {
            if (pause != 0u) __builtin_amdgcn_s_sleep(8);
            return RUN;
        }
        set_hungry(true);
        uint32_t streak = hz::wave::first(_thd_ctx.idle_streak);
        if (streak >= IDLE_EXIT_POLLS) return DONE;
        // makers only falls to zero after every maker's last units increment
        // (its decrement is a release), so once a relaxed read of makers sees
        // zero, one acquire makes those increments visible; from then on
        // units is read relaxed: with no makers it can only fall, and it
        // stays at zero once there.
        int makers = 0, units = 0;
        unsigned int flags = 0;
        if (hz::wave::lane() == 0) {
            if (!g.no_makers) {
                makers = hz::wave::atomic_load(&_dev_ctx.status->makers);
                if (makers == 0) {
                    hz::wave::fence_acquire();
                    g.no_makers = 1;
                }
            }
            units = hz::wave::atomic_load(&_dev_ctx.status->units);
            flags = hz::wave::atomic_load(&_dev_ctx.status->flags);
        }
        _thd_ctx.idle += 1u;
        _thd_ctx.idle_streak = streak + 1u;
        makers = hz::wave::first(makers);
        units = hz::wave::first(units);
        flags = hz::wave::first(flags);
        if ((makers == 0 && units == 0) || (flags & ~COMPLETION_FLAG) != 0) return DONE;
        // Pause between idle polls; each poll reads queue and status lines
        // that busy groups update. The pause may grow with the idle streak
        // (IDLE_BACKOFF_MAX).
        uint32_t pauses = streak / 16u + 1u;
        if (pauses > IDLE_BACKOFF_MAX) pauses = IDLE_BACKOFF_MAX;
        for (uint32_t i = 0; i < pauses; ++i) __builtin_amdgcn_s_sleep(4);
        return RUN;
    }

    __device__ HZ_INLINE void init_group() {
        GroupContext& g = _grp_ctx;
        if (hz::wave::lane() == 0) {
            g.free_mask = ALL_LINKS;
            g.ready_mask = 0;
            g.seq = 0;
            g.level = 0;
            g.pf_link = hz::LINK_NONE;
            for (uint32_t d = 0; d < OPS * LEVELS; ++d) g.open[d] = hz::LINK_NONE;
            g.spare_count = 0;
            g.holding = 0;
            g.can_make_work = 1;
            g.maker = 1;
            g.home = blockIdx.x % _dev_ctx.queue_count;
            g.spills = g.fills = 0;
            g.share_check = 0;
            g.hungry = 0;
            g.no_makers = 0;
#ifdef HZ_PHASE_TIMING
            for (int i = 0; i < 8; ++i) g.phase[i] = 0;
#endif
            hz::wave::atomic_add_acqrel(&_dev_ctx.status->makers, 1);
        }
        hz::wave::sync();
    }

    __device__ HZ_INLINE void init_thread() {
        _thd_ctx.thread_id = blockIdx.x * blockDim.x + threadIdx.x;
        _thd_ctx.rand_state = _thd_ctx.thread_id;
        _thd_ctx.appended = 0;
        _thd_ctx.executed = 0;
        _thd_ctx.runs = 0;
        _thd_ctx.made = 0;
        _thd_ctx.idle = 0;
        _thd_ctx.idle_streak = 0;
        _thd_ctx.queue_check = 0;
        _thd_ctx.make_pause = 0;
        _thd_ctx.fwd_on = 0;
        _thd_ctx.fwd_op = hz::LINK_NONE;
        _thd_ctx.forwarded = 0;
    }

    // Leaves nothing in LDS: every stash link goes back to the arena
    // (reference cleanup_runtime, program/async.h:2323-2393).
    __device__ HZ_INLINE void cleanup_runtime() {
        GroupContext& g = _grp_ctx;
        const uint32_t lane = hz::wave::lane();
        // No load into the staging block may still be in flight when the
        // workgroup's LDS is released.
        if constexpr (HAS_LAZY) __builtin_amdgcn_s_waitcnt(0);
        if (lane == 0) {
            for (uint32_t d = 0; d < OPS * LEVELS; ++d) g.open[d] = hz::LINK_NONE;
        }
        hz::wave::sync();
        while (true) {
            uint32_t used = ALL_LINKS & ~hz::wave::first(g.free_mask);
            if (used == 0) break;
            if (!spill_link(lowest(used))) {
                // Arena exhausted: the flag is set and the promises are lost.
                if (lane == 0) g.free_mask = ALL_LINKS;
                hz::wave::sync();
                break;
            }
        }
        // Sum the lanes' append counts (every lane is active here).
        unsigned long long appended = _thd_ctx.appended;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) appended += __shfl_xor(appended, d, 64);
        if (lane == 0) {
            if (g.hungry) {
                g.hungry = 0;
                hz::wave::atomic_add(&_dev_ctx.status->hungry, -1);
            }
            if (g.holding) {
                g.holding = 0;
                add_units(-1);
            }
            while (g.spare_count > 0) {
                g.spare_count -= 1;
                queue_push(_dev_ctx.free_queue, _dev_ctx.free_slots, g.spare[g.spare_count], false);
            }
            if (g.maker) {
                if (g.can_make_work) set_flags(EARLY_HALT_FLAG);
                g.maker = 0;
                hz::wave::atomic_add_acqrel(&_dev_ctx.status->makers, -1);
            }
            hz::AsyncStatus* st = _dev_ctx.status;
            hz::wave::atomic_add(&st->promises, _thd_ctx.executed);
            hz::wave::atomic_add(&st->links_run, _thd_ctx.runs);
            if (appended) hz::wave::atomic_add(&st->appended, appended);
            if (_thd_ctx.forwarded) hz::wave::atomic_add(&st->forwarded, _thd_ctx.forwarded);
            if (g.spills) hz::wave::atomic_add(&st->spills, g.spills);
            if (g.fills) hz::wave::atomic_add(&st->fills, g.fills);
            if (_thd_ctx.made) hz::wave::atomic_add(&st->make_work_calls, _thd_ctx.made);
            if (_thd_ctx.idle) hz::wave::atomic_add(&st->idle_polls, _thd_ctx.idle);
#ifdef HZ_PHASE_TIMING
            for (int i = 0; i < 8; ++i) hz::wave::atomic_add(&st->phase_cycles[i], g.phase[i]);
#endif
            unsigned int order = hz::wave::atomic_add_acqrel(&st->checkout, 1u);
            if (order == gridDim.x * GROUPS_PER_BLOCK - 1) {
                unsigned int old = hz::wave::atomic_and(&st->flags, ~EARLY_HALT_FLAG);
                int units = hz::wave::atomic_load_acquire(&st->units);
                if ((old & EARLY_HALT_FLAG) == 0 && units == 0) set_flags(COMPLETION_FLAG);
                hz::wave::atomic_add(&st->launches, 1ull);
                hz::wave::atomic_store(&st->checkout, 0u);
            }
        }
        hz::wave::sync();
    }

   public:
    __device__ void init_program() {}

    // The persistent loop (reference program/async.h:2625-2668).
    __device__ HZ_INLINE void exec(unsigned int cycle_count) {
        init_group();
        init_thread();
        // The spec hooks are inlined: a hook left out of line would take the
        // program object and push the state it refers to into scratch memory.
        [[clang::always_inline]] PROGRAM_SPEC::initialize(*this);
        for (unsigned int cycle = 0; cycle < cycle_count; ++cycle) {
            if ((cycle & 63u) == 63u) {
                unsigned int flags = 0;
                if (hz::wave::lane() == 0) flags = hz::wave::atomic_load(&_dev_ctx.status->flags);
                flags = hz::wave::first(flags);
                if (flags & ~COMPLETION_FLAG) break;
            }
            if (exec_cycle() == DONE) break;
        }
        [[clang::always_inline]] PROGRAM_SPEC::finalize(*this);
        cleanup_runtime();
    }

    template <typename TYPE, typename... ARGS>
    __device__ HZ_INLINE void async(ARGS... args) {
        async_call_cast<TYPE>(0, Promise<TYPE>(args...));
    }

    template <typename TYPE>
    __device__ HZ_INLINE void async_call(Promise<TYPE> promise) {
        async_call_cast<TYPE>(0, promise);
    }

    template <typename TYPE, typename... ARGS>
    __device__ HZ_INLINE void sync(ARGS... args) {
        immediate_call_cast<TYPE>(Promise<TYPE>(args...));
    }

    template <typename TYPE>
    __device__ HZ_INLINE void sync_call(Promise<TYPE> promise) {
        immediate_call_cast<TYPE>(promise);
    }

    // Queues a promise held as a PromiseEnum (UnitBarrier release).
    __device__ HZ_INLINE void async_call(PromiseEnum<OpSet> promise) { async_enum<0>(promise); }

    // ------------------------------------------------------ remapping barrier
    // (protocol described in barrier.h). link_aux[2L] = next link in a
    // barrier's full list; link_aux[2L+1] = (target << 16) | filled.

    // Every lane that awaits places its promise within one retry loop whose
    // exit is wave-uniform: a lane waiting for a link that a lane of its own
    // wave installs must see that installation in the same pass. (With a
    // per-lane loop the installing lane's exit branch would run only after
    // all of the wave's lanes had left the loop.)
    template <typename BARRIER, typename OP_TYPE>
    __device__ HZ_INLINE void barrier_await(BARRIER& b, Promise<OP_TYPE> promise) {
        constexpr OpDisc D = Lookup<OP_TYPE>::type::DISC;
        constexpr unsigned int RING = BARRIER::RING;
        bool run_now = hz::wave::atomic_load(&b.semaphore.sem) == 0u;
        bool pending = !run_now;
        unsigned int ticket = 0u;
        if (pending) {
            // One atomic per await, or one per wavefront when every awaiting
            // lane names the same barrier: the lanes then take consecutive
            // tickets in lane order, as separate draws in that order would.
            // Acquire pairs with the release that set RELEASED_BIT, after
            // which the promise runs at once.
            const uint64_t act = hz::wave::active();
            const uint64_t addr = (uint64_t)(uintptr_t)&b.tickets[D];
            if (hz::wave::ballot(addr == hz::wave::first(addr)) == act) {
                const uint32_t r = hz::wave::rank(act);
                unsigned int base = 0u;
                if (r == 0u) base = hz::wave::atomic_add_acqrel(&b.tickets[D], hz::wave::popc(act));
                base = hz::wave::first(base);
                ticket = (base & BARRIER::RELEASED_BIT) ? base : base + r;
            } else {
                ticket = hz::wave::atomic_add_acqrel(&b.tickets[D], 1u);
            }
            if (ticket & BARRIER::RELEASED_BIT) {
                run_now = true;
                pending = false;
            }
        }
        PromiseUnionType pu;
        pu.template cast<OP_TYPE>() = promise;
        uint32_t words[WORDS];
        __builtin_memcpy(words, &pu, sizeof(PromiseUnionType));
        const unsigned int block = ticket / 64u;
        const unsigned int slot = ticket % 64u;
        unsigned long long* entry = &b.ring[D][block % RING];
        bool held = false;  // this lane's promise went into a barrier link
        for (uint32_t attempt = 0; hz::wave::ballot(pending) != 0ull; ++attempt) {
            if (attempt >= SLOT_SPIN_LIMIT) {
                if (pending) stall(hz::STALL_BARRIER_WAIT, hz::wave::atomic_load(entry));
                break;
            }
            if (pending) {
                // Relaxed (coherent) polls: an acquire on every poll would
                // invalidate the L2 each time (the barrier demo took 12.1-12.4
                // instead of 10.6-11.1 ms at 1,024 barriers). The installer
                // needs only the entry's value; a lane that finds its block's
                // link acquires once before it writes into the link.
                const unsigned long long seen = hz::wave::atomic_load(entry);
                if (slot == 0u) {
                    // This lane installs the block's link once the block RING
                    // before it has filled and left the entry.
                    const unsigned long long free_for_me =
                        ((unsigned long long)(block - RING) << 32) | hz::BARRIER_NO_LINK;
                    if (seen == free_for_me) {
                        pending = false;
                        unsigned int fresh = arena_alloc_any();
                        if (fresh == hz::LINK_NONE) {
                            set_flags(ARENA_EXHAUSTED_FLAG);
                        } else {
                            hz::wave::atomic_store(&_dev_ctx.link_aux[2u * fresh + 1u], 64u << 16);
                            write_arena_slot(fresh, 0u, words);
                            // The count word and slot 0 are written before the
                            // block's other lanes read the link from the entry.
                            stores_done();
                            hz::wave::atomic_store(entry, ((unsigned long long)block << 32) | fresh);
                            barrier_finish(b, fresh, D, block);
                            held = true;
                        }
                    }
                } else if ((unsigned int)(seen >> 32) == block && (unsigned int)seen != hz::BARRIER_NO_LINK) {
                    // Pairs with the installer's release: the link's count
                    // word and slot 0 are written.
                    hz::wave::fence_acquire();
                    barrier_write(b, (unsigned int)seen, slot, D, words, block);
                    pending = false;
                    held = true;
                }
            }
            if (hz::wave::ballot(pending) != 0ull) hz::wave::sleep_short();
        }
        if (held) _thd_ctx.appended += 1u;
        if (run_now) {
            // Pairs with the release of the semaphore update that released
            // the barrier, so the promise sees what was written before it.
            hz::wave::fence_acquire();
            async_call_cast<OP_TYPE>(0, promise);
        }
    }

    template <typename BARRIER>
    __device__ HZ_INLINE void barrier_release(BARRIER& b) {
        constexpr unsigned int NOPS = BARRIER::TYPE_COUNT;
        constexpr unsigned int RING = BARRIER::RING;
        for (unsigned int d = 0; d < NOPS; ++d) {
            // Close the ticket counter; the tickets handed out fill whole
            // blocks and, unless T is a multiple of 64, part of block T / 64.
            const unsigned int old = hz::wave::atomic_or(&b.tickets[d], BARRIER::RELEASED_BIT);
            if (old & BARRIER::RELEASED_BIT) continue;
            const unsigned int block = old / 64u, part = old % 64u;
            if (part == 0u) continue;
            // Wait for the last block's link (its slot-0 lane holds a ticket
            // below T, so it installs it), then lower its target to `part`.
            unsigned long long* entry = &b.ring[d][block % RING];
            unsigned long long seen = 0ull;
            bool found = false;
            for (uint32_t spins = 0; spins < SLOT_SPIN_LIMIT; ++spins) {
                seen = hz::wave::atomic_load_acquire(entry);
                if ((unsigned int)(seen >> 32) == block && (unsigned int)seen != hz::BARRIER_NO_LINK) {
                    found = true;
                    break;
                }
                hz::wave::sleep_short();
            }
            if (!found) {
                stall(hz::STALL_BARRIER_INSTALL, seen);
                continue;
            }
            const unsigned int L = (unsigned int)seen;
            const unsigned int before =
                hz::wave::atomic_add_acqrel(&_dev_ctx.link_aux[2u * L + 1u], (part - 64u) << 16);
            if ((before & 0xFFFFu) == part) barrier_complete(b, L, d, part, block);
        }
        const unsigned int head = atomicExch(&b.full_head, hz::BARRIER_RELEASED);
        hz::wave::fence_acquire();
        // No link joins the list after the exchange, so it is counted first:
        // one unit update covers every link before the first is queued, and
        // one tail draw places them all. The links are then placed in chunks,
        // each read off the list before any of it is queued (a queued link may
        // be reused), with the chunk's ring-slot checks issued together and
        // one count update publishing it. (A unit update waited for, a tail
        // draw and a slot check per link made the walk four round trips per
        // link: 12 ms of the barrier demo at 512 links per barrier. A unit
        // update and tail draw per chunk of four cost more than they saved.)
        // A chunk's ids and slot checks are held in registers: four per chunk
        // keep the barrier demo's kernel at 66 VGPRs (7 waves per SIMD), where
        // 8 and 16 take 82 and 122.
        unsigned int n = 0;
        for (unsigned int L = head; L < _dev_ctx.links && n < _dev_ctx.links; ++n) {
            L = hz::wave::atomic_load(&_dev_ctx.link_aux[2u * L]);
        }
        if (n == 0u) return;
        add_units((int)n);
        // The units, and the links' slots and metadata (written through by
        // the lanes that filled them), are complete before any link is queued.
        stores_done();
        constexpr unsigned int CHUNK = 4;
        hz::LinkQueue* q = work_queue(_grp_ctx.home);
        unsigned int* slots = work_slots(_grp_ctx.home);
        const unsigned int t0 = hz::wave::atomic_add(&q->tail, n);
        unsigned int h = head;
        for (unsigned int i = 0; i < n; i += CHUNK) {
            const unsigned int k = (n - i < CHUNK) ? n - i : CHUNK;
            unsigned int ids[CHUNK];
            unsigned int seen[CHUNK];
#pragma unroll
            for (unsigned int j = 0; j < CHUNK; ++j) {
                if (j < k) {
                    ids[j] = h;
                    h = hz::wave::atomic_load(&_dev_ctx.link_aux[2u * h]);
                }
            }
#pragma unroll
            for (unsigned int j = 0; j < CHUNK; ++j) {
                if (j < k) seen[j] = hz::wave::atomic_load(slots + ((t0 + i + j) & _dev_ctx.queue_mask));
            }
#pragma unroll
            for (unsigned int j = 0; j < CHUNK; ++j) {
                if (j < k) {
                    unsigned int* slot = slots + ((t0 + i + j) & _dev_ctx.queue_mask);
                    uint32_t spins = 0;
                    while (seen[j] != hz::LINK_NONE) {
                        if (++spins > SLOT_SPIN_LIMIT) {
                            stall(hz::STALL_QUEUE_PUSH, queue_detail(q, t0 + i + j));
                            return;
                        }
                        hz::wave::sleep_short();
                        seen[j] = hz::wave::atomic_load(slot);
                    }
                    hz::wave::atomic_store(slot, ids[j]);
                }
            }
            // Per chunk, so that consumers free ring slots while a list longer
            // than the ring is still being placed.
            hz::wave::atomic_add(&q->count, (int)k);
        }
    }

   protected:
    template <unsigned int I>
    __device__ HZ_INLINE void async_enum(PromiseEnum<OpSet>& pe) {
        if constexpr (I < OPS) {
            using Op = typename hz::OpOf<I, OpSet>::type;
            if (pe.disc == I) {
                async_call_cast<Op>(0, pe.data.template cast<Op>());
            } else {
                async_enum<I + 1>(pe);
            }
        }
    }

    // Any-lane arena allocation (the wave's spare cache is single-lane).
    __device__ HZ_INLINE unsigned int arena_alloc_any() {
        unsigned int id = queue_pop(_dev_ctx.free_queue, _dev_ctx.free_slots);
        if (id != hz::LINK_NONE) return id;
        if (hz::wave::atomic_load(&_dev_ctx.status->claim) < _dev_ctx.links) {
            unsigned int claim = hz::wave::atomic_add(&_dev_ctx.status->claim, 1u);
            if (claim < _dev_ctx.links) return claim;
        }
        return queue_pop(_dev_ctx.free_queue, _dev_ctx.free_slots);
    }

    __device__ HZ_INLINE void write_arena_slot(unsigned int L, unsigned int slot, const uint32_t* words) {
        unsigned int* dst = arena_link(L);
#pragma unroll
        for (uint32_t w = 0; w < WORDS; ++w) {
            __hip_atomic_store(&dst[word_index(slot, w)], words[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    template <typename BARRIER>
    __device__ HZ_INLINE void barrier_write(BARRIER& b, unsigned int L, unsigned int slot, OpDisc d,
                                            const uint32_t* words, unsigned int block) {
        write_arena_slot(L, slot, words);
        barrier_finish(b, L, d, block);
    }

    // Counts one written slot; the lane that completes the link hands it on.
    template <typename BARRIER>
    __device__ HZ_INLINE void barrier_finish(BARRIER& b, unsigned int L, OpDisc d, unsigned int block) {
        // The slot was written through; once it has completed, the count
        // that the link's completer reads covers it.
        stores_done();
        unsigned int before = hz::wave::atomic_add(&_dev_ctx.link_aux[2u * L + 1u], 1u);
        unsigned int filled = (before & 0xFFFFu) + 1u;
        unsigned int target = before >> 16;
        if (filled == target) barrier_complete(b, L, d, target, block);
    }

    // A completed link leaves its ring entry to the block RING after it, and
    // goes onto the barrier's full list, or to the work queues once the
    // barrier is released.
    template <typename BARRIER>
    __device__ HZ_INLINE void barrier_complete(BARRIER& b, unsigned int L, OpDisc d, unsigned int count,
                                               unsigned int block) {
        hz::wave::atomic_store(&b.ring[d][block % BARRIER::RING],
                               ((unsigned long long)block << 32) | hz::BARRIER_NO_LINK);
        hz::wave::atomic_store(&_dev_ctx.link_meta[L], (unsigned int)(d | (count << 16)));
        stores_done();
        unsigned int h = hz::wave::atomic_load(&b.full_head);
        for (uint32_t spins = 0; spins < SLOT_SPIN_LIMIT; ++spins) {
            if (h == hz::BARRIER_RELEASED) {
                add_units(1);
                stores_done();
                queue_push(work_queue(_grp_ctx.home), work_slots(_grp_ctx.home), L, false);
                return;
            }
            hz::wave::atomic_store(&_dev_ctx.link_aux[2u * L], h);
            stores_done();
            unsigned int seen = atomicCAS(&b.full_head, h, L);
            if (seen == h) return;
            h = seen;
        }
        stall(hz::STALL_BARRIER_LIST, ((unsigned long long)L << 32) | h);
    }

   public:
    // The async program has no queue fill fraction (reference async.h:2900-2904).
    template <typename TYPE>
    __device__ HZ_INLINE float load_fraction() {
        return NAN;
    }

    __device__ void halt_early() { set_flags(EARLY_HALT_FLAG); }

    // Promises can be held on a RemappingBarrier (futures may forward).
    static constexpr bool REMAPPING_AWAIT = true;
    // Records a misuse of the future API; complete() reports it.
    __device__ void usage_error() { set_flags(USAGE_FLAG); }

    // One workgroup = GROUPS_PER_BLOCK independent groups (wavefronts).
    __device__ static HZ_INLINE void kernel_exec(DeviceContext ctx, DeviceState state, unsigned int cycles) {
        __shared__ GroupContext gctx[GROUPS_PER_BLOCK];
        __shared__ GroupState group_state[GROUPS_PER_BLOCK];
        const unsigned int g = (GROUPS_PER_BLOCK == 1) ? 0u : (threadIdx.x / 64u);
        ThreadContext tctx;
        ThreadState thread_state;
        ProgramType prog(ctx, gctx[g], tctx, state, group_state[g], thread_state);
        if constexpr (HAS_LAZY) {
            __shared__ uint32_t lazy_stage[GROUPS_PER_BLOCK][LAZY_WORDS][64];
            prog._lazy_stage = lazy_stage[g];
        }
        prog.exec(cycles);
    }

    // Empties the queues and status of an instance; grid-stride over the slot arrays.
    __device__ static void kernel_reset(DeviceContext ctx) {
        size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        size_t stride = (size_t)gridDim.x * blockDim.x;
        size_t cap = (size_t)ctx.queue_mask + 1u;
        const size_t rings = (size_t)ctx.queue_count * LEVELS;
        size_t nslots = cap * rings;
        for (size_t i = tid; i < nslots; i += stride) ctx.queue_slots[i] = hz::LINK_NONE;
        for (size_t i = tid; i < cap; i += stride) ctx.free_slots[i] = hz::LINK_NONE;
        for (size_t i = tid; i < rings; i += stride) {
            ctx.queues[i].tail = 0;
            ctx.queues[i].head = 0;
            ctx.queues[i].count = 0;
        }
        if (tid == 0) {
            ctx.free_queue->tail = 0;
            ctx.free_queue->head = 0;
            ctx.free_queue->count = 0;
            hz::AsyncStatus* st = ctx.status;
            st->flags = 0;
            st->stall_site = 0;
            st->stall_detail = 0;
            st->units = 0;
            st->makers = 0;
            st->hungry = 0;
            st->checkout = 0;
            st->claim = 0;
            st->promises = st->spills = st->fills = st->make_work_calls = st->launches = st->idle_polls = 0;
            st->links_run = 0;
            st->appended = 0;
            st->forwarded = 0;
            for (int i = 0; i < 8; ++i) st->phase_cycles[i] = 0;
        }
    }
};
