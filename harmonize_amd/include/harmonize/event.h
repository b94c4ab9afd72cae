// EventProgram<Spec>: the level-synchronous program type.
//
// Semantics follow the reference's program/event.h:383-518: one launch
// evaluates one generation. Each group drains the per-operation input queues
// in chunks; a group that finds no input calls the spec's make_work until
// make_work returns false or an output queue is half full; then every queue
// is flipped (outputs become the next generation's inputs). The host loops
// `exec` until a generation produces no promises.
//
// Layout and mechanism are gfx950-native:
//  - a group is one 64-lane wavefront;
//  - queues hold promises as 32-bit words in structure-of-arrays form
//    ([op][word][capacity]), so a wavefront reads or writes 64 consecutive
//    words per row: fully coalesced 256-byte accesses;
//  - each operation's queue is split into SEGMENTS equal segments, each with
//    its own counters on its own 128-byte line. A group pushes to, and first
//    drains, its home segment (blockIdx % SEGMENTS) and moves on to other
//    segments only when its own is full or empty. Device-scope atomics on one
//    address serialise across the XCDs at a few tens of ns each, so a single
//    counter per queue would cap a generation at tens of millions of pushes
//    per second;
//  - an async call costs one atomic per wavefront per call site (ballot +
//    mbcnt), not one atomic and two system fences per lane as in the
//    reference (util/iter.h:197-214); inputs are claimed chunk*64 at a time;
//  - the flip runs in a one-group kernel after each generation (the
//    reference's last group checks out through a counter that every group
//    increments). It also publishes a bitmask of the non-empty input
//    segments, so sparse generations do not scan every counter.
#pragma once

#include "host.h"
#include "iter.h"
#include "spec.h"

template <typename PROGRAM_SPEC>
class EventProgram {
   public:
    typedef EventProgram<PROGRAM_SPEC> ProgramType;

    using AdrType = typename hz::spec::AdrTypeOr<PROGRAM_SPEC, unsigned int>::type;
    using OpSet = hz::spec::ProgramOpSet<PROGRAM_SPEC>;
    using DeviceState = typename hz::spec::DeviceStateOr<PROGRAM_SPEC, VoidState>::type;
    using GroupState = typename hz::spec::GroupStateOr<PROGRAM_SPEC, VoidState>::type;
    using ThreadState = typename hz::spec::ThreadStateOr<PROGRAM_SPEC, VoidState>::type;

    typedef PromiseUnion<OpSet> PromiseUnionType;

    template <typename TYPE>
    struct Lookup {
        typedef typename PromiseUnionType::template Lookup<TYPE>::type type;
    };

    static_assert(!hz::spec::GROUP_SIZEOr<PROGRAM_SPEC, 64>::given ||
                      hz::spec::GROUP_SIZEOr<PROGRAM_SPEC, 64>::value == 64,
                  "On gfx950 a group is one 64-lane wavefront; GROUP_SIZE must be 64 if given.");

    static const size_t GROUP_SIZE = 64;
    static const size_t WORK_GROUP_SIZE = 64;
    // Groups live for one generation only.
    static constexpr bool PERSISTENT = false;

    static const unsigned int BAD_FUNC_ID_FLAG = 0x00000001;
    static const unsigned int OVERFLOW_FLAG = 0x00000004;
    static const unsigned int USAGE_FLAG = 0x00000020;
    static const unsigned int EARLY_HALT_FLAG = 0x40000000;
    static const unsigned int COMPLETION_FLAG = 0x80000000;

    static const unsigned char FN_ID_COUNT = PromiseUnionType::Info::COUNT;
    static constexpr uint32_t OPS = PromiseUnionType::Info::COUNT;
    static constexpr uint32_t OPS1 = OPS > 0 ? OPS : 1;
    static constexpr uint32_t WORDS = hz::PromiseWords<OpSet>::value;
    static constexpr uint32_t SEGMENTS = 64;

    static_assert(std::is_trivially_copyable<PromiseUnionType>::value,
                  "promise arguments must be trivially copyable");

    struct ThreadContext {
        unsigned int thread_id;
        unsigned int rand_state;
    };

    struct GroupContext {
        unsigned int toggle;
        unsigned long long int nonempty[OPS1];
    };

    // Counters of one segment of one operation's queue (one 128-byte line).
    struct Segment {
        unsigned int in_count;   // promises in this segment's input
        unsigned int in_claim;   // input claimed so far
        unsigned int out_count;  // output slots claimed (may pass the capacity)
        unsigned int pad[29];
    };

    // Queue counters and run status, kept in HBM.
    struct Status {
        Segment seg[OPS1][SEGMENTS];
        unsigned long long int nonempty[OPS1];  // bit s: input segment s is non-empty
        unsigned int toggle;
        unsigned int flags;
        unsigned int flip_count;
        unsigned int pad;
        unsigned long long int promises;
        unsigned long long int make_work_calls;
    };

    struct DeviceContext {
        typedef ProgramType ParentProgramType;
        Status* status;
        unsigned int* queues[2];  // each [OPS][WORDS][capacity]
        unsigned int capacity;    // per operation, SEGMENTS * seg_capacity
        unsigned int seg_capacity;
    };

    // Owns the queues and status of one program instance (reference
    // program/event.h:139-230). `io_size` is the per-operation capacity of
    // each queue, in promises (rounded up to a multiple of SEGMENTS).
    struct Instance {
        unsigned int seg_capacity;
        unsigned int capacity;
        util::host::DevBuf<Status> status;
        util::host::DevBuf<unsigned int> queue_a;
        util::host::DevBuf<unsigned int> queue_b;
        DeviceState device_state;
        Status host_status{};

        static unsigned int seg_cap_for(size_t io_size) {
            if (io_size == 0 || io_size > 0x7FFFFFFFu) {
                throw std::runtime_error("EventProgram::Instance: io_size must be in [1, 2^31)");
            }
            return (unsigned int)((io_size + SEGMENTS - 1) / SEGMENTS);
        }

        Instance(size_t io_size, DeviceState gs)
            : seg_capacity(seg_cap_for(io_size)),
              capacity(seg_capacity * SEGMENTS),
              status((size_t)1),
              queue_a((size_t)OPS1 * WORDS * (size_t)seg_capacity * SEGMENTS),
              queue_b((size_t)OPS1 * WORDS * (size_t)seg_capacity * SEGMENTS),
              device_state(gs) {
            util::host::auto_throw(hipMemset(status.data(), 0, sizeof(Status)));
        }

        DeviceContext to_context() {
            DeviceContext result;
            result.status = status.data();
            result.queues[0] = queue_a.data();
            result.queues[1] = queue_b.data();
            result.capacity = capacity;
            result.seg_capacity = seg_capacity;
            return result;
        }

        Status read_status() {
            util::host::auto_throw(hipMemcpy(&host_status, status.data(), sizeof(Status), hipMemcpyDeviceToHost));
            return host_status;
        }

        // Reads only the flags word (the per-generation completion check).
        unsigned int read_flags() {
            unsigned int flags = 0;
            util::host::auto_throw(
                hipMemcpy(&flags, &status.data()->flags, sizeof(flags), hipMemcpyDeviceToHost));
            host_status.flags = flags;
            return flags;
        }

        // True once a generation has produced no promises. Raises if an
        // output queue overflowed (promises would have been lost).
        bool complete() {
            unsigned int flags = read_flags();
            if (flags & OVERFLOW_FLAG) {
                throw std::runtime_error(
                    "EventProgram: an output queue overflowed; construct the Instance with a larger io_size");
            }
            if (flags & USAGE_FLAG) {
                throw std::runtime_error(
                    "EventProgram: a Return forwarded a future that was not ready (EventProgram cannot hold "
                    "promises on a barrier)");
            }
            return (flags & COMPLETION_FLAG) != 0;
        }

        void clear_flags() {
            unsigned int zero = 0;
            util::host::auto_throw(
                hipMemcpy(&status.data()->flags, &zero, sizeof(unsigned int), hipMemcpyHostToDevice));
        }

        // Clears every queue counter, the flags and the statistics with a
        // kernel on `stream`, so that the instance can run a new job without
        // a synchronous copy (queued promises are discarded).
        void reset(hipStream_t stream);

        unsigned long long int promises() { return read_status().promises; }
    };

    DeviceContext& _dev_ctx;
    GroupContext& _grp_ctx;
    ThreadContext& _thd_ctx;

    DeviceState& device;
    GroupState& group;
    ThreadState& thread;

    __device__ EventProgram(DeviceContext& d_c, GroupContext& g_c, ThreadContext& t_c, DeviceState& d,
                            GroupState& g, ThreadState& t)
        : _dev_ctx(d_c), _grp_ctx(g_c), _thd_ctx(t_c), device(d), group(g), thread(t) {}

   protected:
    __device__ unsigned int* queue_row(unsigned int which, OpDisc op, unsigned int word) {
        return _dev_ctx.queues[which] + ((size_t)op * WORDS + word) * _dev_ctx.capacity;
    }

    __device__ Segment& segment(OpDisc op, uint32_t s) { return _dev_ctx.status->seg[op][s]; }

    __device__ uint32_t home_segment() const { return blockIdx.x % SEGMENTS; }

    __device__ void set_flags(unsigned int bits) { hz::wave::atomic_or(&_dev_ctx.status->flags, bits); }

   public:
    // Appends one promise per active lane to the output queue of TYPE: one
    // atomic per wavefront on the home segment, spilling to the following
    // segments only when it is full. The depth delta is accepted and ignored:
    // a generation runs every queued promise, whatever its depth (as in the
    // reference, program/event.h:326-358).
    template <typename TYPE>
    __device__ void async_call_cast(int /*depth_delta*/, Promise<TYPE> param_value) {
        constexpr OpDisc D = Lookup<TYPE>::type::DISC;
        static_assert(Lookup<TYPE>::type::CONTAINED, "operation is not part of this program's OpSet");
        PromiseUnionType pu;
        pu.template cast<TYPE>() = param_value;
        uint32_t words[WORDS];
        __builtin_memcpy(words, &pu, sizeof(PromiseUnionType));
        const unsigned int out = _grp_ctx.toggle ^ 1u;
        const uint32_t cap = _dev_ctx.seg_capacity;
        bool pending = true;
        uint32_t s = home_segment();
        for (uint32_t attempt = 0; attempt < SEGMENTS; ++attempt) {
            uint64_t want = hz::wave::ballot(pending);
            if (want == 0) break;
            uint32_t r = hz::wave::rank(want);
            uint32_t base = 0;
            if (pending && r == 0) base = hz::wave::atomic_add(&segment(D, s).out_count, hz::wave::popc(want));
            base = hz::wave::first(hz::wave::read_lane(base, (uint32_t)__builtin_ctzll(want)));
            if (pending) {
                uint32_t idx = base + r;
                if (idx < cap) {
                    size_t at = (size_t)s * cap + idx;
#pragma unroll
                    for (uint32_t w = 0; w < WORDS; ++w) queue_row(out, D, w)[at] = words[w];
                    pending = false;
                }
            }
            s = (s + 1 == SEGMENTS) ? 0 : s + 1;
        }
        if (hz::wave::ballot(pending) != 0 && hz::wave::leader()) set_flags(OVERFLOW_FLAG);
    }

    // Depth level of the promise being evaluated: a level-synchronous
    // program keeps no levels, so always 0.
    __device__ unsigned int depth() { return 0u; }

   protected:
    template <typename TYPE>
    __device__ void immediate_call_cast(Promise<TYPE> param_value) {
        param_value(*this);
    }

    __device__ void eval_words(OpDisc op, const uint32_t* words) {
        PromiseUnionType pu;
        __builtin_memcpy(&pu, words, sizeof(PromiseUnionType));
        pu.template loose_eval<ProgramType>(*this, op);
    }

    // Claims up to `span` input promises of operation `op` in segment `s`.
    __device__ bool claim(uint32_t op, uint32_t s, uint32_t span, uint32_t& base, uint32_t& end) {
        Segment& g = segment(op, s);
        uint32_t limit = hz::wave::first(hz::wave::atomic_load(&g.in_count));
        if (hz::wave::first(hz::wave::atomic_load(&g.in_claim)) >= limit) return false;
        uint32_t b = 0;
        if (hz::wave::lane() == 0) b = hz::wave::atomic_add(&g.in_claim, span);
        b = hz::wave::first(b);
        if (b >= limit) return false;
        base = s * _dev_ctx.seg_capacity + b;
        end = s * _dev_ctx.seg_capacity + ((limit - b > span) ? b + span : limit);
        return true;
    }

   public:
    __device__ void init_program() {}

    // One generation (reference program/event.h:383-518).
    __device__ void exec(unsigned int chunk_size) {
        const uint32_t lane = hz::wave::lane();
        _thd_ctx.thread_id = blockIdx.x * 64u + lane;
        _thd_ctx.rand_state = _thd_ctx.thread_id;
        Status* st = _dev_ctx.status;
        if (lane == 0) {
            _grp_ctx.toggle = hz::wave::atomic_load(&st->toggle);
            for (uint32_t i = 0; i < OPS; ++i) _grp_ctx.nonempty[i] = hz::wave::atomic_load(&st->nonempty[i]);
        }
        hz::wave::sync();
        const unsigned int in = hz::wave::first(_grp_ctx.toggle);
        const unsigned int span = (chunk_size == 0 ? 1u : chunk_size) * 64u;
        const uint32_t home = home_segment();
        unsigned long long made = 0;

        // The spec hooks are inlined: a hook left out of line would take the
        // program object and push the state it refers to into scratch memory.
        [[clang::always_inline]] PROGRAM_SPEC::initialize(*this);

        while (true) {
            // Claim a chunk: home segment first, then the other non-empty ones.
            uint32_t op = OPS, base = 0, end = 0;
            for (uint32_t i = 0; i < OPS && op == OPS; ++i) {
                uint64_t live = hz::wave::first((uint64_t)_grp_ctx.nonempty[i]);
                if (live == 0) continue;
                if ((live >> home) & 1ull) {
                    if (claim(i, home, span, base, end)) {
                        op = i;
                        break;
                    }
                    live &= ~(1ull << home);
                }
                while (live != 0) {
                    uint32_t s = (uint32_t)__builtin_ctzll(live);
                    if (claim(i, s, span, base, end)) {
                        op = i;
                        break;
                    }
                    live &= live - 1;
                }
                if (op == OPS && hz::wave::lane() == 0) _grp_ctx.nonempty[i] = 0;
                hz::wave::sync();
            }
            if (op == OPS) {
                // No input left for this group: make work until the spec says
                // it has none or this group's output segment is half full.
                bool more = (hz::wave::first(hz::wave::atomic_load(&st->flags)) & EARLY_HALT_FLAG) == 0;
                while (more) {
                    for (uint32_t i = 0; i < OPS; ++i) {
                        uint32_t n = hz::wave::first(hz::wave::atomic_load(&segment(i, home).out_count));
                        if (n >= _dev_ctx.seg_capacity / 2) more = false;
                    }
                    if (!more) break;
                    bool mine;
                    [[clang::always_inline]] mine = PROGRAM_SPEC::make_work(*this);
                    more = hz::wave::ballot(mine) != 0;
                    if (more) made++;
                }
                break;
            }
            for (uint32_t i = base + lane; i < end; i += 64u) {
                uint32_t words[WORDS];
#pragma unroll
                for (uint32_t w = 0; w < WORDS; ++w) words[w] = queue_row(in, op, w)[i];
                eval_words(op, words);
            }
        }

        [[clang::always_inline]] PROGRAM_SPEC::finalize(*this);
        if (lane == 0 && made) hz::wave::atomic_add(&st->make_work_calls, made);
    }

    // Flips every queue for the next generation (reference event.h:500-515,
    // util/iter.h:437-445) and records completion when nothing was produced.
    // Launched as one 64-lane group after each generation; lane s handles
    // segment s of every operation.
    static_assert(SEGMENTS == 64, "flip assigns one segment per lane");
    // Zeroes the Status struct (grid-stride over its 32-bit words).
    __device__ static void kernel_reset(DeviceContext ctx) {
        unsigned int* w = (unsigned int*)ctx.status;
        const size_t words = sizeof(Status) / sizeof(unsigned int);
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x) {
            w[i] = 0u;
        }
    }

    __device__ static void flip(DeviceContext ctx) {
        Status* st = ctx.status;
        const uint32_t lane = hz::wave::lane();
        unsigned long long executed = 0, produced = 0;
        for (uint32_t i = 0; i < OPS; ++i) {
            Segment& g = st->seg[i][lane];
            uint32_t done_here = g.in_count;
            uint32_t n = g.out_count < ctx.seg_capacity ? g.out_count : ctx.seg_capacity;
            g.in_count = n;
            g.in_claim = 0;
            g.out_count = 0;
            uint64_t live = hz::wave::ballot(n != 0);
            for (uint32_t l = 0; l < 64; ++l) {
                executed += hz::wave::read_lane(done_here, l);
                produced += hz::wave::read_lane(n, l);
            }
            if (lane == 0) st->nonempty[i] = live;
        }
        if (lane == 0) {
            st->promises += executed;
            st->toggle ^= 1u;
            st->flip_count += 1;
            unsigned int flags = st->flags;
            if (produced == 0 && (flags & EARLY_HALT_FLAG) == 0) flags |= COMPLETION_FLAG;
            st->flags = flags & ~EARLY_HALT_FLAG;
        }
    }

    template <typename TYPE, typename... ARGS>
    __device__ void async(ARGS... args) {
        async_call_cast<TYPE>(0, Promise<TYPE>(args...));
    }

    template <typename TYPE>
    __device__ void async_call(Promise<TYPE> promise) {
        async_call_cast<TYPE>(0, promise);
    }

    template <typename TYPE, typename... ARGS>
    __device__ void sync(ARGS... args) {
        immediate_call_cast<TYPE>(Promise<TYPE>(args...));
    }

    template <typename TYPE>
    __device__ void sync_call(Promise<TYPE> promise) {
        immediate_call_cast<TYPE>(promise);
    }

    // Fill fraction of TYPE's output in this group's home segment (reference
    // event.h:540-544 reports the whole queue's).
    template <typename TYPE>
    __device__ float load_fraction() {
        unsigned int n =
            hz::wave::atomic_load(&segment(Lookup<TYPE>::type::DISC, home_segment()).out_count);
        return (float)n / (float)_dev_ctx.seg_capacity;
    }

    __device__ void halt_early() { set_flags(EARLY_HALT_FLAG); }

    // Promises cannot be held on a RemappingBarrier (see barrier_await).
    static constexpr bool REMAPPING_AWAIT = false;
    // Records a misuse of the future API; complete() reports it.
    __device__ void usage_error() { set_flags(USAGE_FLAG); }

    // Queues a promise held as a PromiseEnum (UnitBarrier release).
    __device__ HZ_INLINE void async_call(PromiseEnum<OpSet> promise) { async_enum<0>(promise); }

    // Holding promises on a RemappingBarrier needs the async program's link
    // arena, as in the reference (sync/barrier.h uses AsyncProgram internals
    // throughout), so awaiting is rejected at compile time. Releasing one is
    // allowed and has nothing to run: no promise can have been held. (A
    // future backed by a RemappingBarrier still carries its value and its
    // ready state under EventProgram.)
    template <typename BARRIER, typename OP_TYPE>
    __device__ void barrier_await(BARRIER&, Promise<OP_TYPE>) {
        static_assert(sizeof(OP_TYPE) == 0, "RemappingBarrier::await is supported by AsyncProgram, not EventProgram");
    }
    template <typename BARRIER>
    __device__ void barrier_release(BARRIER&) {}

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

   public:

    __device__ static void kernel_exec(DeviceContext ctx, DeviceState state, unsigned int chunk) {
        __shared__ GroupContext gctx;
        __shared__ GroupState group_state;
        ThreadContext tctx;
        ThreadState thread_state;
        ProgramType prog(ctx, gctx, tctx, state, group_state, thread_state);
        prog.exec(chunk);
    }
};
