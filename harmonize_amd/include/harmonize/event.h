// EventProgram<Spec>: the level-synchronous program type.
//
// Semantics follow the reference's program/event.h:383-518: one kernel launch
// evaluates one generation. Each group drains the per-operation input queues
// in chunks; a group that finds no input calls the spec's make_work until
// make_work returns false or some output queue is half full; then every queue
// is flipped (outputs become the next generation's inputs). The host loops
// `exec` until a generation produces no promises.
//
// Layout and mechanism are gfx950-native:
//  - a group is one 64-lane wavefront;
//  - queues hold promises as 32-bit words in structure-of-arrays form
//    ([op][word][capacity]), so a wavefront reads or writes 64 consecutive
//    words per row: fully coalesced 256-byte accesses;
//  - an async call costs one global atomic per wavefront per call site
//    (ballot + mbcnt), not one atomic and two system fences per lane as in
//    the reference (util/iter.h:197-214);
//  - inputs are claimed in chunks of chunk*64 promises with one atomic;
//  - the flip runs in a one-group kernel launched after each generation
//    (the reference's last group checks out through a counter that every
//    group increments, which serialises thousands of atomics per launch).
#pragma once

#include "host.h"
#include "iter.h"
#include "spec.h"

template <typename PROGRAM_SPEC>
class EventProgram {
   public:
    typedef EventProgram<PROGRAM_SPEC> ProgramType;

    using AdrType = typename hz::spec::AdrTypeOr<PROGRAM_SPEC, unsigned int>::type;
    using OpSet = typename hz::spec::OpSetOr<PROGRAM_SPEC, OpUnion<>>::type;
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

    static const unsigned int BAD_FUNC_ID_FLAG = 0x00000001;
    static const unsigned int OVERFLOW_FLAG = 0x00000004;
    static const unsigned int EARLY_HALT_FLAG = 0x40000000;
    static const unsigned int COMPLETION_FLAG = 0x80000000;

    static const unsigned char FN_ID_COUNT = PromiseUnionType::Info::COUNT;
    static constexpr uint32_t OPS = PromiseUnionType::Info::COUNT;
    static constexpr uint32_t WORDS = hz::PromiseWords<OpSet>::value;

    static_assert(std::is_trivially_copyable<PromiseUnionType>::value,
                  "promise arguments must be trivially copyable");

    struct ThreadContext {
        unsigned int thread_id;
        unsigned int rand_state;
    };

    struct GroupContext {
        unsigned int toggle;
    };

    // Queue counters and run status, kept in HBM.
    struct Status {
        unsigned int in_count[OPS > 0 ? OPS : 1];
        unsigned int in_claim[OPS > 0 ? OPS : 1];
        unsigned int out_count[OPS > 0 ? OPS : 1];
        unsigned int toggle;
        unsigned int checkout;
        unsigned int flags;
        unsigned int flip_count;
        unsigned long long int promises;
        unsigned long long int make_work_calls;
    };

    struct DeviceContext {
        typedef ProgramType ParentProgramType;
        Status* status;
        unsigned int* queues[2];  // each [OPS][WORDS][capacity]
        unsigned int capacity;
        unsigned int group_count;
    };

    // Owns the queues and status of one program instance (reference
    // program/event.h:139-230). `io_size` is the per-operation capacity of
    // each queue, in promises.
    struct Instance {
        util::host::DevBuf<Status> status;
        util::host::DevBuf<unsigned int> queue_a;
        util::host::DevBuf<unsigned int> queue_b;
        DeviceState device_state;
        unsigned int capacity;
        Status host_status{};
        unsigned int last_groups = 0;

        Instance(size_t io_size, DeviceState gs)
            : status((size_t)1),
              queue_a((size_t)OPS * WORDS * (size_t)io_size),
              queue_b((size_t)OPS * WORDS * (size_t)io_size),
              device_state(gs),
              capacity((unsigned int)io_size) {
            if (io_size == 0 || io_size > 0x7FFFFFFFu) {
                throw std::runtime_error("EventProgram::Instance: io_size must be in [1, 2^31)");
            }
            util::host::auto_throw(hipMemset(status.data(), 0, sizeof(Status)));
        }

        DeviceContext to_context() {
            DeviceContext result;
            result.status = status.data();
            result.queues[0] = queue_a.data();
            result.queues[1] = queue_b.data();
            result.capacity = capacity;
            result.group_count = last_groups;
            return result;
        }

        Status read_status() {
            util::host::auto_throw(hipMemcpy(&host_status, status.data(), sizeof(Status), hipMemcpyDeviceToHost));
            return host_status;
        }

        // True once a generation has produced no promises. Raises if any
        // output queue overflowed (promises would have been lost).
        bool complete() {
            read_status();
            if (host_status.flags & OVERFLOW_FLAG) {
                throw std::runtime_error(
                    "EventProgram: an output queue overflowed; construct the Instance with a larger io_size");
            }
            return (host_status.flags & COMPLETION_FLAG) != 0;
        }

        void clear_flags() {
            read_status();
            host_status.flags = 0;
            util::host::auto_throw(hipMemcpy(&status.data()->flags, &host_status.flags, sizeof(unsigned int),
                                             hipMemcpyHostToDevice));
        }

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

    __device__ void set_flags(unsigned int bits) { hz::wave::atomic_or(&_dev_ctx.status->flags, bits); }

    // Appends one promise per active lane to the output queue of TYPE with a
    // single atomic for the wavefront.
    template <typename TYPE>
    __device__ void async_call_cast(int /*depth_delta*/, Promise<TYPE> param_value) {
        constexpr OpDisc D = Lookup<TYPE>::type::DISC;
        static_assert(Lookup<TYPE>::type::CONTAINED, "operation is not part of this program's OpSet");
        PromiseUnionType pu;
        pu.template cast<TYPE>() = param_value;
        uint32_t words[WORDS];
        __builtin_memcpy(words, &pu, sizeof(PromiseUnionType) < sizeof(words) ? sizeof(PromiseUnionType)
                                                                                : sizeof(words));
        uint64_t act = hz::wave::active();
        uint32_t r = hz::wave::rank(act);
        uint32_t base = 0;
        if (r == 0) base = hz::wave::atomic_add(&_dev_ctx.status->out_count[D], hz::wave::popc(act));
        base = hz::wave::first(base);
        uint32_t idx = base + r;
        if (idx < _dev_ctx.capacity) {
            unsigned int out = _grp_ctx.toggle ^ 1u;
#pragma unroll
            for (uint32_t w = 0; w < WORDS; ++w) queue_row(out, D, w)[idx] = words[w];
        } else {
            set_flags(OVERFLOW_FLAG);
        }
    }

    template <typename TYPE>
    __device__ void immediate_call_cast(Promise<TYPE> param_value) {
        param_value(*this);
    }

    __device__ void eval_words(OpDisc op, const uint32_t* words) {
        PromiseUnionType pu;
        __builtin_memcpy(&pu, words, sizeof(PromiseUnionType));
        pu.template loose_eval<ProgramType>(*this, op);
    }

   public:
    __device__ void init_program() {}

    // One generation (reference program/event.h:383-518).
    __device__ void exec(unsigned int chunk_size) {
        const uint32_t lane = hz::wave::lane();
        _thd_ctx.thread_id = blockIdx.x * 64u + lane;
        _thd_ctx.rand_state = _thd_ctx.thread_id;
        if (lane == 0) _grp_ctx.toggle = hz::wave::atomic_load(&_dev_ctx.status->toggle);
        hz::wave::sync();
        const unsigned int in = hz::wave::first(_grp_ctx.toggle);
        const unsigned int span = (chunk_size == 0 ? 1u : chunk_size) * 64u;
        Status* st = _dev_ctx.status;
        unsigned long long made = 0;

        PROGRAM_SPEC::initialize(*this);

        while (true) {
            // Claim a chunk of input from the first operation that has any.
            uint32_t op = OPS, base = 0, end = 0;
            for (uint32_t i = 0; i < OPS; ++i) {
                uint32_t limit = hz::wave::first(hz::wave::atomic_load(&st->in_count[i]));
                if (hz::wave::first(hz::wave::atomic_load(&st->in_claim[i])) >= limit) continue;
                uint32_t b = 0;
                if (lane == 0) b = hz::wave::atomic_add(&st->in_claim[i], span);
                b = hz::wave::first(b);
                if (b < limit) {
                    op = i;
                    base = b;
                    end = (limit - b > span) ? b + span : limit;
                    break;
                }
            }
            if (op == OPS) {
                // No input left for this group: make work until the spec says
                // it has none or an output queue is half full.
                bool more = (hz::wave::atomic_load(&st->flags) & EARLY_HALT_FLAG) == 0;
                while (more) {
                    for (uint32_t i = 0; i < OPS; ++i) {
                        if (hz::wave::atomic_load(&st->out_count[i]) >= _dev_ctx.capacity / 2) more = false;
                    }
                    if (!more) break;
                    bool mine = PROGRAM_SPEC::make_work(*this);
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

        PROGRAM_SPEC::finalize(*this);

        // The queues are flipped by flip_kernel after this launch: the kernel
        // boundary orders every group's pushes before the flip, so no group
        // has to check out through a contended counter.
        if (lane == 0 && made) hz::wave::atomic_add(&st->make_work_calls, made);
    }

    // Flips every queue for the next generation (reference event.h:500-515,
    // util/iter.h:437-445) and records completion when nothing was produced.
    // Launched as one 64-lane group after each generation.
    __device__ static void flip(DeviceContext ctx) {
        if (hz::wave::lane() != 0) return;
        Status* st = ctx.status;
        unsigned long long executed = 0, produced_total = 0;
        for (uint32_t i = 0; i < OPS; ++i) {
            executed += st->in_count[i];
            unsigned int produced = st->out_count[i];
            unsigned int n = produced < ctx.capacity ? produced : ctx.capacity;
            st->in_count[i] = n;
            st->in_claim[i] = 0;
            st->out_count[i] = 0;
            produced_total += n;
        }
        st->promises += executed;
        st->toggle ^= 1u;
        st->flip_count += 1;
        unsigned int flags = st->flags;
        if (produced_total == 0 && (flags & EARLY_HALT_FLAG) == 0) flags |= COMPLETION_FLAG;
        st->flags = flags & ~EARLY_HALT_FLAG;
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

    // Fill fraction of TYPE's output queue (reference event.h:540-544).
    template <typename TYPE>
    __device__ float load_fraction() {
        unsigned int n = hz::wave::atomic_load(&_dev_ctx.status->out_count[Lookup<TYPE>::type::DISC]);
        return (float)n / (float)_dev_ctx.capacity;
    }

    __device__ void halt_early() { set_flags(EARLY_HALT_FLAG); }

    // Kernel bodies used by the launch helpers.
    __device__ static void kernel_exec(DeviceContext ctx, DeviceState state, unsigned int chunk) {
        __shared__ GroupContext gctx;
        __shared__ GroupState group_state;
        ThreadContext tctx;
        ThreadState thread_state;
        ProgramType prog(ctx, gctx, tctx, state, group_state, thread_state);
        prog.exec(chunk);
    }
};
