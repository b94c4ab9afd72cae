// ORACLE — test infrastructure only (see host_shim.h).
//
// HostProgram<Spec>: a sequential host executor of a program specification,
// with the same device-side API as the GPU program types (prog.device,
// prog.group, prog.thread, async<Op>, sync<Op>, load_fraction<Op>,
// halt_early), so an unmodified specification runs on the CPU. It restates
// the execution contract of the reference's program types:
//  - initialize once, then evaluate queued promises; when none is queued,
//    call make_work until it returns false (program/async.h:2201-2222,
//    program/event.h:445-472); then finalize (async.h:2656, event.h:493);
//  - an operation is dispatched by its discriminant, the position of its
//    type in OpSet (async/promise.h:65-91, :647-658);
//  - every promise is evaluated exactly once; sync<Op> evaluates in place
//    (program/event.h:361-367);
//  - a promise's depth level is its creator's level plus the call's depth
//    delta, clamped to the specification's STACK_SIZE levels (make_work's
//    promises are at level 0), and prog.depth() reports it
//    (program/async.h:1395-1411, mem/work.h:133-164).
// Promises are evaluated first-in first-out, one at a time (the order does
// not change a deterministic specification's results).
//
// OmpProgram<Spec>: the same contract on all host threads. Each thread has
// a private queue and its own group/thread state; surplus promises move to
// a shared pool that idle threads draw from; the run ends when every
// thread is idle, the pool is empty and no thread can make more work.
#pragma once

#include <omp.h>

#include <atomic>
#include <cmath>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "host_shim.h"
#include "harmonize/promise.h"
#include "harmonize/iter.h"
#include "harmonize/spec.h"
#include "harmonize/barrier.h"

namespace oracle {

struct RunStats {
    uint64_t promises = 0;
    uint64_t make_work_calls = 0;
    uint64_t peak_queue = 0;
    uint64_t halts = 0;
};

template <typename SPEC>
class HostProgram {
   public:
    typedef HostProgram<SPEC> ProgramType;
    using OpSet = hz::spec::ProgramOpSet<SPEC>;
    using DeviceState = typename hz::spec::DeviceStateOr<SPEC, VoidState>::type;
    using GroupState = typename hz::spec::GroupStateOr<SPEC, VoidState>::type;
    using ThreadState = typename hz::spec::ThreadStateOr<SPEC, VoidState>::type;
    typedef PromiseUnion<OpSet> PromiseUnionType;
    // A queued promise: discriminant, payload and depth level.
    struct Entry {
        OpDisc first;
        PromiseUnionType second;
        uint32_t level;
        Entry(OpDisc d, const PromiseUnionType& p, uint32_t lv = 0) : first(d), second(p), level(lv) {}
    };
    static constexpr uint32_t LEVELS =
        hz::spec::STACK_SIZEOr<SPEC, 0>::value == 0 ? 1u : (uint32_t)hz::spec::STACK_SIZEOr<SPEC, 0>::value;

    template <typename TYPE>
    struct Lookup {
        typedef OpUnionLookup<TYPE, OpSet> type;
    };

    static constexpr bool PERSISTENT = false;

    struct Runner {
        std::deque<Entry> queue;
        RunStats stats;
        uint32_t level = 0;  // level of the promise being evaluated
    };

    DeviceState& device;
    GroupState& group;
    ThreadState& thread;
    Runner& _run;

    HostProgram(DeviceState& d, GroupState& g, ThreadState& t, Runner& r)
        : device(d), group(g), thread(t), _run(r) {}

    template <typename TYPE, typename... ARGS>
    void async(ARGS... args) {
        async_call<TYPE>(Promise<TYPE>(args...));
    }

    template <typename TYPE>
    void async_call(Promise<TYPE> promise) {
        async_call_cast<TYPE>(0, promise);
    }

    template <typename TYPE>
    void async_call_cast(int depth_delta, Promise<TYPE> promise) {
        static_assert(Lookup<TYPE>::type::CONTAINED, "operation is not part of this program's OpSet");
        PromiseUnionType pu;
        pu.template cast<TYPE>() = promise;
        int lv = (int)_run.level + depth_delta;
        lv = lv < 0 ? 0 : (lv >= (int)LEVELS ? (int)LEVELS - 1 : lv);
        _run.queue.emplace_back(Lookup<TYPE>::type::DISC, pu, (uint32_t)lv);
        if (_run.queue.size() > _run.stats.peak_queue) _run.stats.peak_queue = _run.queue.size();
    }

    unsigned int depth() const { return _run.level; }

    template <typename TYPE, typename... ARGS>
    void sync(ARGS... args) {
        Promise<TYPE>(args...)(*this);
    }

    template <typename TYPE>
    void sync_call(Promise<TYPE> promise) {
        promise(*this);
    }

    // A host executor has no bounded output queue; report it empty.
    template <typename TYPE>
    float load_fraction() {
        return 0.0f;
    }

    // A GPU pass halted early is resumed by the host loop; the sequential
    // executor therefore just records the request and carries on.
    void halt_early() { _run.stats.halts++; }

    // Promises can be held on a RemappingBarrier (futures may forward).
    static constexpr bool REMAPPING_AWAIT = true;
    void usage_error() { throw std::logic_error("HostProgram: a Return forwarded a unit future that was not ready"); }

    void async_call(PromiseEnum<OpSet> pe) {
        _run.queue.emplace_back(pe.disc, pe.data);
        if (_run.queue.size() > _run.stats.peak_queue) _run.stats.peak_queue = _run.queue.size();
    }

    // RemappingBarrier (reference sync/barrier.h:460-639): promises awaited
    // on an unreleased barrier wait in a side table shared by all host
    // threads; release moves them to the releasing thread's queue. The
    // semaphore test and the insertion happen under the table's lock, and a
    // release zeroes the semaphore before it takes the lock, so no promise
    // can be inserted after its barrier was drained.
    static std::mutex& barrier_lock() {
        static std::mutex m;
        return m;
    }
    static std::unordered_map<const void*, std::vector<Entry>>& barrier_table() {
        static std::unordered_map<const void*, std::vector<Entry>> t;
        return t;
    }

    template <typename BARRIER, typename OP_TYPE>
    void barrier_await(BARRIER& b, Promise<OP_TYPE> promise) {
        {
            std::lock_guard<std::mutex> guard(barrier_lock());
            if (__atomic_load_n(&b.semaphore.sem, __ATOMIC_SEQ_CST) != 0u) {
                PromiseUnionType pu;
                pu.template cast<OP_TYPE>() = promise;
                barrier_table()[&b].emplace_back(Lookup<OP_TYPE>::type::DISC, pu);
                return;
            }
        }
        async_call<OP_TYPE>(promise);
    }

    template <typename BARRIER>
    void barrier_release(BARRIER& b) {
        std::vector<Entry> held;
        {
            std::lock_guard<std::mutex> guard(barrier_lock());
            auto it = barrier_table().find(&b);
            if (it != barrier_table().end()) {
                held.swap(it->second);
                barrier_table().erase(it);
            }
        }
        for (auto& e : held) _run.queue.push_back(e);
        if (_run.queue.size() > _run.stats.peak_queue) _run.stats.peak_queue = _run.queue.size();
    }

    // Evaluates one queued promise; false if the queue is empty.
    bool step() {
        if (_run.queue.empty()) return false;
        Entry e = _run.queue.front();
        _run.queue.pop_front();
        _run.level = e.level;
        e.second.template loose_eval<ProgramType>(*this, e.first);
        _run.stats.promises++;
        return true;
    }

    // Runs a specification to completion on one host thread.
    static RunStats run(DeviceState state) {
        hz_host_group_slot() = 0;
        Runner r;
        GroupState g{};
        ThreadState t{};
        ProgramType prog(state, g, t, r);
        SPEC::initialize(prog);
        bool can_make = true;
        while (true) {
            if (prog.step()) continue;
            if (can_make) {
                r.stats.make_work_calls++;
                r.level = 0;
                can_make = SPEC::make_work(prog);
                continue;
            }
            break;
        }
        SPEC::finalize(prog);
        return r.stats;
    }
};

template <typename SPEC>
struct OmpProgram {
    using Host = HostProgram<SPEC>;
    using DeviceState = typename Host::DeviceState;
    using GroupState = typename Host::GroupState;
    using ThreadState = typename Host::ThreadState;
    using Entry = typename Host::Entry;

    static constexpr size_t SHARE_ABOVE = 4096;  // local queue length that triggers sharing
    static constexpr size_t TAKE = 1024;         // promises taken from the pool at once

    static RunStats run(DeviceState state, int threads) {
        if (threads <= 0) threads = omp_get_max_threads();
        std::mutex lock;
        std::vector<Entry> pool;
        int busy = threads;
        RunStats total;

#pragma omp parallel num_threads(threads)
        {
            hz_host_group_slot() = omp_get_thread_num();
            typename Host::Runner r;
            GroupState g{};
            ThreadState t{};
            DeviceState local_state = state;
            Host prog(local_state, g, t, r);
            SPEC::initialize(prog);
            bool can_make = true;
            while (true) {
                if (prog.step()) {
                    if (r.queue.size() > SHARE_ABOVE) {
                        std::lock_guard<std::mutex> guard(lock);
                        size_t n = r.queue.size() / 2;
                        for (size_t i = 0; i < n; ++i) {
                            pool.push_back(r.queue.back());
                            r.queue.pop_back();
                        }
                    }
                    continue;
                }
                if (can_make) {
                    r.stats.make_work_calls++;
                    r.level = 0;
                    can_make = SPEC::make_work(prog);
                    continue;
                }
                // Idle: draw from the pool, or finish once nobody can produce work.
                bool done = false;
                bool idle_marked = false;
                while (true) {
                    {
                        std::lock_guard<std::mutex> guard(lock);
                        if (!pool.empty()) {
                            size_t n = pool.size() < TAKE ? pool.size() : TAKE;
                            for (size_t i = 0; i < n; ++i) {
                                r.queue.push_back(pool.back());
                                pool.pop_back();
                            }
                            if (idle_marked) busy++;
                            break;
                        }
                        if (!idle_marked) {
                            busy--;
                            idle_marked = true;
                        }
                        if (busy == 0) {
                            done = true;
                            break;
                        }
                    }
                    std::this_thread::yield();
                }
                if (done) break;
            }
            SPEC::finalize(prog);
            std::lock_guard<std::mutex> guard(lock);
            total.promises += r.stats.promises;
            total.make_work_calls += r.stats.make_work_calls;
            total.halts += r.stats.halts;
            if (r.stats.peak_queue > total.peak_queue) total.peak_queue = r.stats.peak_queue;
        }
        return total;
    }

};

}  // namespace oracle
