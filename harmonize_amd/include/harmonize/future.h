// Futures and value-returning operations (reference async/future.h:7-60,
// async/promise.h:326-434, sync/barrier.h:648-699).
//
// The reference declares these pieces but leaves their resolution unfinished
// (Return::resolve is commented out, async/promise.h:415-432). This file
// completes them with the following semantics.
//
// An operation with `using Type = R(*)(Args...)`, R not void, is
// value-returning. Its promises carry a ReturnAdr<R> `ret`: `base` is the
// TaggedSemaphore that starts a UnitBarrier or a RemappingBarrier (the tag
// says which) and `dest` is where the value goes. When a program type
// evaluates such a promise, hz::resolve delivers what eval returned:
//  - a value (an R, or a Return<P,R> built from one) is stored at `dest`,
//    then `base`'s barrier is decremented once;
//  - a promise (a Return<P,R> built from a Promise<Op> whose Op returns R)
//    takes over the return address and is queued, so that a chain of
//    promises resolves the original address once, at its end;
//  - a future (a Return<P,R> built from a future's ReturnAdr) forwards that
//    future's value: a ReturnOp promise is held on the future's
//    RemappingBarrier and, once the future resolves, copies the value to
//    `dest` and decrements `base`. Only a RemappingBarrier future can hold
//    that promise, and only AsyncProgram and the host executors can hold
//    promises on a barrier; a unit future is forwarded only if it is
//    already ready. Otherwise the program reports a usage error.
// A null `base` (the default of a promise built from its arguments) drops
// the result; a null `dest` only decrements.
//
// Every decrement is a release and the one that reaches zero an acquire
// (barrier.h), so a promise released by a future sees the future's value.
#pragma once

#include "promise.h"

namespace hz {
namespace detail {

template <typename BARRIER>
struct IsRemappingBarrier : std::false_type {};
template <typename OP_SET, typename ADR_TYPE>
struct IsRemappingBarrier<RemappingBarrier<OP_SET, ADR_TYPE>> : std::true_type {};

template <typename T>
struct IsReturn : std::false_type {};
template <typename PROGRAM, typename TYPE>
struct IsReturn<::Return<PROGRAM, TYPE>> : std::true_type {};

}  // namespace detail
}  // namespace hz

// A value tied to a barrier (reference async/future.h:7-32). With a
// RemappingBarrier, any number of promises may await the value; with a
// UnitBarrier, the future may queue one promise, fixed at construction.
// `data` is meaningful once the future is ready.
template <typename TYPE, typename BARRIER>
struct Future {
    using BarrierType = BARRIER;
    using Type = TYPE;

    BarrierType barrier;
    Type data;

    Future() = default;

    // A future that becomes ready after `semaphore_value` fulfilments
    // (reference future.h:18-20).
    explicit HZ_HD Future(unsigned int semaphore_value) : barrier(make(semaphore_value)), data() {}

    // A UnitBarrier future that queues `then` when it becomes ready.
    template <typename OP_TYPE>
    HZ_HD Future(Promise<OP_TYPE> then, unsigned int semaphore_value) : barrier(then, semaphore_value), data() {
        static_assert(!hz::detail::IsRemappingBarrier<BarrierType>::value,
                      "a RemappingBarrier future holds promises with await(), not at construction");
    }

    // The return address a value-returning promise resolves this future with.
    HZ_HD ReturnAdr<TYPE> adr() { return ReturnAdr<TYPE>{&barrier.semaphore, &data}; }

    // Holds `promise` until the future is ready; runs it at once if it is
    // already (reference future.h:22-25).
    template <typename PROGRAM, typename OP_TYPE>
    __device__ HZ_INLINE void await(PROGRAM program, Promise<OP_TYPE> promise) {
        static_assert(hz::detail::IsRemappingBarrier<BarrierType>::value,
                      "only a RemappingBarrier future can hold awaiting promises");
        barrier.await(program, promise);
    }

    // Counts one fulfilment (reference future.h:28-31).
    template <typename PROGRAM>
    __device__ HZ_INLINE void fulfill(PROGRAM program) {
        barrier.sub_semaphore(program, 1u);
    }

    // Stores the value, then counts one fulfilment.
    template <typename PROGRAM>
    __device__ HZ_INLINE void fulfill(PROGRAM program, TYPE value) {
        data = value;
        fulfill(program);
    }

    // True once every fulfilment has been counted; `data` may then be read.
    __device__ HZ_INLINE bool ready() const {
        bool r = hz_load(&barrier.semaphore.sem) == 0u;
        if (r) hz_fence_acquire();
        return r;
    }

   private:
    static HZ_HD BarrierType make(unsigned int semaphore_value) {
        if constexpr (hz::detail::IsRemappingBarrier<BarrierType>::value) {
            return BarrierType::blank(semaphore_value);
        } else {
            return BarrierType(semaphore_value);
        }
    }
};

// The result of a value-returning operation's eval (reference
// async/promise.h:350-434): a value, a promise of an operation in the
// program's set that returns the same type, or the address of a future.
template <typename PROGRAM, typename TYPE>
struct Return {
    using Type = TYPE;
    using ProgramType = PROGRAM;
    using OpSet = typename PROGRAM::OpSet;
    using EnumType = PromiseEnum<OpSet>;
    using RetAdrType = ReturnAdr<TYPE>;
    using RFutureType = Future<TYPE, RemappingBarrier<OpSet>>;
    using UFutureType = Future<TYPE, UnitBarrier<OpSet>>;

    static_assert(std::is_trivially_copyable<TYPE>::value, "a returned value must be trivially copyable");

    enum Form { VALUE, PROMISE, FUTURE };

    union Data {
        Type value;
        EnumType promise;
        RetAdrType future;
        HZ_HD Data() {}
    };

    Form form;
    Data data;

    // Rejects, at compile time, a promise whose operation is not in the
    // program's set or does not return TYPE (reference promise.h:379-396).
    template <typename OP_TYPE>
    static HZ_HD Promise<OP_TYPE> return_guard(Promise<OP_TYPE> promise) {
        static_assert(PromiseUnion<OpSet>::template Lookup<OP_TYPE>::type::CONTAINED,
                      "the returned promise's operation is not part of the program's OpSet");
        static_assert(std::is_same<typename Promise<OP_TYPE>::Return, TYPE>::value,
                      "the returned promise's operation does not return the same type as the returning operation");
        return promise;
    }

    template <typename OP_TYPE>
    HZ_HD Return(Promise<OP_TYPE> promise) : form(PROMISE) {
        data.promise = EnumType(return_guard(promise));
    }

    HZ_HD Return(Type value) : form(VALUE) { data.value = value; }

    // `future` must be the address of a RemappingBarrier future (its adr()).
    HZ_HD Return(RetAdrType future) : form(FUTURE) { data.future = future; }

    template <typename OP_SET, typename ADR_TYPE>
    HZ_HD Return(Future<TYPE, RemappingBarrier<OP_SET, ADR_TYPE>>& future) : form(FUTURE) {
        data.future = future.adr();
    }
};

namespace hz {

// Counts one fulfilment of the barrier that starts with `sem`; nothing for a
// null semaphore (a dropped result).
template <typename PROGRAM>
__device__ __host__ HZ_INLINE void fulfill(PROGRAM program, TaggedSemaphore* sem) {
    if (sem == nullptr) return;
    using Set = typename PROGRAM::OpSet;
    if (sem->tag == UNIT_BARRIER_TAG) {
        reinterpret_cast<UnitBarrier<Set>*>(sem)->sub_semaphore(program, 1u);
    } else {
        reinterpret_cast<RemappingBarrier<Set>*>(sem)->sub_semaphore(program, 1u);
    }
}

namespace detail {

// Queues the promise held in `pe` with `adr` as its return address. Each
// candidate operation is a branch on the discriminant that copies the promise
// out as its own type, so no union member is written through a run-time
// choice (which would keep the union in scratch memory).
template <unsigned int I, typename PROGRAM, typename ENUM, typename TYPE>
__device__ __host__ HZ_INLINE void forward_promise(PROGRAM program, ENUM pe, ReturnAdr<TYPE> adr) {
    using Set = typename PROGRAM::OpSet;
    if constexpr (I < OpCount<Set>::value) {
        using Op = typename OpOf<I, Set>::type;
        if constexpr (std::is_same<typename Promise<Op>::Return, TYPE>::value) {
            if (pe.disc == I) {
                Promise<Op> p = pe.data.template cast<Op>();
                p.ret = adr;
                program.async_call(p);
                return;
            }
        }
        forward_promise<I + 1>(program, pe, adr);
    }
}

}  // namespace detail

template <typename PROGRAM, typename TYPE>
__device__ __host__ HZ_INLINE void resolve_value(PROGRAM program, ReturnAdr<TYPE> adr, const TYPE& value) {
    if (adr.dest != nullptr) *adr.dest = value;
    fulfill(program, adr.base);
}

template <typename PROGRAM, typename TYPE, typename RESULT>
__device__ __host__ HZ_INLINE void resolve(PROGRAM program, ReturnAdr<TYPE> adr, RESULT result) {
    if constexpr (detail::IsReturn<RESULT>::value) {
        static_assert(std::is_same<typename RESULT::Type, TYPE>::value,
                      "eval returned a Return of another type than its operation's");
        static_assert(std::is_same<typename RESULT::ProgramType, PROGRAM>::value,
                      "eval returned a Return for another program type");
        if (result.form == RESULT::VALUE) {
            resolve_value(program, adr, result.data.value);
        } else if (result.form == RESULT::PROMISE) {
            // The returned promise (of an operation returning TYPE, as Return's
            // constructor checked) takes over this operation's return address.
            detail::forward_promise<0>(program, result.data.promise, adr);
        } else {
            ReturnAdr<TYPE> src = result.data.future;
            Promise<ReturnOp> forward((TaggedSemaphore*)adr.base, (void*)adr.dest, (void*)src.dest, sizeof(TYPE));
            bool unit = src.base == nullptr || src.base->tag == UNIT_BARRIER_TAG;
            if (unit) {
                // A unit future cannot hold the forwarding promise; it can
                // only be forwarded once ready (a null base counts as ready).
                if (src.base == nullptr || hz_load(&src.base->sem) == 0u) {
                    hz_fence_acquire();
                    program.async_call(forward);
                } else {
                    program.usage_error();
                }
            } else if constexpr (PROGRAM::REMAPPING_AWAIT) {
                reinterpret_cast<RemappingBarrier<typename PROGRAM::OpSet>*>(src.base)->await(program, forward);
            } else {
                program.usage_error();
            }
        }
    } else {
        resolve_value(program, adr, (TYPE)result);
    }
}

}  // namespace hz

// Copies the resolved value and counts one fulfilment of the destination
// (reference async/future.h:41-58).
template <typename PROGRAM>
__device__ HZ_INLINE void ReturnOp::eval(PROGRAM program, TaggedSemaphore* sem, void* dst, void* src, size_t size) {
    if (dst != nullptr && src != nullptr) {
        unsigned char* d = (unsigned char*)dst;
        const unsigned char* s = (const unsigned char*)src;
        for (size_t i = 0; i < size; ++i) d[i] = s[i];
    }
    hz::fulfill(program, sem);
}
