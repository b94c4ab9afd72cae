// Operation sets, promises and promise unions: the type-level ABI that program
// specifications are written against.
//
// Same names and meaning as the reference's async/promise.h (OpUnion :17,
// OpUnionLookup :65-91, ArgTuple :247-305, OpType :312-319, Promise :439-490,
// PromiseUnion :503-696, PromiseEnum :705-742), so specifications written for
// the reference compile unchanged. The storage is laid out differently:
//  - a Promise of a void operation carries only its arguments (the reference
//    adds an unused return address), so the Collatz promise is 16 bytes, not 32;
//  - ArgTuple has no trailing empty member, so argument packs are packed the
//    way the platform ABI packs a plain struct of the same fields.
// Packed promises are copied word by word into structure-of-arrays work links
// (one 32-bit row per word, one column per lane of a wavefront).
#pragma once

#include "portable.h"

using OpDisc = unsigned int;

template <typename... OPS>
struct OpUnion {};

template <typename OPERATION>
struct Promise;

template <typename OP_UNION>
union PromiseUnion;

template <typename OP_UNION>
struct PromiseEnum;

struct VoidState {};

template <typename OP_SET, typename ADR_TYPE = unsigned int>
struct RemappingBarrier;

template <typename OP_SET, typename ADR_TYPE = unsigned int>
struct UnitBarrier;

template <typename TYPE, typename BARRIER>
struct Future;

// Internal operation that fills in a future's value and resolves its barrier
// (reference async/future.h:37-60); defined in future.h.
struct ReturnOp;

// What a value-returning operation's eval may return: a value, a promise of
// an operation returning the same type, or the address of a future whose value
// becomes the result (reference async/promise.h:350-434); defined in future.h.
template <typename PROGRAM, typename TYPE>
struct Return;

namespace hz {
constexpr unsigned int BARRIER_NO_LINK = 0xFFFFFFFFu;   // no link / empty list
constexpr unsigned int BARRIER_RELEASED = 0xFFFFFFFEu;  // partial word or list head after release
constexpr unsigned int BARRIER_TAG = 1u;
constexpr unsigned int UNIT_BARRIER_TAG = 0u;
}  // namespace hz

namespace hz {
namespace detail {

template <typename Q, typename... OPS>
struct IndexOf;

template <typename Q>
struct IndexOf<Q> {
    static constexpr bool found = false;
    static constexpr unsigned value = 0;
};

template <typename Q, typename H, typename... T>
struct IndexOf<Q, H, T...> {
    static constexpr bool found = std::is_same<Q, H>::value || IndexOf<Q, T...>::found;
    static constexpr unsigned value =
        std::is_same<Q, H>::value ? 0u : 1u + IndexOf<Q, T...>::value;
};

template <unsigned I, typename... OPS>
struct OpAt;

template <typename H, typename... T>
struct OpAt<0, H, T...> {
    using type = H;
};

template <unsigned I, typename H, typename... T>
struct OpAt<I, H, T...> {
    using type = typename OpAt<I - 1, T...>::type;
};

}  // namespace detail

// The operation with discriminant I in an OpUnion.
template <unsigned I, typename OP_SET>
struct OpOf;
template <unsigned I, typename... OPS>
struct OpOf<I, OpUnion<OPS...>> {
    using type = typename detail::OpAt<I, OPS...>::type;
};
}  // namespace hz

// Compile-time facts about one operation type relative to an operation set.
// DISC is the operation's position in the set (first operation = 0), which is
// the discriminant stored with every work link.
template <typename QUERY, typename OP_UNION>
struct OpUnionLookup;

template <typename QUERY, typename... OPS>
struct OpUnionLookup<QUERY, OpUnion<OPS...>> {
    static constexpr bool CONTAINED = hz::detail::IndexOf<QUERY, OPS...>::found;
    static constexpr OpDisc DEPTH = sizeof...(OPS);
    static constexpr OpDisc DISC = hz::detail::IndexOf<QUERY, OPS...>::value;
};

template <typename HEAD, typename TAIL_UNION>
struct OpUnionAppend;

template <typename HEAD, typename... TAIL>
struct OpUnionAppend<HEAD, OpUnion<TAIL...>> {
    using Type = OpUnion<HEAD, TAIL...>;
};

// Subset / equality relations between two operation sets.
template <typename LEFT, typename RIGHT>
struct OpUnionPair;

template <typename... L, typename... R>
struct OpUnionPair<OpUnion<L...>, OpUnion<R...>> {
    static constexpr bool LEFT_SUBSET = (true && ... && OpUnionLookup<L, OpUnion<R...>>::CONTAINED);
    static constexpr bool RIGHT_SUBSET = (true && ... && OpUnionLookup<R, OpUnion<L...>>::CONTAINED);
    static constexpr bool EQUAL = LEFT_SUBSET && RIGHT_SUBSET;
};

// A semaphore that also records what kind of object owns it.
struct TaggedSemaphore {
    unsigned int sem;
    unsigned int tag;

    TaggedSemaphore() = default;
    HZ_HD TaggedSemaphore(unsigned int sem_val, unsigned int tag_val) : sem(sem_val), tag(tag_val) {}
};

// Internal operation that resolves a future: it copies `size` bytes of the
// value from `src` to `dst` and then decrements the barrier whose semaphore
// is `sem` (reference async/future.h:37-60). Program types add it to the
// operation set of a specification that has value-returning operations.
struct ReturnOp {
    using Type = void (*)(TaggedSemaphore*, void*, void*, size_t);
    template <typename PROGRAM>
    __device__ HZ_INLINE static void eval(PROGRAM program, TaggedSemaphore* sem, void* dst, void* src, size_t size);
};

// An argument filled in from device memory (reference LazyLoad,
// async/promise.h:204-240, and ASYNC_LOADS, :284-303). It is built at the
// call site from a reference to the value in device memory.
//  - Passed where the operation takes a plain TYPE (the reference's form
//    without ASYNC_LOADS), it converts at once: the value is read when the
//    promise is built.
//  - Declared as the operation's parameter (`LazyLoad<TYPE>`), the promise
//    carries the address. AsyncProgram loads the values of a whole link with
//    direct-to-LDS loads issued while the link before it is evaluated, and
//    hands them to the operation in `value`; elsewhere, or if the link was
//    not prefetched, get() reads the source when called.
// TYPE must be trivially copyable, 4-byte aligned and a whole number of
// 32-bit words. Bit 0 of `source` (free, by that alignment) marks a filled
// `value`.
template <typename TYPE>
struct LazyLoad {
    using Type = TYPE;
    static_assert(alignof(TYPE) >= 4 && sizeof(TYPE) % 4 == 0, "LazyLoad values are whole aligned 32-bit words");
    uintptr_t source;
    TYPE value;
    LazyLoad() = default;
    HZ_HD LazyLoad(const TYPE& ref) : source((uintptr_t)&ref), value() {}
    HZ_HD TYPE get() const { return (source & 1u) ? value : *(const TYPE*)source; }
    HZ_HD operator TYPE() const { return get(); }
    // Hands over the loaded value.
    HZ_HD void fill(const TYPE& v) {
        value = v;
        source |= 1u;
    }
};

template <typename... TYPES>
HZ_HD bool has_load(TYPES...) { return false; }
template <typename HEAD, typename... TAIL>
HZ_HD bool has_load(LazyLoad<HEAD>, TAIL...) { return true; }
template <typename HEAD, typename... TAIL>
HZ_HD bool has_load(HEAD, TAIL... tail) { return has_load(tail...); }

// Packed argument storage. A one-element tuple has no tail member so that a
// pack of (u64, u32, u32) occupies exactly 16 bytes.
template <typename... TAIL>
struct ArgTuple;

template <>
struct ArgTuple<> {
    ArgTuple() = default;
};

template <typename HEAD>
struct ArgTuple<HEAD> {
    using Head = HEAD;
    Head head;
    ArgTuple() = default;
    HZ_HD ArgTuple(HEAD h) : head(h) {}
};

template <typename HEAD, typename NEXT, typename... TAIL>
struct ArgTuple<HEAD, NEXT, TAIL...> {
    using Head = HEAD;
    using Tail = ArgTuple<NEXT, TAIL...>;
    Head head;
    Tail tail;
    ArgTuple() = default;
    HZ_HD ArgTuple(HEAD h, NEXT n, TAIL... t) : head(h), tail(n, t...) {}
};

namespace hz {
// The first LazyLoad<T> parameter of an argument tuple, if any: HAS, the
// loaded type Value, and of(t), the member itself.
template <typename TUPLE>
struct FirstLazy {
    static constexpr bool HAS = false;
    using Value = unsigned int;
};
template <typename T>
struct FirstLazy<ArgTuple<LazyLoad<T>>> {
    static constexpr bool HAS = true;
    using Value = T;
    static HZ_HD LazyLoad<T>& of(ArgTuple<LazyLoad<T>>& t) { return t.head; }
};
template <typename T, typename N, typename... R>
struct FirstLazy<ArgTuple<LazyLoad<T>, N, R...>> {
    static constexpr bool HAS = true;
    using Value = T;
    static HZ_HD LazyLoad<T>& of(ArgTuple<LazyLoad<T>, N, R...>& t) { return t.head; }
};
template <typename H, typename N, typename... R>
struct FirstLazy<ArgTuple<H, N, R...>> {
    using Rest = FirstLazy<ArgTuple<N, R...>>;
    static constexpr bool HAS = Rest::HAS;
    using Value = typename Rest::Value;
    static HZ_HD auto& of(ArgTuple<H, N, R...>& t) { return Rest::of(t.tail); }
};
}  // namespace hz

// Deduces the return and argument types of an operation from `using Type = R(*)(Args...)`.
template <typename TYPE>
struct OpType;

template <typename RETURN, typename... ARGS>
struct OpType<RETURN (*)(ARGS...)> {
    typedef RETURN Return;
    typedef ArgTuple<ARGS...> Args;
};

// Where a value-returning operation delivers its result: the semaphore of the
// future/barrier to decrement and the value slot to fill.
template <typename TYPE>
struct ReturnAdr {
    using Type = TYPE;
    TaggedSemaphore* base;
    Type* dest;
};

template <>
struct ReturnAdr<void> {
    using Type = void;
    TaggedSemaphore* base;
};

namespace hz {

// Delivers the result of a value-returning operation to its return address:
// a value is stored and the address's barrier decremented; a returned promise
// inherits the address and is queued; a returned future forwards its value
// when it resolves. Defined in future.h.
template <typename PROGRAM, typename TYPE, typename RESULT>
__device__ __host__ HZ_INLINE void resolve(PROGRAM program, ReturnAdr<TYPE> adr, RESULT result);

namespace detail {

// Void operations carry no return address at all (empty base).
template <typename RETURN>
struct PromiseReturn {
    ReturnAdr<RETURN> ret;
};
template <>
struct PromiseReturn<void> {};

template <typename OP, typename PROGRAM, typename TUPLE, typename... UNROLLED>
struct Unpack;

template <typename OP, typename PROGRAM, typename... UNROLLED>
struct Unpack<OP, PROGRAM, ArgTuple<>, UNROLLED...> {
    __device__ __host__ HZ_INLINE static decltype(auto) call(PROGRAM prog, const ArgTuple<>&, UNROLLED... u) {
        return OP::template eval<PROGRAM>(prog, u...);
    }
};

template <typename OP, typename PROGRAM, typename H, typename... UNROLLED>
struct Unpack<OP, PROGRAM, ArgTuple<H>, UNROLLED...> {
    __device__ __host__ HZ_INLINE static decltype(auto) call(PROGRAM prog, const ArgTuple<H>& t, UNROLLED... u) {
        return OP::template eval<PROGRAM>(prog, u..., t.head);
    }
};

template <typename OP, typename PROGRAM, typename H, typename N, typename... T, typename... UNROLLED>
struct Unpack<OP, PROGRAM, ArgTuple<H, N, T...>, UNROLLED...> {
    __device__ __host__ HZ_INLINE static decltype(auto) call(PROGRAM prog, const ArgTuple<H, N, T...>& t, UNROLLED... u) {
        return Unpack<OP, PROGRAM, ArgTuple<N, T...>, UNROLLED..., H>::call(prog, t.tail, u..., t.head);
    }
};

}  // namespace detail
}  // namespace hz

template <typename OPERATION>
struct Promise : hz::detail::PromiseReturn<typename OpType<typename OPERATION::Type>::Return> {
    typedef typename OpType<typename OPERATION::Type>::Return Return;
    typedef typename OpType<typename OPERATION::Type>::Args Args;

    Args args;

    Promise() = default;

    // A promise of a value-returning operation starts with a null return
    // address (its result is dropped); set `ret` (Future::adr() gives one) to
    // deliver the result. `Promise<Op>()` value-initializes, so it is null too.
    template <typename... ARGS>
    HZ_HD Promise(ARGS... a) : args(a...) {
        if constexpr (!std::is_void<Return>::value) this->ret = ReturnAdr<Return>{nullptr, nullptr};
    }

    // Evaluates the operation with the stored arguments on the given program
    // and returns what its eval returns.
    template <typename PROGRAM>
    __device__ __host__ HZ_INLINE decltype(auto) operator()(PROGRAM program) {
        return hz::detail::Unpack<OPERATION, PROGRAM, Args>::call(program, args);
    }

    // Evaluates the promise the way the program types do: the result of a
    // value-returning operation goes to `ret` (hz::resolve, future.h).
    template <typename PROGRAM>
    __device__ __host__ HZ_INLINE void run(PROGRAM program) {
        if constexpr (std::is_void<Return>::value) {
            (*this)(program);
        } else {
            hz::resolve(program, this->ret, (*this)(program));
        }
    }
};

// A union over the promise types of an operation set. Which member is live is
// not stored here; work links carry it as their discriminant.
template <>
union PromiseUnion<OpUnion<>> {
    struct Info {
        static constexpr OpDisc COUNT = 0;
    };
    template <typename TYPE>
    struct Lookup {
        typedef OpUnionLookup<TYPE, OpUnion<>> type;
    };
    template <typename PROGRAM, typename TYPE>
    __device__ __host__ void rigid_eval(PROGRAM) {}
    template <typename PROGRAM>
    __device__ __host__ void loose_eval(PROGRAM, OpDisc) {}
    HZ_HD void dyn_copy_as(OpDisc, PromiseUnion<OpUnion<>>&) {}
    PromiseUnion() = default;
};

template <typename HEAD, typename... TAIL>
union PromiseUnion<OpUnion<HEAD, TAIL...>> {
    using Head = Promise<HEAD>;
    using Tail = PromiseUnion<OpUnion<TAIL...>>;
    using OpSet = OpUnion<HEAD, TAIL...>;
    using TailOpSet = OpUnion<TAIL...>;

    Head head_form;
    Tail tail_form;

    struct Info {
        static constexpr OpDisc COUNT = sizeof...(TAIL) + 1;
        static constexpr OpDisc INDEX = 0;
    };

    template <typename TYPE>
    struct Lookup {
        typedef OpUnionLookup<TYPE, OpSet> type;
    };

    PromiseUnion() = default;

    template <typename OP_TYPE>
    HZ_HD PromiseUnion(Promise<OP_TYPE> prom) {
        static_assert(OpUnionLookup<OP_TYPE, OpSet>::CONTAINED,
                      "Promise type is not part of this promise union's operation set.");
        cast<OP_TYPE>() = prom;
    }

    // Views the union as the promise type of TYPE.
    template <typename TYPE>
    HZ_HD HZ_INLINE Promise<TYPE>& cast() {
        static_assert(OpUnionLookup<TYPE, OpSet>::CONTAINED,
                      "Promise type is not part of this promise union's operation set.");
        if constexpr (std::is_same<TYPE, HEAD>::value) {
            return head_form;
        } else {
            return tail_form.template cast<TYPE>();
        }
    }

    // Evaluates the union as a promise of the statically known operation TYPE.
    template <typename PROGRAM, typename TYPE>
    __device__ __host__ HZ_INLINE void rigid_eval(PROGRAM program) {
        cast<TYPE>().run(program);
    }

    // Evaluates the union as a promise of the operation with discriminant `disc`.
    template <typename PROGRAM>
    __device__ __host__ HZ_INLINE void loose_eval(PROGRAM program, OpDisc disc) {
        if (disc == 0) {
            head_form.run(program);
        } else {
            if constexpr (sizeof...(TAIL) > 0) {
                tail_form.template loose_eval<PROGRAM>(program, disc - 1);
            }
        }
    }

    template <typename TYPE>
    HZ_HD void copy_as(PromiseUnion<OpSet>& other) {
        cast<TYPE>() = other.template cast<TYPE>();
    }

    HZ_HD void dyn_copy_as(OpDisc disc, PromiseUnion<OpSet>& other) {
        if (disc == 0) {
            head_form = other.head_form;
        } else {
            tail_form.dyn_copy_as(disc - 1, other.tail_form);
        }
    }

};

// A promise union tagged with its discriminant.
template <typename OP_SET>
struct PromiseEnum {
    using UnionType = PromiseUnion<OP_SET>;
    UnionType data;
    OpDisc disc;

    PromiseEnum() = default;
    HZ_HD PromiseEnum(UnionType uni, OpDisc d) : data(uni), disc(d) {}

    template <typename OP_TYPE>
    HZ_HD PromiseEnum(Promise<OP_TYPE> promise)
        : data(promise), disc(UnionType::template Lookup<OP_TYPE>::type::DISC) {}
};

namespace hz {

// Number of 32-bit rows one promise of OP_SET occupies in a work link.
template <typename OP_SET>
struct PromiseWords {
    static constexpr uint32_t value = WordCount<PromiseUnion<OP_SET>>::value;
};

// The operation set a program type executes for a specification's OpSet:
// ReturnOp is appended when some operation returns a value (futures are
// resolved by ReturnOp promises held on barriers) and is not already there.
// Specifications whose operations all return void keep their set unchanged.
template <typename OP_SET>
struct ExecOpSet;
template <typename... OPS>
struct ExecOpSet<OpUnion<OPS...>> {
    static constexpr bool returns_values =
        (false || ... || !std::is_void<typename OpType<typename OPS::Type>::Return>::value);
    static constexpr bool has_return_op = (false || ... || std::is_same<OPS, ReturnOp>::value);
    using type = typename std::conditional<returns_values && !has_return_op, OpUnion<OPS..., ReturnOp>,
                                           OpUnion<OPS...>>::type;
};

template <typename OP_SET>
struct OpCount;
template <typename... OPS>
struct OpCount<OpUnion<OPS...>> {
    static constexpr uint32_t value = sizeof...(OPS);
};

}  // namespace hz
