// Index iterators that specifications use to claim work (reference
// util/iter.h: Iter :13-60, AtomicIter :138-223, ArrayIter :228-314,
// IOBuffer :320-474).
//
// Portable: the including platform layer provides hz_fetch_add (a
// fetch-and-add; on gfx950 it is aggregated across the active lanes of a
// wavefront) and hz_load (a relaxed atomic load).
#pragma once

#include "portable.h"

namespace util {
namespace iter {

template <typename INDEX_TYPE>
struct Iter {
    typedef INDEX_TYPE IndexType;
    IndexType value;
    IndexType limit;

    Iter() = default;
    HZ_HD Iter(IndexType v, IndexType l) : value(v), limit(l) {}

    HZ_HD void reset(IndexType start_val, IndexType limit_val) {
        value = start_val;
        limit = limit_val;
    }

    // Yields the next index, or false once the range is exhausted.
    HZ_HD bool step(IndexType& iter_val) {
        if (value >= limit) return false;
        iter_val = value;
        value += 1;
        return true;
    }

    // Splits off the next `length` indices (fewer at the end of the range).
    HZ_HD Iter<IndexType> leap(IndexType length) {
        IndexType old_value = value;
        IndexType sub_limit = (limit - value > length) ? value + length : limit;
        if (value >= limit) sub_limit = value;
        value = sub_limit;
        return Iter<IndexType>(old_value, sub_limit);
    }

    HZ_HD bool done() const { return value >= limit; }
};

// A range shared by every thread of a program; claims are atomic.
template <typename INDEX_TYPE>
struct AtomicIter {
    typedef INDEX_TYPE IndexType;
    IndexType value;
    IndexType limit;

    AtomicIter() = default;
    HZ_HD AtomicIter(IndexType start_val, IndexType limit_val) : value(start_val), limit(limit_val) {}
    HZ_HD AtomicIter(Iter<IndexType> it) : value(it.value), limit(it.limit) {}
    HZ_HD AtomicIter& operator=(Iter<IndexType> it) {
        value = it.value;
        limit = it.limit;
        return *this;
    }

    // Claims the next `leap_size` indices. Claims past the limit come back empty.
    __device__ Iter<IndexType> leap(IndexType leap_size) {
        IndexType start_val = 0;
        IndexType limit_val = 0;
        IndexType lim = limit;
        if (hz_load(&value) < lim) {
            start_val = hz_fetch_add(&value, leap_size);
            limit_val = start_val + leap_size;
            // guard against wrap of the counter past the top of IndexType
            if (limit_val < start_val) limit_val = lim;
            start_val = (start_val < lim) ? start_val : lim;
            limit_val = (limit_val < lim) ? limit_val : lim;
        }
        return Iter<IndexType>(start_val, limit_val);
    }

    // Claims a single index.
    __device__ bool step(IndexType& iter_val) {
        if (hz_load(&value) >= limit) return false;
        IndexType try_val = hz_fetch_add(&value, (IndexType)1);
        if (try_val >= limit) return false;
        iter_val = try_val;
        return true;
    }

    __device__ bool done() const { return hz_load(&value) >= limit; }

    // The reference's barrier demo calls sync_done(); same meaning as done().
    __device__ bool sync_done() const { return done(); }
};

template <typename T, template <typename> class ITER_TYPE = Iter, typename INDEX_TYPE = unsigned int>
struct ArrayIter {
    typedef INDEX_TYPE IndexType;
    typedef ITER_TYPE<IndexType> IterType;

    T* array;
    IterType iter;

    ArrayIter() = default;
    HZ_HD ArrayIter(T* adr, IterType other_iter) : array(adr), iter(other_iter) {}

    HZ_HD void reset(T* new_array, IterType new_iter) {
        array = new_array;
        iter = new_iter;
    }

    __device__ bool step_val(T& val) {
        IndexType index;
        if (iter.step(index)) {
            val = array[index];
            return true;
        }
        return false;
    }

    __device__ bool step_idx_val(IndexType& idx, T& val) {
        IndexType index;
        if (iter.step(index)) {
            idx = index;
            val = array[index];
            return true;
        }
        return false;
    }

    __device__ bool step_ptr(T*& val) {
        IndexType index;
        if (iter.step(index)) {
            val = &array[index];
            return true;
        }
        return false;
    }

    __device__ ArrayIter<T, Iter, IndexType> leap(IndexType leap_size) {
        return ArrayIter<T, Iter, IndexType>(array, iter.leap(leap_size));
    }

    __device__ bool done() { return iter.done(); }
};

// The reference's EventProgram queue type. The native EventProgram keeps its
// queues in structure-of-arrays form instead; this type is kept so that
// specifications that name it (as a field type) still compile.
template <typename T, typename INDEX_TYPE = unsigned int>
struct IOBuffer {
    typedef INDEX_TYPE IndexType;
    bool toggle;
    T* data_a;
    T* data_b;
    IndexType capacity;
    AtomicIter<IndexType> input_iter;
    AtomicIter<IndexType> output_iter;
};

}  // namespace iter
}  // namespace util
