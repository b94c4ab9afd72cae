// Portable base of the spec API: the pieces that compile both for gfx950 (hipcc)
// and for a host compiler. It expects __host__/__device__ to be defined by the
// including platform layer (hip_runtime.h on the device side).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <utility>

#define HZ_HD __host__ __device__
#define HZ_INLINE __forceinline__

namespace hz {

template <typename T>
HZ_HD constexpr T ceil_div(T a, T b) { return (a + b - 1) / b; }

template <typename T>
HZ_HD constexpr T round_up(T a, T b) { return ceil_div(a, b) * b; }

// Number of 32-bit words a trivially copyable value occupies once padded.
template <typename T>
struct WordCount {
    static constexpr uint32_t value = (uint32_t)((sizeof(T) + 3) / 4);
};

}  // namespace hz
