// Seeded Collatz start generator, shared by the Collatz entry points and the
// serving loop: start i = (Philox4x32-10 of counter (i, 0, 0, 0) under key
// seed, words 1:0 as a u64) mod (2^bits - 1) + 1, uniform in [1, 2^bits).
#pragma once

#include "philox.h"

namespace hz {
namespace collatz {

static __global__ void __launch_bounds__(256)
    collatz_starts_kernel(unsigned long long seed, unsigned int bits, unsigned int count, unsigned long long* out) {
    const unsigned long long modulus = (1ull << bits) - 1ull;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        hz::philox::Ctr r = hz::philox::philox4x32_10(i, 0u, 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
        unsigned long long x = ((unsigned long long)r.v[1] << 32) | r.v[0];
        out[i] = x % modulus + 1ull;
    }
}

// Starts first .. first + count - 1 of the same sequence.
static __global__ void __launch_bounds__(256) collatz_starts_range_kernel(unsigned long long seed, unsigned int bits,
                                                                           unsigned int first, unsigned int count,
                                                                           unsigned long long* out) {
    const unsigned long long modulus = (1ull << bits) - 1ull;
    for (unsigned int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const unsigned int i = first + t;
        hz::philox::Ctr r = hz::philox::philox4x32_10(i, 0u, 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
        unsigned long long x = ((unsigned long long)r.v[1] << 32) | r.v[0];
        out[i] = x % modulus + 1ull;
    }
}

}  // namespace collatz
}  // namespace hz

