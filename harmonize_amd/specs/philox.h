// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), written from the
// published round function: per round, two 32x32->64 multiplies of counter
// words 0 and 2, an xor of the high halves with counter words 1 and 3 and
// the key, and a Weyl increment of the key between rounds.
//
// Portable: compiles for gfx950 and for the host, with identical results,
// so host executors and GPU programs draw the same numbers for the same
// (key, counter).
#pragma once

#include <cstdint>

#ifndef HZ_PHILOX_HD
#if defined(__HIPCC__)
#define HZ_PHILOX_HD __host__ __device__ __forceinline__
#else
#define HZ_PHILOX_HD inline
#endif
#endif

namespace hz {
namespace philox {

struct Ctr {
    uint32_t v[4];
};
struct Key {
    uint32_t v[2];
};

// Round multipliers and key increments of Philox4x32 from the paper.
constexpr uint32_t M0 = 0xD2511F53u;
constexpr uint32_t M1 = 0xCD9E8D57u;
constexpr uint32_t W0 = 0x9E3779B9u;
constexpr uint32_t W1 = 0xBB67AE85u;

HZ_PHILOX_HD void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

HZ_PHILOX_HD Ctr round(Ctr c, Key k) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, c.v[0], hi0, lo0);
    mulhilo(M1, c.v[2], hi1, lo1);
    Ctr r;
    r.v[0] = hi1 ^ c.v[1] ^ k.v[0];
    r.v[1] = lo1;
    r.v[2] = hi0 ^ c.v[3] ^ k.v[1];
    r.v[3] = lo0;
    return r;
}

// Philox4x32 with `rounds` rounds (10 for Philox4x32-10).
HZ_PHILOX_HD Ctr generate(Ctr c, Key k, int rounds = 10) {
    for (int i = 0; i < rounds; ++i) {
        if (i > 0) {
            k.v[0] += W0;
            k.v[1] += W1;
        }
        c = round(c, k);
    }
    return c;
}

HZ_PHILOX_HD Ctr philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    Ctr c{{c0, c1, c2, c3}};
    Key k{{k0, k1}};
    return generate(c, k, 10);
}

// Uniform float in (0, 1]: the top 24 bits of x, plus one, times 2^-24.
// Exact in binary32, never 0, so -log(u) is finite.
HZ_PHILOX_HD float u01_open_closed(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

// Uniform float in [0, 1): the top 24 bits of x times 2^-24.
HZ_PHILOX_HD float u01_closed_open(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// Uniform double in (0, 1] from 53 bits of a and b.
HZ_PHILOX_HD double u01d_open_closed(uint32_t a, uint32_t b) {
    uint64_t m = ((uint64_t)(a >> 6) << 27) | (uint64_t)(b >> 5);  // 26 + 27 = 53 bits
    return (double)(m + 1u) * (1.0 / 9007199254740992.0);
}

}  // namespace philox
}  // namespace hz
