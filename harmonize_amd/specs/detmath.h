// Deterministic elementary functions for specifications whose host and GPU
// results must agree bit for bit. They use only IEEE-754 binary64/binary32
// additions, multiplications, conversions and integer bit operations, which
// gfx950 and x86-64 both round to nearest even, and are compiled with
// floating-point contraction off on both sides. No libm, no hardware
// approximation instructions.
//
//  - det_log: natural logarithm of a positive normal double. Exponent and
//    mantissa are split from the bit pattern, the mantissa is folded into
//    (sqrt(1/2), sqrt(2)], and log1p of the remainder f (|f| <= 0.4143) is
//    the Maclaurin series f - f^2/2 + f^3/3 - ... to 30 terms (truncation
//    below 1e-13 relative), evaluated with Estrin's scheme; plus e * ln 2.
//  - det_sqrt: square root of a positive normal double from Newton's
//    iteration for 1/sqrt(x), y <- y * (3/2 - x/2 * y^2), started from the
//    classic shifted-exponent estimate (Lomont, "Fast inverse square root",
//    2003, constant for binary64) and run 4 times (quadratic convergence
//    from a 3.5% start leaves a relative error below 1e-20), then
//    multiplied by x.
#pragma once

#include <cstdint>
#include <cstring>

#ifndef HZ_DET_HD
#if defined(__HIPCC__)
#define HZ_DET_HD __host__ __device__ __forceinline__
#else
#define HZ_DET_HD inline
#endif
#endif

namespace hz {
namespace det {

HZ_DET_HD uint64_t bits_of(double x) {
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    return b;
}

HZ_DET_HD double double_of(uint64_t b) {
    double x;
    std::memcpy(&x, &b, sizeof x);
    return x;
}

constexpr double LN2 = 0.6931471805599453;  // ln 2 rounded to binary64
constexpr double SQRT2 = 1.4142135623730951;
constexpr int LOG_TERMS = 30;

// Series coefficient (-1)^(n+1)/n, rounded once by the compiler's constant
// evaluator (correctly rounded in both compilers), never divided at run time.
template <int N>
struct LogCoef {
    static constexpr double value = ((N & 1) ? 1.0 : -1.0) / (double)N;
};

// Estrin's scheme for P(f) = sum_{n=1}^{30} c_n f^(n-1): a fixed expression
// tree whose dependency depth is about 10 operations (Horner's rule would
// chain 60). The tree is spelled out, so both compilers evaluate the same
// operations in the same order.
HZ_DET_HD double log_series(double f) {
    const double f2 = f * f;
    const double f4 = f2 * f2;
    const double f8 = f4 * f4;
    const double f16 = f8 * f8;
#define HZ_PAIR(a) (LogCoef<a>::value + f * LogCoef<a + 1>::value)
    // 15 pairs (c_a + c_{a+1} f), a = 1, 3, ..., 29
    const double q1 = HZ_PAIR(1) + f2 * HZ_PAIR(3);
    const double q5 = HZ_PAIR(5) + f2 * HZ_PAIR(7);
    const double q9 = HZ_PAIR(9) + f2 * HZ_PAIR(11);
    const double q13 = HZ_PAIR(13) + f2 * HZ_PAIR(15);
    const double q17 = HZ_PAIR(17) + f2 * HZ_PAIR(19);
    const double q21 = HZ_PAIR(21) + f2 * HZ_PAIR(23);
    const double q25 = HZ_PAIR(25) + f2 * HZ_PAIR(27);
    const double q29 = HZ_PAIR(29);
#undef HZ_PAIR
    const double o1 = q1 + f4 * q5;
    const double o9 = q9 + f4 * q13;
    const double o17 = q17 + f4 * q21;
    const double o25 = q25 + f4 * q29;
    const double h1 = o1 + f8 * o9;
    const double h17 = o17 + f8 * o25;
    return h1 + f16 * h17;
}

// log(x) for positive normal x.
HZ_DET_HD double det_log(double x) {
    uint64_t b = bits_of(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = double_of((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);  // [1, 2)
    if (m > SQRT2) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;  // exact: m in (0.70, 1.42]
    return (double)e * LN2 + f * log_series(f);
}

// sqrt(x) for positive normal x.
HZ_DET_HD double det_sqrt(double x) {
    double y = double_of(0x5FE6EB50C7B537A9ull - (bits_of(x) >> 1));
    double hx = 0.5 * x;
    for (int i = 0; i < 4; ++i) {
        y = y * (1.5 - hx * (y * y));
    }
    return x * y;
}

}  // namespace det
}  // namespace hz
