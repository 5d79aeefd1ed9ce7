// fixed_sum.hpp — exact sums of f64 terms in fixed point, the row accumulator
// of PageRank's binned sweep (pagerank.hip) and of ursa_hip_exact_sum.
//
// A term x with |x| < 4 becomes the integer round_toward_zero(x * 2^100),
// split as lo + hi * 2^40 with lo in [0, 2^40) and hi signed; sums of the two
// parts are kept separately (lo in a u64, hi in an i64), so two independent
// integer additions accumulate a term and the result does not depend on the
// order of the additions. A u64 low sum absorbs 2^24 terms without overflow;
// the i64 high sum holds sums below 8 in magnitude (2^63 units of 2^-60).
// fx_value rounds the exact sum to the nearest f64 (ties to even).
//
// Plain C++ as well as HIP, so the host tests can check it against exact
// rational sums.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FX_HD __host__ __device__ __forceinline__
#else
#define FX_HD inline
#endif

namespace ursa {

constexpr int FX_LO_BITS = 40;
constexpr unsigned long long FX_LO_MASK = (1ull << FX_LO_BITS) - 1;
constexpr int FX_FRAC = 100;                 // fixed point: units of 2^-100

// x (|x| < 4) as lo + hi * 2^40 in units of 2^-100, lo in [0, 2^40): the
// magnitude truncated below 2^-100, then negated in two's complement.
FX_HD void fx_split(double x, unsigned long long& lo, long long& hi) {
    unsigned long long bits;
    __builtin_memcpy(&bits, &x, 8);
    const int be = (int)((bits >> 52) & 0x7FF);
    unsigned long long mant = bits & ((1ull << 52) - 1);
    int sh = 1 - 975;  // subnormal: mant * 2^(1-1075) = mant * 2^(sh) units
    if (be) {
        mant |= 1ull << 52;
        sh = be - 975;
    }
    unsigned long long L, H;
    if (sh >= FX_LO_BITS) {  // at most 49 for |x| < 4
        L = 0;
        H = mant << (sh - FX_LO_BITS);
    } else if (sh >= 0) {
        L = (mant << sh) & FX_LO_MASK;
        H = mant >> (FX_LO_BITS - sh);
    } else {
        const unsigned long long t = sh > -64 ? mant >> (-sh) : 0ull;
        L = t & FX_LO_MASK;
        H = t >> FX_LO_BITS;
    }
    if (bits >> 63) {
        if (L) {
            lo = (1ull << FX_LO_BITS) - L;
            hi = -(long long)H - 1;
        } else {
            lo = 0;
            hi = -(long long)H;
        }
    } else {
        lo = L;
        hi = (long long)H;
    }
}

// u64 -> f64 rounded to nearest even: both halves convert exactly, one add rounds.
FX_HD double u64_to_f64_rn(unsigned long long v) {
    return __builtin_fma((double)(uint32_t)(v >> 32), 4294967296.0, (double)(uint32_t)v);
}

// (lo + hi * 2^40) * 2^-100 rounded to nearest even; lo may hold any u64 sum.
FX_HD double fx_value(unsigned long long lo, long long hi) {
    hi += (long long)(lo >> FX_LO_BITS);
    lo &= FX_LO_MASK;
    const bool neg = hi < 0;
    unsigned long long mh = (unsigned long long)hi, ml = lo;
    if (neg) {
        if (lo) {
            mh = (unsigned long long)(-(hi + 1));
            ml = (1ull << FX_LO_BITS) - lo;
        } else {
            mh = (unsigned long long)(-hi);
            ml = 0;
        }
    }
    // magnitude = A * 2^64 + B
    const unsigned long long A = mh >> (64 - FX_LO_BITS), B = (mh << FX_LO_BITS) | ml;
    double r;
    if (A == 0) {
        r = __builtin_ldexp(u64_to_f64_rn(B), -FX_FRAC);
    } else {
        const int p = 127 - __builtin_clzll(A);  // top bit of the 128-bit magnitude (>= 64)
        const int s = p - 63;                    // 1 .. 39
        const unsigned long long T = (A << (64 - s)) | (B >> s);
        const unsigned long long sticky = (B << (64 - s)) != 0 ? 1ull : 0ull;
        r = __builtin_ldexp(u64_to_f64_rn(T | sticky), s - FX_FRAC);  // the sticky bit keeps the rounding exact
    }
    return neg ? -r : r;
}

}  // namespace ursa
