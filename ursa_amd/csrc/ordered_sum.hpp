// ordered_sum.hpp — a left-to-right f64 sum, s = ((0 + x0) + x1) + ..., computed
// in parallel with exactly the value the sequential loop rounds to.
//
// The reference sums PageRank's dangling mass that way (pagerank.rs:76-78,
// `dangling.iter().map(|&u| rank[u]).sum()`), and at R-MAT scale 24 the
// rounding of that loop moves ranks by up to 1e-12 relative from the exact
// sum, so a tree sum would not agree with it to 1e-12.
//
// Why it parallelises: while the running sum s stays in one binade
// [2^e, 2^(e+1)), every representable value there is a multiple of
// u = 2^(e-52), so the rounded step s + x lands on s + q(x)*u with
// q(x) = round(x / u) — a function of x and e alone, unless x / u lies exactly
// halfway between integers (round-half-even then reads the last bit of s).
// Between binade changes the sum is therefore an exact integer sum of the
// quotients q, which may be formed in any order. The few steps that leave the
// binade (the sum grows past 2^(e+1)), meet a halfway case, start from zero or
// see a negative or non-finite element are taken one at a time with an
// ordinary add, which is the reference's own operation.
//
// Two launches:
//   build  one wave per chunk of OS_CHUNK summands: gathers them contiguously
//          into xd, and records for each of OS_BINS candidate binades
//          E0 .. E0+15 the chunk's quotient sum and a stop bit (some summand
//          needs a single step in that binade), plus the chunk's plain f64 sum.
//   walk   one workgroup. From a prefix of the plain chunk sums it predicts
//          each chunk's binade and which chunks will need single steps (the
//          running sum changes binade inside them, or a stop bit is set), and
//          stages those tables and chunks in LDS. One wave then walks the
//          chunks in order, adding whole chunks as exact integer sums and
//          stepping through the others 64 summands at a time. A wrong
//          prediction only sends the walk to global memory; the result is
//          the sequential loop's either way.
// E0 comes from the previous sum, so the candidates cover every running sum
// above 2^-13 of it.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ursa {

constexpr int OS_BINS = 16;
constexpr int OS_CHUNK = 256;                   // summands per chunk (one wave, 4 per lane)
constexpr int OS_PER_LANE = OS_CHUNK / 64;
constexpr int OS_BUILD_THREADS = 256;           // 4 chunks per build workgroup
constexpr int OS_WALK_THREADS = 1024;
constexpr int OS_TMAX = 6144;                   // chunk tables staged in LDS
constexpr int OS_PMAX = 24;                     // predicted single-step chunks staged in LDS
constexpr int OS_PER_THREAD_MAX = OS_TMAX / OS_WALK_THREADS;
constexpr uint64_t OS_LIMIT = 1ull << 53;
constexpr uint64_t OS_MANT = (1ull << 52) - 1;

// q = round(x / 2^(e-52)) when the step s + x with s in binade e can be taken
// as s + q * 2^(e-52): x is +0 or positive finite, not halfway, and x alone
// stays below 2^(e+1). Returns false otherwise.
__device__ __forceinline__ bool os_quot(double x, int e, uint64_t& q) {
    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    const uint32_t ex = (uint32_t)(bits >> 52);  // sign bit lands at 0x800
    q = 0;
    if (ex >= 0x7FF) return false;               // negative, -0, inf or NaN
    uint64_t mant = bits & OS_MANT;
    int scale;
    if (ex == 0) {
        if (mant == 0) return true;  // +0
        scale = -1074;
    } else {
        mant |= 1ull << 52;
        scale = (int)ex - 1075;
    }
    const int sh = scale - e + 52;  // x / u = mant * 2^sh
    if (sh >= 0) {
        if (sh > 10) return false;
        const uint64_t v = mant << sh;
        if (v >= OS_LIMIT) return false;
        q = v;
        return true;
    }
    const int k = -sh;
    if (k > 54) return true;  // x < u / 4: rounds to nothing
    const uint64_t half = 1ull << (k - 1);
    const uint64_t rem = mant & ((half << 1) - 1);
    q = mant >> k;
    if (rem == half) return false;
    q += rem > half ? 1u : 0u;
    return true;
}

// os_quot without branches for the common case: x positive and normal, and
// 1 <= k <= 64 where x / u = mant / 2^k. Returns false for everything else
// (the caller then takes the step with an ordinary add, which is always exact).
__device__ __forceinline__ bool os_quot_fast(double x, int e, uint64_t& q) {
    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    const int ex = (int)(bits >> 52);  // sign bit lands at 0x800
    const uint64_t mant = (bits & OS_MANT) | (1ull << 52);
    const int k = e + 1023 - ex;
    const int sh = (k - 1) & 63;
    q = ((mant >> sh) + 1) >> 1;
    return ex >= 1 && ex < 0x7FF && k >= 1 && k <= 64 && sh != __builtin_ctzll(mant);
}

// Quotients of x for all candidate binades e0 + b, added to acc[b]; bit b of
// stop is set where os_quot would refuse. With x / u = mant / 2^k,
// round(x / u) = (floor(mant / 2^(k-1)) + 1) / 2, and the halfway case is the
// one binade where bit k-1 is mant's lowest set bit.
__device__ __forceinline__ void os_quot_all(double x, int e0, uint64_t (&acc)[OS_BINS], uint32_t& stop) {
    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    const uint32_t ex = (uint32_t)(bits >> 52);
    if (ex >= 0x7FF) {
        stop |= (1u << OS_BINS) - 1;
        return;
    }
    uint64_t mant = bits & OS_MANT;
    int scale;
    if (ex == 0) {
        if (mant == 0) return;
        scale = -1074;
    } else {
        mant |= 1ull << 52;
        scale = (int)ex - 1075;
    }
    const int k0 = e0 - 52 - scale;  // k for binade e0
    if (k0 < 1) {                    // x is large against some candidate binade
#pragma unroll
        for (int b = 0; b < OS_BINS; ++b) {
            uint64_t q;
            if (os_quot(x, e0 + b, q))
                acc[b] += q;
            else
                stop |= 1u << b;
        }
        return;
    }
    const int tz = __builtin_ctzll(mant);
#pragma unroll
    for (int b = 0; b < OS_BINS; ++b) {
        const int s = k0 + b - 1;
        const uint64_t r = s < 64 ? (mant >> s) : 0;
        acc[b] += (r + 1) >> 1;
        if (s == tz) stop |= 1u << b;
    }
}

// Running sum in the fast form: binade e and integer significand in [2^52, 2^53).
// `fast` is false while the sum is zero, tiny (below 2^-1000), negative or not finite.
struct OsState {
    int e;
    uint64_t sig;
    bool fast;
};

__device__ __forceinline__ OsState os_state(double s) {
    const uint64_t b = (uint64_t)__double_as_longlong(s);
    const uint32_t ex = (uint32_t)(b >> 52);
    OsState st;
    st.fast = ex >= 23 && ex < 0x7FF;
    st.e = (int)ex - 1023;
    st.sig = (b & OS_MANT) | (1ull << 52);
    return st;
}

__device__ __forceinline__ double os_value(int e, uint64_t sig) {
    return __longlong_as_double((long long)(((uint64_t)(e + 1023) << 52) | (sig & OS_MANT)));
}

// Inclusive sum over the wave's 64 lanes with DPP row shifts and row
// broadcasts (lanes that receive nothing read 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t os_dpp(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROWS, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, ROWS, 0xF, false);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t os_wave_incl_scan(uint64_t v) {
    v += os_dpp<0x111, 0xF>(v);  // row_shr:1
    v += os_dpp<0x112, 0xF>(v);  // row_shr:2
    v += os_dpp<0x114, 0xF>(v);  // row_shr:4
    v += os_dpp<0x118, 0xF>(v);  // row_shr:8
    v += os_dpp<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += os_dpp<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ uint64_t os_readlane(uint64_t v, uint32_t lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)lane);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double os_readlane(double v, uint32_t lane) {
    return __longlong_as_double((long long)os_readlane((uint64_t)__double_as_longlong(v), lane));
}

// Adds the window's summands (lane i holds x_i, i < cnt <= 64) to s in order;
// one wave, uniform result.
__device__ inline double os_walk_window(double x, uint32_t cnt, double s) {
    const int lane = threadIdx.x & 63;
    uint32_t o = 0;
    while (o < cnt) {
        const OsState st = os_state(s);
        if (!st.fast) {
            s = s + os_readlane(x, o);
            ++o;
            continue;
        }
        const bool live = (uint32_t)lane >= o && (uint32_t)lane < cnt;
        uint64_t q = 0;
        bool ok = true;
        if (live) ok = os_quot(x, st.e, q);
        const uint64_t p = os_wave_incl_scan(live ? q : 0);
        const bool stop = live && (!ok || st.sig + p >= OS_LIMIT);
        const uint64_t mask = __ballot(stop);
        const uint32_t f = mask ? (uint32_t)__builtin_ctzll(mask) : cnt;
        if (f > o) s = os_value(st.e, st.sig + os_readlane(p, f - 1));
        if (f < cnt) {
            s = s + os_readlane(x, f);
            o = f + 1;
        } else {
            o = cnt;
        }
    }
    return s;
}

// Build: chunk c = blockIdx.x * 4 + wave; summand j of the chunk is src[idx[j]].
__device__ inline void os_build(const double* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t d, int e0,
                                double* __restrict__ xd, uint64_t* __restrict__ cq, uint32_t* __restrict__ cm,
                                double* __restrict__ cs) {
    const int lane = threadIdx.x & 63;
    const uint64_t c = (uint64_t)blockIdx.x * (OS_BUILD_THREADS / 64) + (threadIdx.x >> 6);
    const uint64_t j0 = c * OS_CHUNK;
    if (j0 >= d) return;  // wave-uniform
    double xs[OS_PER_LANE];
#pragma unroll
    for (int i = 0; i < OS_PER_LANE; ++i) {
        const uint64_t j = j0 + (uint64_t)i * 64 + lane;
        xs[i] = j < d ? src[idx[j]] : 0.0;
    }
    uint64_t acc[OS_BINS];
#pragma unroll
    for (int b = 0; b < OS_BINS; ++b) acc[b] = 0;
    uint32_t stop = 0;
    double approx = 0.0;
#pragma unroll
    for (int i = 0; i < OS_PER_LANE; ++i) {
        const uint64_t j = j0 + (uint64_t)i * 64 + lane;
        if (j < d) {
            xd[j] = xs[i];
            os_quot_all(xs[i], e0, acc, stop);
            approx += xs[i];
        }
    }
    // Per lane acc[b] < 8 * 2^53, so the wave sums stay below 2^62.
#pragma unroll
    for (int b = 0; b < OS_BINS; ++b)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[b] += __shfl_xor(acc[b], o, 64);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        stop |= __shfl_xor(stop, o, 64);
        approx += __shfl_xor(approx, o, 64);
    }
    if (lane < OS_BINS) {
        uint64_t v = 0;
#pragma unroll
        for (int b = 0; b < OS_BINS; ++b) v = (b == lane) ? acc[b] : v;
        cq[c * OS_BINS + lane] = v < OS_LIMIT ? v : OS_LIMIT;
    }
    if (lane == 0) {
        cm[c] = stop;
        cs[c] = approx;
    }
}

struct OsLds {
    double xs[OS_PMAX * OS_CHUNK];  // staged chunks
    uint64_t pre[OS_TMAX];          // quotient sums, then their prefix within the chunk's segment
    int8_t tb[OS_TMAX];             // predicted binade of chunk c (candidate index), or -1: step through it
    int16_t slot[OS_TMAX];          // staging slot of chunk c, or -1
    uint16_t head[OS_TMAX];         // first chunk of c's segment
    uint16_t last[OS_TMAX];         // for a segment's first chunk: its last chunk
    uint32_t plist[OS_PMAX];
    uint64_t wq[OS_WALK_THREADS / 64];
    int32_t wh[OS_WALK_THREADS / 64];
    int32_t wf[OS_WALK_THREADS / 64];
    double wsum[OS_WALK_THREADS / 64];
    uint32_t nslots;
};

// Adds the summands of lane f (xv[0 .. cnt), uniform index) to s one by one.
__device__ __forceinline__ double os_lane_steps(const double (&xv)[OS_PER_LANE], uint32_t f, uint32_t cnt, double s) {
#pragma unroll
    for (int k = 0; k < OS_PER_LANE; ++k) {
        if ((uint32_t)k < cnt) {
            const double x = os_readlane(xv[k], f);
            const OsState st = os_state(s);
            uint64_t q;
            if (st.fast && os_quot_fast(x, st.e, q) && st.sig + q < OS_LIMIT)
                s = os_value(st.e, st.sig + q);
            else
                s = s + x;
        }
    }
    return s;
}

// Steps through chunk c from LDS or global memory. Lane l holds summands
// 4l .. 4l+3; a pass adds every lane whose four quotients fit, up to the first
// that does not, whose summands are then taken one by one.
__device__ inline double os_walk_chunk(const double* __restrict__ xd, uint64_t d, uint64_t c, const OsLds& L,
                                       double s) {
    const uint64_t j0 = c * OS_CHUNK;
    const uint32_t len = (uint32_t)(d - j0 < (uint64_t)OS_CHUNK ? d - j0 : (uint64_t)OS_CHUNK);
    const double* base = xd + j0;
    if (c < (uint64_t)OS_TMAX && L.slot[c] >= 0) base = L.xs + (uint64_t)L.slot[c] * OS_CHUNK;
    const uint32_t lane = threadIdx.x & 63;
    double xv[OS_PER_LANE];
#pragma unroll
    for (int k = 0; k < OS_PER_LANE; ++k) {
        const uint32_t j = lane * OS_PER_LANE + k;
        xv[k] = j < len ? base[j] : 0.0;
    }
    const uint32_t nl = (len + OS_PER_LANE - 1) / OS_PER_LANE;
    uint32_t o = 0;
    while (o < nl) {
        const uint32_t cnt_o = len - o * OS_PER_LANE < (uint32_t)OS_PER_LANE ? len - o * OS_PER_LANE : OS_PER_LANE;
        const OsState st = os_state(s);
        if (!st.fast) {
            s = os_lane_steps(xv, o, cnt_o, s);
            ++o;
            continue;
        }
        const bool live = lane >= o && lane < nl;
        uint64_t q = 0;
        bool ok = true;
        if (live) {
#pragma unroll
            for (int k = 0; k < OS_PER_LANE; ++k) {
                if (lane * OS_PER_LANE + k < len) {
                    uint64_t qk;
                    ok = os_quot_fast(xv[k], st.e, qk) && ok;
                    q += qk;
                }
            }
        }
        const uint64_t p = os_wave_incl_scan(live && ok ? q : 0);  // < 64 * 4 * 2^53
        const bool stop = live && (!ok || st.sig + p >= OS_LIMIT);
        const uint64_t mask = __ballot(stop);
        const uint32_t f = mask ? (uint32_t)__builtin_ctzll(mask) : nl;
        if (f > o) s = os_value(st.e, st.sig + os_readlane(p, f - 1));
        if (f < nl) {
            const uint32_t cnt_f = len - f * OS_PER_LANE < (uint32_t)OS_PER_LANE ? len - f * OS_PER_LANE : OS_PER_LANE;
            s = os_lane_steps(xv, f, cnt_f, s);
            o = f + 1;
        } else {
            o = nl;
        }
    }
    return s;
}

// Chunks c0 .. nch from the global tables, 64 at a time (chunks past the LDS tables).
__device__ inline double os_walk_global(const double* __restrict__ xd, uint64_t d, const uint64_t* __restrict__ cq,
                                        const uint32_t* __restrict__ cm, uint64_t c0, uint64_t nch, int e0,
                                        const OsLds& L, double s) {
    const int lane = threadIdx.x & 63;
    for (; c0 < nch; c0 += 64) {
        const uint32_t cnt = (uint32_t)(nch - c0 < 64 ? nch - c0 : 64);
        uint32_t o = 0;
        while (o < cnt) {
            const OsState st = os_state(s);
            const int b = st.e - e0;
            const bool inwin = st.fast && b >= 0 && b < OS_BINS;
            const uint64_t c = c0 + lane;
            const bool live = (uint32_t)lane >= o && (uint32_t)lane < cnt;
            uint64_t q = 0;
            bool ok = false;
            if (live && inwin) {
                ok = !((cm[c] >> b) & 1u);
                q = ok ? cq[c * OS_BINS + b] : 0;
            }
            const uint64_t p = os_wave_incl_scan(live && ok ? q : 0);
            const bool stop = live && (!ok || st.sig + p >= OS_LIMIT);
            const uint64_t mask = __ballot(stop);
            const uint32_t f = mask ? (uint32_t)__builtin_ctzll(mask) : cnt;
            if (f > o) s = os_value(st.e, st.sig + os_readlane(p, f - 1));
            if (f < cnt) {
                s = os_walk_chunk(xd, d, c0 + f, L, s);
                o = f + 1;
            } else {
                o = cnt;
            }
        }
    }
    return s;
}

__device__ __forceinline__ uint64_t os_sat_add(uint64_t a, uint64_t b) {
    const uint64_t v = a + b;  // a, b <= 2^53: no wrap
    return v < OS_LIMIT ? v : OS_LIMIT;
}

// Walk: the sequential sum of xd[0 .. d). All OS_WALK_THREADS threads call it;
// the result is valid in wave 0.
//
// Prediction (all threads): a prefix of the plain chunk sums gives each
// chunk's binade. A chunk is stepped through when it is the first, when the
// prefix crosses a power of two inside it (with a 1e-9 margin), or when a
// summand needs a single step in its binade; the others ("regular") are split
// into segments of consecutive regular chunks with one binade, and the
// segment-wise prefix of their quotient sums is kept, so the walk adds a
// whole segment at once and stops only at stepped chunks.
__device__ inline double os_walk(const double* __restrict__ xd, uint64_t d, const uint64_t* __restrict__ cq,
                                 const uint32_t* __restrict__ cm, const double* __restrict__ cs, int e0, OsLds& L) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t nch = (d + OS_CHUNK - 1) / OS_CHUNK;
    const uint64_t lim = nch < (uint64_t)OS_TMAX ? nch : (uint64_t)OS_TMAX;

    // 1. Approximate running sums at chunk starts, over thread-contiguous ranges.
    const uint64_t per = (lim + OS_WALK_THREADS - 1) / OS_WALK_THREADS;
    const uint64_t c_lo = (uint64_t)tid * per < lim ? (uint64_t)tid * per : lim;
    const uint64_t c_hi = c_lo + per < lim ? c_lo + per : lim;
    double part = 0.0;
    for (uint64_t c = c_lo; c < c_hi; ++c) part += cs[c];
    double incl = part;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) L.wsum[wave] = incl;
    if (tid == 0) L.nslots = 0;
    __syncthreads();
    double pre = incl - part;
    for (int w = 0; w < wave; ++w) pre += L.wsum[w];

    // 2. Predicted binade, stepped chunks and their staging slots. The table
    //    loads of a thread's chunks are issued together.
    int pb[OS_PER_THREAD_MAX];
    uint32_t pm[OS_PER_THREAD_MAX];
    uint64_t pq[OS_PER_THREAD_MAX];
#pragma unroll
    for (int i = 0; i < OS_PER_THREAD_MAX; ++i) {
        const uint64_t c = c_lo + i;
        pb[i] = -1;
        if (c < c_hi) {
            const double s0 = pre, s1 = pre + cs[c];
            pre = s1;
            if (c != 0 && s0 > 1.0e-300 && s1 < 1.0e300) {  // normal: the exponent field is the binade
                const int e_lo = (int)((uint64_t)__double_as_longlong(s0 * (1.0 - 1e-9)) >> 52) - 1023;
                const int e_hi = (int)((uint64_t)__double_as_longlong(s1 * (1.0 + 1e-9)) >> 52) - 1023;
                const int b = e_lo - e0;
                if (e_lo == e_hi && b >= 0 && b < OS_BINS) pb[i] = b;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < OS_PER_THREAD_MAX; ++i) {
        const uint64_t c = c_lo + i;
        pm[i] = 0;
        pq[i] = 0;
        if (pb[i] >= 0) {
            pm[i] = cm[c];
            pq[i] = cq[c * OS_BINS + pb[i]];
        }
    }
#pragma unroll
    for (int i = 0; i < OS_PER_THREAD_MAX; ++i) {
        const uint64_t c = c_lo + i;
        if (c >= c_hi) continue;
        const int b = pb[i];
        const bool step = b < 0 || ((pm[i] >> b) & 1u);
        L.tb[c] = step ? (int8_t)-1 : (int8_t)b;
        L.pre[c] = step ? 0 : pq[i];
        L.slot[c] = -1;
        if (step) {
            const uint32_t k = atomicAdd(&L.nslots, 1u);
            if (k < (uint32_t)OS_PMAX) {
                L.plist[k] = (uint32_t)c;
                L.slot[c] = (int16_t)k;
            }
        }
    }
    __syncthreads();

    // 3. Segments: a regular chunk starts one unless its predecessor is regular
    //    with the same binade. Segmented prefix of the quotient sums, clamped
    //    at 2^53 (a segment that reaches it leaves the binade somewhere), and
    //    each chunk's segment head, carried across threads by a scan.
    uint64_t run = 0;
    int32_t hd = -1;
    bool any_head = false;
    for (uint64_t c = c_lo; c < c_hi; ++c) {
        if (L.tb[c] < 0) {
            run = 0;
            hd = -1;
            any_head = true;
            continue;
        }
        if (c == 0 || L.tb[c - 1] != L.tb[c]) {
            run = 0;
            hd = (int32_t)c;
            any_head = true;
        }
        run = os_sat_add(run, L.pre[c]);
    }
    // Carry scan over threads: (run, hd, any_head) combine left to right.
    uint64_t crun = run;
    int32_t chd = hd;
    bool cflag = any_head;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t r2 = __shfl_up(crun, o, 64);
        const int32_t h2 = __shfl_up(chd, o, 64);
        const bool f2 = __shfl_up((int)cflag, o, 64) != 0;
        if (lane >= o && !cflag) {
            crun = os_sat_add(crun, r2);
            chd = h2;
            cflag = f2;
        }
    }
    if (lane == 63) {
        L.wq[wave] = crun;
        L.wh[wave] = chd;
        L.wf[wave] = cflag ? 1 : 0;
    }
    __syncthreads();
    // Exclusive carry into this thread: the lanes before it, then earlier waves.
    uint64_t in_run = __shfl_up(crun, 1, 64);
    int32_t in_hd = __shfl_up(chd, 1, 64);
    bool in_flag = __shfl_up((int)cflag, 1, 64) != 0;
    if (lane == 0) {
        in_run = 0;
        in_hd = -1;
        in_flag = false;
    }
    if (!in_flag) {
        for (int w = wave - 1; w >= 0; --w) {
            in_run = os_sat_add(in_run, L.wq[w]);
            if (L.wf[w]) {
                in_hd = L.wh[w];
                break;
            }
        }
    }
    run = in_run;
    hd = in_hd;
    for (uint64_t c = c_lo; c < c_hi; ++c) {
        if (L.tb[c] < 0) {
            run = 0;
            hd = -1;
            continue;
        }
        if (c == 0 || L.tb[c - 1] != L.tb[c]) {
            run = 0;
            hd = (int32_t)c;
        }
        run = os_sat_add(run, L.pre[c]);
        L.pre[c] = run;
        L.head[c] = (uint16_t)hd;
        if (c + 1 == lim || L.tb[c + 1] != L.tb[c]) L.last[hd] = (uint16_t)c;
    }
    __syncthreads();

    // 4. Stage the stepped chunks.
    const uint32_t ns = L.nslots < (uint32_t)OS_PMAX ? L.nslots : (uint32_t)OS_PMAX;
    for (uint32_t i = tid; i < ns * OS_CHUNK; i += OS_WALK_THREADS) {
        const uint64_t j = (uint64_t)L.plist[i / OS_CHUNK] * OS_CHUNK + i % OS_CHUNK;
        L.xs[i] = j < d ? xd[j] : 0.0;
    }
    __syncthreads();
    if (wave != 0) return 0.0;

    // 5. The walk (uniform across wave 0).
    double s = 0.0;
    uint64_t c = 0;
    while (c < lim) {
        const int b = L.tb[c];
        if (b >= 0) {
            const OsState st = os_state(s);
            const uint64_t h = L.head[c], e = L.last[h];
            const uint64_t before = c > h ? L.pre[c - 1] : 0;
            // Mid-segment, differences of the clamped prefix are exact only below the clamp.
            if (st.fast && st.e - e0 == b && (c == h || L.pre[e] < OS_LIMIT)) {
                if (st.sig + (L.pre[e] - before) < OS_LIMIT) {
                    s = os_value(st.e, st.sig + (L.pre[e] - before));
                    c = e + 1;
                    continue;
                }
                // The sum leaves the binade inside the segment: find the chunk.
                uint64_t lo = c, hi = e;
                while (lo < hi) {
                    const uint64_t mid = (lo + hi) >> 1;
                    if (st.sig + (L.pre[mid] - before) >= OS_LIMIT)
                        hi = mid;
                    else
                        lo = mid + 1;
                }
                if (lo > c) s = os_value(st.e, st.sig + (L.pre[lo - 1] - before));
                c = lo;
            }
        }
        s = os_walk_chunk(xd, d, c, L, s);
        ++c;
    }
    if (lim < nch) s = os_walk_global(xd, d, cq, cm, lim, nch, e0, L, s);
    return s;
}

}  // namespace ursa
