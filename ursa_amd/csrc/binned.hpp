// binned.hpp — propagation-blocked sweeps over PageRank's renumbered in-CSR,
// shared by PageRank (pagerank.hip) and personalized PageRank (ppr.hip).
//
// The pull sweep gathers one contribution per in-slot; on R-MAT scale 24 the
// L2 request rate of those 8-byte gathers, not HBM, sets its time. The binned
// sweep instead streams every byte it touches. Sources are cut into blocks of
// BN_SSPAN (new ids) and rows into bins of at most BN_RSPAN consecutive rows
// and at most BN_T2 in-slots (a row with more in-slots is a bin of its own),
// so that every bin is one work item of phase 2; tile (c, b) holds the
// in-slots from source block c into bin b, rows ascending, padded to whole
// chunks of BN_CH slots (padding slots: source 0, row BN_PAD). The tiles are
// stored twice over: source-block-major for phase 1 (a u16 source index per
// slot, and per chunk its place in the other order) and bin-major for phase 2
// (a u16 row index per slot, and the term array). Phase 1 (k_bn_scatter)
// copies a source block's contributions into LDS, forms the terms of a chunk
// (contribution, times its weight) per thread and stores them at the chunk's
// bin-major place; consecutive chunks of a tile are consecutive in both
// orders, so the stores are runs. Phase 2 (k_bn_gather) takes one bin and
// streams its chunks: a thread adds a chunk's terms row by row in registers
// (a tile lists its rows in ascending order, so a hub row's terms come in
// long runs) and adds each run into the row's accumulator in LDS; a wave whose
// 64 chunks all hold one row adds it once. Then the bin's rows get the pull
// sweep's epilogue.
//
// Row sums are exact (fixed_sum.hpp): every term is a fixed-point integer in
// units of 2^-100 (terms are at most 1 in magnitude: a contribution is
// rank / out-degree or w * rank / out-strength with w <= out-strength, a
// correction is below 1, and a row sums to at most 1), kept as a 40-bit low
// part and a signed high part that integer adds accumulate. Integer sums do
// not depend on the order of the adds, so each row gets the exact sum of its
// terms (each truncated below 2^-100), correctly rounded to f64 — bit-identical
// run to run, whichever wave adds first. A workgroup takes a whole bin, so no
// partial sums cross workgroups (a cross-workgroup merge would need device-
// scope fences, which write back and invalidate L2).
//
// HBM traffic per f64 sweep: 2 (source index) + 8 (term written) + 8 (term
// read) + 2 (row index) bytes per slot, half a byte of chunk places, plus the
// node terms; f32 sweeps write and read 4-byte terms.
#pragma once

#include "common.hpp"
#include "fixed_sum.hpp"
#include "topology.hpp"

namespace ursa {

constexpr int BN_SBITS = 13;
constexpr uint32_t BN_SSPAN = 1u << BN_SBITS;  // sources per block (their contributions fill LDS)
constexpr uint32_t BN_RSPAN = 4096;  // rows per bin at most (16-byte accumulators: 64 KB of LDS)
constexpr uint32_t BN_CH = 8;                  // slots per chunk; tiles are padded to whole chunks
constexpr uint32_t BN_PAD = 0xFFFFu;           // row index of a padding slot
constexpr int BN_P1_BLOCK = 512;
constexpr int BN_P2_BLOCK = 512;
constexpr int BN_EP = BN_RSPAN / BN_P2_BLOCK;  // epilogue rows per thread
constexpr uint32_t BN_T1 = 1u << 17;         // slots per phase-1 item at most
constexpr uint32_t BN_T2 = 1u << 18;         // in-slots per bin at most, unless one row has more
constexpr uint64_t BN_MAX_ROW = 1ull << 24;  // in-degree limit of the f64 binned sweep (low parts of a row stay below 2^64)
constexpr uint64_t BN_MAX_TILES = 1ull << 26;
constexpr uint64_t BN_MIN_BLOCKS = 32;       // automatic choice: binned from 32 source blocks up
static_assert(BN_RSPAN % (4 * BN_P2_BLOCK) == 0, "epilogue rows per thread: groups of four");

// Lane-to-slot mapping of both phases: a wave takes a window of BN_WIN
// consecutive slots, and in step i (0..3) lane L handles the slot pair
// 128 i + 2 L, 128 i + 2 L + 1 of the window, so every load and store
// instruction of the wave covers one contiguous span.
constexpr uint32_t BN_WIN = 512;

// 8 consecutive values from a 16-byte aligned address in 16-byte loads. A
// lane's pieces are 32 or 64 bytes apart, so each load instruction of a wave
// covers its lines only partly and the following ones complete them. For f64
// (four instructions per line set) plain loads keep the lines in L2 in
// between, where streaming loads let them go and read them again (30 % more
// bytes, 1.6x the time at R-MAT scale 24); f32 (two instructions) streams.
template <class T>
__device__ __forceinline__ void load8(const T* p, T (&v)[8]) {
    if constexpr (sizeof(T) == 8) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2* q = reinterpret_cast<const d2*>(p);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const d2 t = q[i];
            v[2 * i] = t.x;
            v[2 * i + 1] = t.y;
        }
    } else {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4* q = reinterpret_cast<const f4*>(p);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const f4 t = __builtin_nontemporal_load(q + i);
            v[4 * i] = t.x;
            v[4 * i + 1] = t.y;
            v[4 * i + 2] = t.z;
            v[4 * i + 3] = t.w;
        }
    }
}

// G consecutive values of T (G * sizeof(T) = 4, 8 or 16 bytes) at an aligned address.
template <class T, int G>
struct Vec {
    T v[G];
};
template <class T, int G>
__device__ __forceinline__ Vec<T, G> load_vec(const T* p) {
    typedef T V __attribute__((ext_vector_type(G)));
    const V x = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
    Vec<T, G> r;
#pragma unroll
    for (int i = 0; i < G; ++i) r.v[i] = x[i];
    return r;
}
template <class T, int G>
__device__ __forceinline__ void store_vec(T* p, const Vec<T, G>& a) {
    typedef T V __attribute__((ext_vector_type(G)));
    V x;
#pragma unroll
    for (int i = 0; i < G; ++i) x[i] = a.v[i];
    *reinterpret_cast<V*>(p) = x;  // plain store: the wave's stores to a tile's run merge in L2
}

typedef uint32_t bn_u32x4 __attribute__((ext_vector_type(4)));

// Row accumulators of phase 2. f64 sweeps (PARTS = 2) keep the exact
// fixed-point sum of fixed_sum.hpp: a low part (units of 2^-100) and a high
// part (units of 2^-60) per row. f32 sweeps (PARTS = 1) keep one signed
// 64-bit integer per row in units of 2^-K, K chosen per sweep so that no row
// sum can reach 2^60 units (K = 60 - ceil(log2 B) for a bound B on the L1 norm
// of the terms); each term is floor(|x| * 2^K), sign applied, computed with
// f32 operations. Both are integer sums, so neither depends on the order of
// the additions.
template <class CT>
struct BnAcc;

template <>
struct BnAcc<double> {
    static constexpr int PARTS = 2;
    unsigned long long lo = 0;
    long long hi = 0;
    __device__ __forceinline__ void add(double x, float) {
        unsigned long long l;
        long long h;
        fx_split(x, l, h);
        lo += l;
        hi += h;
    }
    __device__ __forceinline__ void reset() {
        lo = 0;
        hi = 0;
    }
    __device__ __forceinline__ void flush(uint32_t r, unsigned long long* s) const {
        atomicAdd(&s[r], lo);
        atomicAdd(&s[BN_RSPAN + r], (unsigned long long)hi);
    }
    // this += the value of lane (lane - off), for lanes at or above off when take.
    __device__ __forceinline__ void add_up(int off, bool take) {
        const unsigned long long l = __shfl_up(lo, off, 64);
        const long long h = __shfl_up(hi, off, 64);
        if (take) {
            lo += l;
            hi += h;
        }
    }
};

template <>
struct BnAcc<float> {
    static constexpr int PARTS = 1;
    unsigned long long t = 0;
    // scale = 2^(K-32): |x| * scale < 2^28 by the choice of K.
    __device__ __forceinline__ void add(float x, float scale) {
        const float a = fabsf(x) * scale;
        const float f = floorf(a);
        const unsigned long long v =
            ((unsigned long long)(uint32_t)f << 32) | (uint32_t)((a - f) * 4294967296.0f);
        t += x < 0.f ? 0ull - v : v;
    }
    __device__ __forceinline__ void reset() { t = 0; }
    __device__ __forceinline__ void flush(uint32_t r, unsigned long long* s) const { atomicAdd(&s[r], t); }
    __device__ __forceinline__ void add_up(int off, bool take) {
        const unsigned long long v = __shfl_up(t, off, 64);
        if (take) t += v;
    }
};

// Adds the 8 consecutive slots of one lane (rows d, terms x; padding rows
// BN_PAD at the end) into the LDS accumulators of type A (f32 terms may go
// into the exact f64 accumulators). A tile lists its rows in
// ascending order, so a row's terms come in runs: each run is summed in
// registers; runs that end inside the lane are added at once, and the lane's
// last run goes through a segmented scan over the wave keyed by row, so that
// consecutive lanes holding the same row (a hub's long run) add it once, from
// the segment's last lane. Without the scan those lanes would all add to one
// LDS address, which serialises the atomics. Called by all 64 lanes.
template <class CT, class A = CT>
__device__ __forceinline__ void bn_add_lane(const uint32_t (&d)[8], const CT (&x)[8], uint32_t lane, float scale,
                                            unsigned long long* s) {
    BnAcc<A> acc;
    uint32_t cur = BN_PAD;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (d[j] == BN_PAD) break;
        if (d[j] != cur) {
            if (cur != BN_PAD) acc.flush(cur, s);
            acc.reset();
            cur = d[j];
        }
        acc.add((A)x[j], scale);
    }
    const uint32_t next = __shfl_down(cur, 1, 64);
    if (__ballot(lane < 63 && next == cur && cur != BN_PAD) == 0) {
        // No last run continues into the next lane (the usual case off the hub
        // rows): every lane adds its own.
        if (cur != BN_PAD) acc.flush(cur, s);
        return;
    }
    const uint32_t prev = __shfl_up(cur, 1, 64);
    bool start = lane == 0 || prev != cur;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const bool start_up = __shfl_up(start ? 1 : 0, off, 64) != 0;
        const bool take = lane >= (uint32_t)off && !start;
        acc.add_up(off, take);
        start = start || (lane >= (uint32_t)off && start_up);
    }
    if ((lane == 63 || next != cur) && cur != BN_PAD) acc.flush(cur, s);
}

// Phase 1 of one launch: claims items (block c, source-major chunks [q0, q1))
// through ctr[0] until none is left; for each, copies the block's
// contributions contrib[v * cstride] (converted to CT) into s_c and stores
// each slot's term (times its weight) at place[chunk] * BN_CH + (slot % BN_CH).
// A wave takes windows of BN_WIN slots; in each of BN_WIN / (64 G) steps lane
// L handles the G consecutive slots at step * 64 G + G L (G = 16 bytes / term
// size), so every load and store instruction of the wave covers one
// contiguous span. s_c (BN_SSPAN values) and s_it are LDS.
template <class CC, class CT, class W, bool WEIGHTED>
__device__ __forceinline__ void bn_scatter_items(const uint16_t* __restrict__ bsrc,
                                                 const uint32_t* __restrict__ place, const W* __restrict__ w,
                                                 const uint32_t* __restrict__ items, uint32_t nitems, uint64_t n,
                                                 const CC* __restrict__ contrib, uint32_t cstride,
                                                 CT* __restrict__ term, uint32_t* __restrict__ ctr, CT* s_c,
                                                 uint32_t* s_it) {
    constexpr int WAVES = BN_P1_BLOCK / 64;
    constexpr int G = 16 / sizeof(CT);
    constexpr int STEPS = BN_WIN / (64 * G);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) *s_it = atomicAdd(&ctr[0], 1u);
        __syncthreads();
        const uint32_t it = *s_it;
        if (it >= nitems) break;
        const uint32_t c = items[4 * (uint64_t)it];
        const uint64_t k0 = (uint64_t)items[4 * (uint64_t)it + 1] * BN_CH, k1 = (uint64_t)items[4 * (uint64_t)it + 2] * BN_CH;
        const uint64_t s0 = (uint64_t)c << BN_SBITS;
        const uint32_t cnt = (uint32_t)(n - s0 < BN_SSPAN ? n - s0 : BN_SSPAN);
        for (uint32_t i = threadIdx.x; i < cnt; i += BN_P1_BLOCK) s_c[i] = (CT)contrib[(s0 + i) * cstride];
        __syncthreads();
        for (uint64_t W0 = k0 + (uint64_t)wave * BN_WIN; W0 < k1; W0 += (uint64_t)WAVES * BN_WIN) {
            // Loads without branches (slots past the item read its first slot and
            // store nothing), so that all of them are issued before the first wait.
            Vec<uint16_t, G> sv[STEPS];
            Vec<W, G> wv[STEPS];
            uint32_t pl[STEPS];
#pragma unroll
            for (int i = 0; i < STEPS; ++i) {
                const uint64_t sl = W0 + (uint64_t)i * 64 * G + (uint64_t)G * lane;
                const uint64_t sa = sl < k1 ? sl : k0;
                sv[i] = load_vec<uint16_t, G>(bsrc + sa);
                pl[i] = place[sa / BN_CH];
                if constexpr (WEIGHTED) wv[i] = load_vec<W, G>(w + sa);
            }
#pragma unroll
            for (int i = 0; i < STEPS; ++i) {
                const uint64_t sl = W0 + (uint64_t)i * 64 * G + (uint64_t)G * lane;
                if (sl < k1) {
                    Vec<CT, G> v;
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        v.v[g] = s_c[sv[i].v[g]];
                        if constexpr (WEIGHTED) v.v[g] = (CT)wv[i].v[g] * v.v[g];
                    }
                    store_vec<CT, G>(term + (uint64_t)pl[i] * BN_CH + (sl % BN_CH), v);
                }
            }
        }
        __syncthreads();
    }
}

// Phase 2's accumulation of one item's bin-major slots [k0, k1) into the
// bin's LDS accumulators (type A: BnAcc<A>). A wave takes windows of BN_WIN
// slots, BN_P2_BLOCK / 64 windows apart; lane L takes the chunk at window
// offset 8 L (row indexes in one 16-byte load, terms in 16-byte loads).
// Chunks lie inside one tile and windows start on chunk boundaries, so a
// lane's 8 slots are all before the item's end or all after it; a lane past
// the end loads the item's first chunk and treats it as padding, so that no
// load sits behind a branch. NW windows are loaded before any is added.
template <class CT, class A>
__device__ __forceinline__ void bn_gather_windows(const uint16_t* __restrict__ bdst, const CT* __restrict__ term,
                                                  uint64_t k0, uint64_t k1, uint32_t wave, uint32_t lane, float scale,
                                                  unsigned long long* s_acc) {
    constexpr int WAVES = BN_P2_BLOCK / 64;
    constexpr int NW = sizeof(CT) == 8 ? 2 : 4;
    constexpr uint64_t STRIDE = (uint64_t)WAVES * BN_WIN;
    for (uint64_t W = k0 + (uint64_t)wave * BN_WIN; W < k1; W += NW * STRIDE) {
        uint32_t d[NW][8];
        CT x[NW][8];
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint64_t sl = W + w * STRIDE + 8 * lane;
            const uint64_t sa = sl < k1 ? sl : k0;
            const bn_u32x4 dv = __builtin_nontemporal_load(reinterpret_cast<const bn_u32x4*>(bdst + sa));
            d[w][0] = dv.x & 0xFFFFu; d[w][1] = dv.x >> 16;
            d[w][2] = dv.y & 0xFFFFu; d[w][3] = dv.y >> 16;
            d[w][4] = dv.z & 0xFFFFu; d[w][5] = dv.z >> 16;
            d[w][6] = dv.w & 0xFFFFu; d[w][7] = dv.w >> 16;
            load8(term + sa, x[w]);
        }
#pragma unroll
        for (int w = 0; w < NW; ++w)
            if (W + w * STRIDE + 8 * lane >= k1) d[w][0] = BN_PAD;
#pragma unroll
        for (int w = 0; w < NW; ++w) bn_add_lane<CT, A>(d[w], x[w], lane, scale, s_acc);
    }
}

// Compute units of the current device (grid sizes of the work-claiming kernels).
inline unsigned cu_count() {
    static int cus = [] {
        int dev = 0, v = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        return v > 0 ? v : 256;
    }();
    return (unsigned)cus;
}

// Largest in-degree of PageRank's renumbered in-CSR (cached on the topology).
uint64_t pr_max_in_degree(ursa_hip_topo* t, hipStream_t s);
// Whether the binned layout fits t: at most BN_MAX_TILES tiles and every
// in-degree below BN_MAX_ROW (needs PageRank's working copy).
bool bn_fits(ursa_hip_topo* t, hipStream_t s);
// Builds the binned layout of PageRank's working copy (cached on the topology).
void bn_prepare(ursa_hip_topo* t, hipStream_t s);

}  // namespace ursa
