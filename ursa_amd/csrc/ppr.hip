// ppr.hip — personalized PageRank, up to 16 queries on one graph per solve.
//
// ursa-core has no personalized PageRank; the definition is the global
// iteration of pagerank.rs:63-124 with the teleport and dangling mass sent to
// the query's seed set (networkx `pagerank(personalization=...)` semantics):
//   x0[u] = 1/n
//   next[u] = d * sum_{v in in(u)} x[v] / od[v] + ((1 - d) + d * dm) * p[u]
//   dm = sum of x over dangling vertices, p = uniform over the distinct seeds,
// stopping after the first iteration whose L1 delta is below tol. A query
// with no seed in range returns x0 after zero iterations.
//
// Layout: the sweeps run over PageRank's working copy of the in-CSR (vertices
// renumbered with the most gathered sources first, each row's sources in
// ascending new id; pagerank.hip), so the hot contributions share cache lines.
// The queries of one solve are lanes q of a vertex-major state, value of
// (renumbered) vertex v for lane q at v*KP + q (KP = lane count rounded up to
// a power of two), so one in-edge gathers KP adjacent values for all queries at
// once. Seeds come in and results go out in the original numbering.
//
// Precision: a lane with tol < 2e-6 and max_iter > 16 first runs f32 sweeps
// with compensated (Neumaier) row sums until its delta is below 2e-6 or stops
// falling, then f64 sweeps refine the same iterate until delta < tol; the
// first f64 sweep reads the f32 iterate widened. The f32 phase always leaves
// at least one f64 sweep before max_iter, so every lane returns an f64
// iterate. Any other lane (a coarse tol, or a max_iter too small for the f32
// phase to pay) runs f64 sweeps from the uniform start, the iteration above
// term for term, so its iteration count and values follow the f64 solve.
// Lanes switch and finish independently.
//
// Work split: a group of PPR_G = 8 adjacent threads sums one row, thread g
// taking in-slots g, g+8, ... (so the group reads the slot list 32 bytes at a
// time) and the group's eight partial sums are combined in a fixed butterfly.
// A row with more than PPR_HEAVY in-edges is cut into segments fixed by its
// degree alone; a group sums each segment and the segment partials are added
// in segment order. A workgroup sweeps PPR_ROWS rows and reduces the L1 delta
// and dangling mass of those rows in a fixed tree; the host-side reduction adds
// the per-workgroup values in workgroup order.
//
// Binned schedule: on graphs whose contributions do not stay in cache (from
// 32 blocks of 8192 vertices up, no row above PPR_HEAVY in-edges, i.e. no
// hubs whose sources share lines, and a mean in-degree of at least
// PPR_BN_MIN_DEGREE), each lane's sweep is PageRank's
// propagation-blocked pair of passes over the same working copy (binned.hpp):
// phase 1 streams the lane's contributions into terms in bin-major order,
// phase 2 adds each bin's terms into exact fixed-point row sums in LDS (f32
// phase: units of 2^-PPR_BN_K; f64 phases: fixed_sum.hpp) and applies the
// epilogue below. A lane's delta and dangling mass are reduced per bin.
// The schedule is a property of the graph (and of ursa_hip_ppr_schedule),
// never of the batch, and each lane runs its own passes.
//
// Determinism: a lane's arithmetic does not depend on the other lanes (every
// split above is fixed by the graph), so a query gives the same bits alone, in
// any batch and in any lane.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "binned.hpp"
#include "ppr.hpp"
#include "topology.hpp"

namespace ursa {

namespace {

constexpr int PB = 256;
constexpr int PPR_G = 8;                     // threads per row
constexpr int PPR_GROUPS = PB / PPR_G;       // rows per workgroup pass
constexpr uint32_t PPR_ROWS = 128;           // rows per workgroup (one partial each)
constexpr uint32_t PPR_HEAVY = 256;
constexpr uint32_t PPR_MAXSEG = 1024;
constexpr int PPR_BATCH = 4;                 // slots read together per thread (2 at four lanes)
constexpr double PPR_F32_TOL = 2e-6;         // the f32 phase ends once delta is below this
constexpr uint32_t PPR_F32_MIN_ITER = 16;    // lanes with max_iter <= this run f64 only
// Binned f32 phase: terms and row sums of a lane are below 4 (its iterate
// sums to 1), so rows sum exactly in 64-bit integers of 2^-58 units.
constexpr int PPR_BN_K = 58;
// Automatic binned schedule: mean in-degree at least this (below it the
// per-row work of the binned passes outweighs the gathers they replace;
// Erdos-Renyi 1.4 M vertices: edge factor 5 pull 84 us vs binned 74 us per
// sweep but slower per solve, edge factor 15: 266 vs 123 us).
constexpr uint64_t PPR_BN_MIN_DEGREE = 8;

// Sweep schedule (ursa_hip_ppr_schedule): 0 automatic, 1 pull, 2 binned.
std::atomic<int> g_ppr_schedule{0};

struct PprLane {
    double d, tol, inv_s, pb, prev, delta;
    uint32_t max_iter, iters, f32_iters;
    int32_t phase;    // 0 f32 sweeps, 1 first f64 sweep, 2 f64 sweeps, 3 done
    int32_t pad[2];
};

struct PprCtl {
    uint32_t active;
};

__device__ __forceinline__ void neu_add(float& s, float& c, float x) {
    const float t = s + x;
    if (fabsf(s) >= fabsf(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

template <int KP>
__device__ __forceinline__ void load_phases(const PprLane* __restrict__ lanes, int* ph) {
#pragma unroll
    for (int q = 0; q < KP; ++q) ph[q] = lanes[q].phase;
}

// KP adjacent values (KP * sizeof(T) bytes, aligned to that size) with
// 16-byte loads where they fit.
template <int KP, class T>
__device__ __forceinline__ void load_lanes(const T* __restrict__ p, T* x) {
    constexpr int PER = 16 / (int)sizeof(T);  // values per 16-byte load
    if constexpr (KP >= PER) {
        using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
#pragma unroll
        for (int i = 0; i < KP / PER; ++i) {
            const V w = reinterpret_cast<const V*>(p)[i];
            if constexpr (sizeof(T) == 4) {
                x[4 * i] = w.x;
                x[4 * i + 1] = w.y;
                x[4 * i + 2] = w.z;
                x[4 * i + 3] = w.w;
            } else {
                x[2 * i] = w.x;
                x[2 * i + 1] = w.y;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < KP; ++q) x[q] = p[q];
    }
}

// This thread's share (slots b+g, b+g+8, ...) of the row sum over in-slots
// [b, e), combined across the group's eight threads; the result is valid in
// the group's thread g = 0.
template <int KP>
__device__ __forceinline__ void ppr_row_sum(uint32_t b, uint32_t e, uint32_t g, const uint32_t* __restrict__ tgt,
                                            const int* ph, const float* __restrict__ c32,
                                            const double* __restrict__ c64, float* s, float* c, double* a) {
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        s[q] = 0.f;
        c[q] = 0.f;
        a[q] = 0.0;
    }
    // Which value arrays any lane reads this sweep (phases are uniform).
    bool any32 = false, any64 = false;
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        any32 |= ph[q] == 0 || ph[q] == 1;
        any64 |= ph[q] == 2;
    }
    auto add = [&](const float* x32, const double* x64) {
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            if (ph[q] == 0)
                neu_add(s[q], c[q], x32[q]);
            else if (ph[q] == 1)
                a[q] += (double)x32[q];
            else if (ph[q] == 2)
                a[q] += x64[q];
        }
    };
    if constexpr (KP <= 4) {
        // Few lanes: this thread's next NB slots are read together (the
        // columns, then their values) and added in slot order as before.
        constexpr int NB = KP <= 2 ? PPR_BATCH : 2;
        for (uint32_t k0 = b + g; k0 < e; k0 += NB * PPR_G) {
            uint64_t vv[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const uint32_t k = k0 + j * PPR_G;
                vv[j] = k < e ? tgt[k] : 0u;
            }
            float x32[NB][KP];
            double x64[NB][KP];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (k0 + j * PPR_G < e) {
                    if (any32) load_lanes<KP>(c32 + vv[j] * KP, x32[j]);
                    if (any64) load_lanes<KP>(c64 + vv[j] * KP, x64[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if (k0 + j * PPR_G < e) add(x32[j], x64[j]);
        }
    } else {
#pragma unroll 2
        for (uint32_t k = b + g; k < e; k += PPR_G) {
            const uint64_t v = tgt[k];
            // The KP values of vertex v are adjacent: read them with wide loads,
            // then add lane by lane (the per-lane arithmetic does not change).
            float x32[KP];
            double x64[KP];
            if (any32) load_lanes<KP>(c32 + v * KP, x32);
            if (any64) load_lanes<KP>(c64 + v * KP, x64);
            add(x32, x64);
        }
    }
    // Butterfly over the group (lanes differing in the low three bits).
#pragma unroll
    for (int off = PPR_G / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            if (ph[q] == 0) {
                const float s2 = __shfl_xor(s[q], off, 64), c2 = __shfl_xor(c[q], off, 64);
                c[q] += c2;
                neu_add(s[q], c[q], s2);
            } else if (ph[q] != 3) {
                a[q] += __shfl_xor(a[q], off, 64);
            }
        }
    }
}

// New value of row u for every running lane; adds |next - cur| to dl and,
// for a dangling row, next to dm.
template <int KP>
__device__ __forceinline__ void ppr_epilogue(uint64_t u, uint32_t seedb, uint32_t odu,
                                             const PprLane* __restrict__ lanes, const int* ph, const float* s,
                                             const float* c, const double* a, float* __restrict__ r32,
                                             float* __restrict__ c32n, double* __restrict__ r64,
                                             double* __restrict__ c64n, double* dl, double* dm) {
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        if (ph[q] == 3) continue;
        const bool seed = (seedb >> q) & 1u;
        const uint64_t i = u * KP + q;
        if (ph[q] == 0) {
            const float acc = s[q] + c[q];
            const float tele = seed ? (float)(lanes[q].pb * lanes[q].inv_s) : 0.f;
            const float nx = fmaf((float)lanes[q].d, acc, tele);
            const float cur = r32[i];
            r32[i] = nx;
            dl[q] += fabs((double)nx - (double)cur);
            if (odu)
                c32n[i] = nx / (float)odu;
            else
                dm[q] += (double)nx;
        } else {
            const double tele = seed ? lanes[q].pb * lanes[q].inv_s : 0.0;
            const double nx = fma(lanes[q].d, a[q], tele);
            const double cur = ph[q] == 1 ? (double)r32[i] : r64[i];
            r64[i] = nx;
            dl[q] += fabs(nx - cur);
            if (odu)
                c64n[i] = nx / (double)odu;
            else
                dm[q] += nx;
        }
    }
}

// Fixed-order block reduction of (delta, dangling) per lane into out[KP*2].
template <int KP>
__device__ __forceinline__ void ppr_block_reduce(double* dl, double* dm, double* __restrict__ out) {
    __shared__ double sh[PB / 64][KP * 2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        double x = dl[q], y = dm[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            x += __shfl_down(x, off, 64);
            y += __shfl_down(y, off, 64);
        }
        if (lane == 0) {
            sh[w][2 * q] = x;
            sh[w][2 * q + 1] = y;
        }
    }
    __syncthreads();
    if (threadIdx.x < KP * 2) {
        double acc = sh[0][threadIdx.x];
#pragma unroll
        for (int j = 1; j < PB / 64; ++j) acc += sh[j][threadIdx.x];
        out[threadIdx.x] = acc;
    }
}

// One sweep over the light rows (blocks [0, light_blocks), PPR_ROWS rows
// each) and the heavy segments (PPR_GROUPS per block after them); bid is the
// block's index in that range (a whole workgroup of PB threads).
template <int KP>
__device__ __forceinline__ void ppr_rows_block(uint32_t bid, uint64_t n, const uint32_t* __restrict__ off,
                                                 const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ od,
                                                 const uint32_t* __restrict__ seedbits,
                                                 const PprLane* __restrict__ lanes, const float* __restrict__ c32,
                                                 float* __restrict__ c32n, float* __restrict__ r32,
                                                 const double* __restrict__ c64, double* __restrict__ c64n,
                                                 double* __restrict__ r64, double* __restrict__ part,
                                                 uint32_t light_blocks, uint64_t nseg,
                                                 const uint32_t* __restrict__ sbeg,
                                                 const uint32_t* __restrict__ send, double* __restrict__ spart) {
    int ph[KP];
    load_phases<KP>(lanes, ph);
    bool any = false;
#pragma unroll
    for (int q = 0; q < KP; ++q) any |= ph[q] != 3;
    if (!any) return;  // every query has finished (a queued sweep after the last one)
    float s[KP], c[KP];
    double a[KP];
    const uint32_t g = threadIdx.x % PPR_G, grp = threadIdx.x / PPR_G;
    if (bid >= light_blocks) {
        const uint64_t sg = (uint64_t)(bid - light_blocks) * PPR_GROUPS + grp;
        if (sg >= nseg) return;  // whole groups leave together
        ppr_row_sum<KP>(sbeg[sg], send[sg], g, tgt, ph, c32, c64, s, c, a);
        if (g == 0) {
#pragma unroll
            for (int q = 0; q < KP; ++q)
                if (ph[q] != 3) spart[sg * KP + q] = ph[q] == 0 ? (double)s[q] + (double)c[q] : a[q];
        }
        return;
    }
    double dl[KP], dm[KP];
#pragma unroll
    for (int q = 0; q < KP; ++q) dl[q] = dm[q] = 0.0;
    const uint64_t row0 = (uint64_t)bid * PPR_ROWS;
    for (uint32_t r = grp; r < PPR_ROWS; r += PPR_GROUPS) {
        const uint64_t u = row0 + r;
        if (u >= n) break;
        const uint32_t b = off[u], e = off[u + 1];
        if (e - b > PPR_HEAVY) continue;
        ppr_row_sum<KP>(b, e, g, tgt, ph, c32, c64, s, c, a);
        if (g == 0) ppr_epilogue<KP>(u, seedbits[u], od[u], lanes, ph, s, c, a, r32, c32n, r64, c64n, dl, dm);
    }
    ppr_block_reduce<KP>(dl, dm, part + (uint64_t)bid * KP * 2);
}

template <int KP>
__global__ __launch_bounds__(PB) void k_ppr_rows(uint64_t n, const uint32_t* __restrict__ off,
                                                 const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ od,
                                                 const uint32_t* __restrict__ seedbits,
                                                 const PprLane* __restrict__ lanes, const float* __restrict__ c32,
                                                 float* __restrict__ c32n, float* __restrict__ r32,
                                                 const double* __restrict__ c64, double* __restrict__ c64n,
                                                 double* __restrict__ r64, double* __restrict__ part,
                                                 uint32_t light_blocks, uint64_t nseg,
                                                 const uint32_t* __restrict__ sbeg,
                                                 const uint32_t* __restrict__ send, double* __restrict__ spart) {
    ppr_rows_block<KP>(blockIdx.x, n, off, tgt, od, seedbits, lanes, c32, c32n, r32, c64, c64n, r64, part,
                       light_blocks, nseg, sbeg, send, spart);
}

// Heavy rows: add the row's segment partials in order, then the epilogue
// (block bid of the heavy-row range).
template <int KP>
__device__ __forceinline__ void ppr_heavy_block(uint32_t bid, uint64_t H, const uint32_t* __restrict__ hrow,
                                                  const uint32_t* __restrict__ hsoff,
                                                  const double* __restrict__ spart, const uint32_t* __restrict__ od,
                                                  const uint32_t* __restrict__ seedbits,
                                                  const PprLane* __restrict__ lanes, float* __restrict__ c32n,
                                                  float* __restrict__ r32, double* __restrict__ c64n,
                                                  double* __restrict__ r64, double* __restrict__ part) {
    int ph[KP];
    load_phases<KP>(lanes, ph);
    bool any = false;
#pragma unroll
    for (int q = 0; q < KP; ++q) any |= ph[q] != 3;
    if (!any) return;
    double dl[KP], dm[KP];
#pragma unroll
    for (int q = 0; q < KP; ++q) dl[q] = dm[q] = 0.0;
    const uint64_t h = (uint64_t)bid * PB + threadIdx.x;
    if (h < H) {
        const uint64_t u = hrow[h];
        const uint32_t sb = hsoff[h], se = hsoff[h + 1];
        float s[KP], c[KP];
        double a[KP];
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            // Segment order; the loads of 16 segments are issued before their adds.
            double acc = 0.0;
            if (ph[q] != 3) {
                for (uint32_t j0 = sb; j0 < se; j0 += 16) {
                    double x[16];
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const uint32_t j = min(j0 + i, se - 1);
                        x[i] = spart[(uint64_t)j * KP + q];
                    }
#pragma unroll
                    for (int i = 0; i < 16; ++i)
                        if (j0 + i < se) acc += x[i];
                }
            }
            a[q] = acc;
            s[q] = (float)acc;
            c[q] = 0.f;
        }
        ppr_epilogue<KP>(u, seedbits[u], od[u], lanes, ph, s, c, a, r32, c32n, r64, c64n, dl, dm);
    }
    ppr_block_reduce<KP>(dl, dm, part + (uint64_t)bid * KP * 2);
}

template <int KP>
__global__ __launch_bounds__(PB) void k_ppr_heavy(uint64_t H, const uint32_t* __restrict__ hrow,
                                                  const uint32_t* __restrict__ hsoff,
                                                  const double* __restrict__ spart, const uint32_t* __restrict__ od,
                                                  const uint32_t* __restrict__ seedbits,
                                                  const PprLane* __restrict__ lanes, float* __restrict__ c32n,
                                                  float* __restrict__ r32, double* __restrict__ c64n,
                                                  double* __restrict__ r64, double* __restrict__ part) {
    ppr_heavy_block<KP>(blockIdx.x, H, hrow, hsoff, spart, od, seedbits, lanes, c32n, r32, c64n, r64, part);
}

// Sums the block partials per lane in a fixed order and advances each lane
// (one workgroup of PB threads).
template <int KP>
__device__ __forceinline__ void ppr_reduce_block(const double* __restrict__ part, uint32_t nparts,
                                                 PprLane* __restrict__ lanes, PprCtl* __restrict__ ctl) {
    __shared__ double sh[PB];
    __shared__ double res[KP * 2];
    // Thread t adds the partials b = t, t + PB, ... in that order for every
    // value (q, w); the loads of U partials are issued before their adds.
    constexpr int V = KP * 2, U = V >= 16 ? 1 : 16 / V;
    double acc[V];
#pragma unroll
    for (int j = 0; j < V; ++j) acc[j] = 0.0;
    for (uint32_t b0 = threadIdx.x; b0 < nparts; b0 += U * PB) {
        double x[U][V];
#pragma unroll
        for (int i = 0; i < U; ++i) {
            const uint32_t b = min(b0 + i * PB, nparts - 1);
#pragma unroll
            for (int j = 0; j < V; ++j) x[i][j] = part[(uint64_t)b * V + j];
        }
#pragma unroll
        for (int i = 0; i < U; ++i)
            if (b0 + i * PB < nparts) {
#pragma unroll
                for (int j = 0; j < V; ++j) acc[j] += x[i][j];
            }
    }
    for (int q = 0; q < KP; ++q) {
        if (lanes[q].phase == 3) continue;
        for (int w = 0; w < 2; ++w) {
            sh[threadIdx.x] = acc[q * 2 + w];
            __syncthreads();
            for (int st = PB / 2; st > 0; st >>= 1) {
                if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
                __syncthreads();
            }
            if (threadIdx.x == 0) res[q * 2 + w] = sh[0];
            __syncthreads();
        }
    }
    if (threadIdx.x != 0) return;
    uint32_t active = 0;
    for (int q = 0; q < KP; ++q) {
        PprLane L = lanes[q];
        if (L.phase == 3) continue;
        const double delta = res[q * 2], dmass = res[q * 2 + 1];
        L.iters += 1;
        L.delta = delta;
        L.pb = (1.0 - L.d) + L.d * dmass;
        if (L.phase == 0) {
            L.f32_iters += 1;
            const bool stalled = L.f32_iters >= 3 && delta > 0.7 * L.prev;
            L.prev = delta;
            // The last sweep max_iter allows is always an f64 one.
            if (delta < PPR_F32_TOL || stalled || L.iters + 1 >= L.max_iter) L.phase = 1;
        } else {
            if (delta < L.tol || L.iters >= L.max_iter) {
                L.phase = 3;
            } else {
                L.phase = 2;
            }
        }
        lanes[q] = L;
        if (L.phase != 3) ++active;
    }
    ctl->active = active;
}

template <int KP>
__global__ __launch_bounds__(PB) void k_ppr_reduce(const double* __restrict__ part, uint32_t nparts,
                                                   PprLane* __restrict__ lanes, PprCtl* __restrict__ ctl) {
    ppr_reduce_block<KP>(part, nparts, lanes, ctl);
}

// Binned sweep of lane q, phase 1: the lane's contributions (c32 in the f32
// phase and the first f64 sweep, c64 after) become terms in bin-major order.
// Also clears phase 2's claim counter ctr[1].
template <int KP>
__global__ __launch_bounds__(BN_P1_BLOCK) void k_ppr_bn_scatter(int q, const PprLane* __restrict__ lanes,
                                                                const uint16_t* __restrict__ bsrc,
                                                                const uint32_t* __restrict__ place,
                                                                const uint32_t* __restrict__ items, uint32_t nitems,
                                                                uint64_t n, const float* __restrict__ c32,
                                                                const double* __restrict__ c64, void* term,
                                                                uint32_t* __restrict__ ctr) {
    __shared__ double s_c[BN_SSPAN];
    __shared__ uint32_t s_it;
    const int ph = lanes[q].phase;
    if (ph == 3) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[1] = 0;
    if (ph == 2)
        bn_scatter_items<double, double, float, false>(bsrc, place, nullptr, items, nitems, n, c64 + q, (uint32_t)KP,
                                                       static_cast<double*>(term), ctr, s_c, &s_it);
    else
        bn_scatter_items<float, float, float, false>(bsrc, place, nullptr, items, nitems, n, c32 + q, (uint32_t)KP,
                                                     static_cast<float*>(term), ctr,
                                                     reinterpret_cast<float*>(s_c), &s_it);
}

// Fixed-order sum over a BN_P2_BLOCK workgroup; valid in thread 0.
__device__ __forceinline__ double ppr_bn_block_sum(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < BN_P2_BLOCK / 64; ++w) tot += lds[w];
    __syncthreads();
    return tot;
}

// Binned sweep of lane q, phase 2: one bin per work item (claimed through
// ctr[1]; also clears phase 1's counter ctr[0]); exact row sums, the lane's
// epilogue (as ppr_epilogue) and the bin's (delta, dangling) partial at
// part[(bin * KP + q) * 2].
template <int KP>
__global__ __launch_bounds__(BN_P2_BLOCK, 4) void k_ppr_bn_gather(
    int q, const PprLane* __restrict__ lanes, const uint16_t* __restrict__ bdst, const void* term,
    const uint32_t* __restrict__ items, uint32_t nitems, const uint32_t* __restrict__ od,
    const uint32_t* __restrict__ seedbits, float* __restrict__ r32, float* __restrict__ c32n,
    double* __restrict__ r64, double* __restrict__ c64n, double* __restrict__ part, uint32_t* __restrict__ ctr) {
    __shared__ unsigned long long s_acc[2 * BN_RSPAN];
    __shared__ double s_red[BN_P2_BLOCK / 64];
    __shared__ uint32_t s_it;
    const int ph = lanes[q].phase;
    if (ph == 3) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[0] = 0;
    const double d = lanes[q].d, tele = lanes[q].pb * lanes[q].inv_s;
    const float scale = __builtin_ldexpf(1.0f, PPR_BN_K - 32);
    const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t parts = ph == 0 ? 1u : 2u;
    for (;;) {
        if (threadIdx.x == 0) s_it = atomicAdd(&ctr[1], 1u);
        for (uint32_t i = threadIdx.x; i < parts * BN_RSPAN; i += BN_P2_BLOCK) s_acc[i] = 0ull;
        __syncthreads();
        const uint32_t it = s_it;
        if (it >= nitems) break;
        const uint32_t* I = items + 8 * (uint64_t)it;
        const uint32_t b = I[0], rows = I[4];
        const uint64_t k0 = (uint64_t)I[1] * BN_CH, k1 = (uint64_t)I[2] * BN_CH, r0 = I[3];
        if (ph == 0)
            bn_gather_windows<float, float>(bdst, static_cast<const float*>(term), k0, k1, wave, lane, scale, s_acc);
        else if (ph == 1)
            bn_gather_windows<float, double>(bdst, static_cast<const float*>(term), k0, k1, wave, lane, scale, s_acc);
        else
            bn_gather_windows<double, double>(bdst, static_cast<const double*>(term), k0, k1, wave, lane, scale,
                                              s_acc);
        __syncthreads();
        double dl = 0.0, dm = 0.0;
        for (uint32_t i = threadIdx.x; i < rows; i += BN_P2_BLOCK) {
            const uint64_t u = r0 + i, x = u * KP + q;
            const bool seed = (seedbits[u] >> q) & 1u;
            const uint32_t odu = od[u];
            if (ph == 0) {
                const float acc = (float)__builtin_ldexp((double)(long long)s_acc[i], -PPR_BN_K);
                const float nx = fmaf((float)d, acc, seed ? (float)tele : 0.f);
                const float cur = r32[x];
                r32[x] = nx;
                dl += fabs((double)nx - (double)cur);
                if (odu)
                    c32n[x] = nx / (float)odu;
                else
                    dm += (double)nx;
            } else {
                const double acc = fx_value(s_acc[i], (long long)s_acc[BN_RSPAN + i]);
                const double nx = fma(d, acc, seed ? tele : 0.0);
                const double cur = ph == 1 ? (double)r32[x] : r64[x];
                r64[x] = nx;
                dl += fabs(nx - cur);
                if (odu)
                    c64n[x] = nx / (double)odu;
                else
                    dm += nx;
            }
        }
        const double bd = ppr_bn_block_sum(dl, s_red);
        const double bm = ppr_bn_block_sum(dm, s_red);
        if (threadIdx.x == 0) {
            part[((uint64_t)b * KP + q) * 2] = bd;
            part[((uint64_t)b * KP + q) * 2 + 1] = bm;
        }
        __syncthreads();
    }
}

// x0 = 1/n for every lane, with its f32 and f64 contributions (lanes in the
// f32 phase read c32, lanes that start in f64 read c64).
template <int KP>
__device__ __forceinline__ void ppr_init_vertex(uint64_t u, uint64_t n, const uint32_t* __restrict__ od,
                                                float* __restrict__ r32, float* __restrict__ c32,
                                                double* __restrict__ r64, double* __restrict__ c64) {
    const float x32 = (float)(1.0 / (double)n);
    const double x64 = 1.0 / (double)n;
    const uint32_t o = od[u];
#pragma unroll
    for (int q = 0; q < KP; ++q) {
        r32[u * KP + q] = x32;
        r64[u * KP + q] = x64;
        c32[u * KP + q] = o ? x32 / (float)o : 0.f;
        c64[u * KP + q] = o ? x64 / (double)o : 0.0;
    }
}

template <int KP>
__global__ void k_ppr_init(uint64_t n, const uint32_t* __restrict__ od, float* __restrict__ r32,
                           float* __restrict__ c32, double* __restrict__ r64, double* __restrict__ c64) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x)
        ppr_init_vertex<KP>(u, n, od, r32, c32, r64, c64);
}

__global__ void k_ppr_seedbits(const uint32_t* __restrict__ seeds, const uint32_t* __restrict__ lane_of,
                               uint32_t count, const uint32_t* __restrict__ new_of_old, uint32_t* __restrict__ bits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) atomicOr(&bits[new_of_old[seeds[i]]], 1u << lane_of[i]);
}

// Every lane ends on its f64 iterate (the uniform start for a lane that never ran).
template <int KP>
__global__ void k_ppr_out(uint64_t n, int q, const uint32_t* __restrict__ new_of_old, const double* __restrict__ r64,
                          double* __restrict__ out) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x)
        out[u] = r64[(uint64_t)new_of_old[u] * KP + q];
}

// ---- several one-lane solves on different graphs, swept together ----
//
// A cross-graph pass runs P one-lane solves (each on its own graph, on the
// pull schedule) in lockstep: one launch per sweep phase covers every
// problem, a workgroup finding its problem in a prefix table of block
// counts. Every workgroup runs the one-graph kernels' block bodies
// (ppr_rows_block, ppr_heavy_block, ppr_reduce_block) with its index inside
// its own problem's grid, so each problem's arithmetic, and its result, is
// that of its own solve alone.
struct PprProb {
    const uint32_t *off, *tgt, *od, *sbeg, *send, *hrow, *hsoff, *new_of_old;
    uint32_t* seedbits;
    PprLane* lanes;
    PprCtl* ctl;
    float *r32, *c32a, *c32b;
    double *r64, *c64a, *c64b, *part, *spart, *out;
    uint64_t n, nseg, H;
    uint32_t light_blocks, nparts;
};

// Largest p < np with start[p] <= bid (start non-decreasing, start[0] = 0):
// the problem whose block range holds bid.
__device__ __forceinline__ uint32_t ppr_prob_of(const uint32_t* __restrict__ start, uint32_t np, uint32_t bid) {
    uint32_t lo = 0, hi = np;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= bid)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(PB) void k_pprm_rows(const PprProb* __restrict__ P, const uint32_t* __restrict__ start,
                                                  uint32_t np, int parity) {
    const uint32_t p = ppr_prob_of(start, np, blockIdx.x);
    const PprProb& G = P[p];
    const float* c32 = parity ? G.c32b : G.c32a;
    float* c32n = parity ? G.c32a : G.c32b;
    const double* c64 = parity ? G.c64b : G.c64a;
    double* c64n = parity ? G.c64a : G.c64b;
    ppr_rows_block<1>(blockIdx.x - start[p], G.n, G.off, G.tgt, G.od, G.seedbits, G.lanes, c32, c32n, G.r32, c64,
                      c64n, G.r64, G.part, G.light_blocks, G.nseg, G.sbeg, G.send, G.spart);
}

__global__ __launch_bounds__(PB) void k_pprm_heavy(const PprProb* __restrict__ P, const uint32_t* __restrict__ start,
                                                   uint32_t np, int parity) {
    const uint32_t p = ppr_prob_of(start, np, blockIdx.x);
    const PprProb& G = P[p];
    float* c32n = parity ? G.c32a : G.c32b;
    double* c64n = parity ? G.c64a : G.c64b;
    ppr_heavy_block<1>(blockIdx.x - start[p], G.H, G.hrow, G.hsoff, G.spart, G.od, G.seedbits, G.lanes, c32n, G.r32,
                       c64n, G.r64, G.part + (uint64_t)G.light_blocks * 2);
}

__global__ __launch_bounds__(PB) void k_pprm_reduce(const PprProb* __restrict__ P) {
    const PprProb& G = P[blockIdx.x];
    ppr_reduce_block<1>(G.part, G.nparts, G.lanes, G.ctl);
}

// x0 and cleared seed bits for every problem (block range start[p].. per problem).
__global__ void k_pprm_init(const PprProb* __restrict__ P, const uint32_t* __restrict__ start, uint32_t np) {
    const uint32_t p = ppr_prob_of(start, np, blockIdx.x);
    const PprProb& G = P[p];
    const uint64_t nb = start[p + 1] - start[p];
    for (uint64_t u = (blockIdx.x - start[p]) * (uint64_t)blockDim.x + threadIdx.x; u < G.n; u += nb * blockDim.x) {
        G.seedbits[u] = 0;
        ppr_init_vertex<1>(u, G.n, G.od, G.r32, G.c32a, G.r64, G.c64a);
    }
}

__global__ void k_pprm_seeds(const PprProb* __restrict__ P, const uint32_t* __restrict__ prob,
                             const uint32_t* __restrict__ seed, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const PprProb& G = P[prob[i]];
        atomicOr(&G.seedbits[G.new_of_old[seed[i]]], 1u);
    }
}

__global__ void k_pprm_out(const PprProb* __restrict__ P, const uint32_t* __restrict__ start, uint32_t np) {
    const uint32_t p = ppr_prob_of(start, np, blockIdx.x);
    const PprProb& G = P[p];
    const uint64_t nb = start[p + 1] - start[p];
    for (uint64_t u = (blockIdx.x - start[p]) * (uint64_t)blockDim.x + threadIdx.x; u < G.n; u += nb * blockDim.x)
        G.out[u] = G.r64[G.new_of_old[u]];
}

int lanes_padded(int k) {
    int kp = 1;
    while (kp < k) kp <<= 1;
    return kp;
}

struct PprBufs {
    DevBuf<float> r32, c32a, c32b;
    DevBuf<double> r64, c64a, c64b, part, spart, term;
    DevBuf<uint32_t> seedbits, seeds, lane_of, bctr;
    DevBuf<PprLane> lanes;
    DevBuf<PprCtl> ctl;
    bool binned = false;
    int k = 0;                 // lanes that hold a query
    unsigned g1 = 1, g2 = 1;   // binned grid sizes
};

template <int KP>
void ppr_iterate(ursa_hip_topo* t, PprBufs& B, uint32_t max_iter_all, hipStream_t s) {
    const uint64_t n = t->n;
    const uint32_t light_blocks = (uint32_t)((n + PPR_ROWS - 1) / PPR_ROWS);
    const uint32_t seg_blocks = (uint32_t)((t->ppr_S + PPR_GROUPS - 1) / PPR_GROUPS);
    const uint32_t heavy_blocks = (uint32_t)((t->ppr_H + PB - 1) / PB);
    const uint32_t nparts = B.binned ? (uint32_t)t->bn_bins : light_blocks + heavy_blocks;
    const double edge_bytes = 4.0 + 4.0 * KP;
    uint32_t it = 0;
    CapturedLaunches graph;  // a full batch always starts at an even sweep
    while (it < max_iter_all) {
        const uint32_t end = std::min<uint32_t>(max_iter_all, it + 8);
        const uint32_t it0 = it;
        auto launches = [&] {
            for (uint32_t i = it0; i < end; ++i) {
                const int p = (int)(i & 1);
                const float* c32 = p ? B.c32b.get() : B.c32a.get();
                float* c32n = p ? B.c32a.get() : B.c32b.get();
                const double* c64 = p ? B.c64b.get() : B.c64a.get();
                double* c64n = p ? B.c64a.get() : B.c64b.get();
                if (B.binned) {
                    // Per lane: 2 + 8 (f64 term) + 8 + 2 bytes per slot at most.
                    timed("ppr_binned", (20.0 * (double)t->bn_slots + 28.0 * (double)n) * B.k, s, [&] {
                        for (int q = 0; q < B.k; ++q) {
                            k_ppr_bn_scatter<KP><<<B.g1, BN_P1_BLOCK, 0, s>>>(
                                q, B.lanes.get(), t->bn_src.get(), t->bn_place.get(), t->bn_item1.get(),
                                (uint32_t)t->bn_n1, n, c32, c64, B.term.get(), B.bctr.get());
                            k_ppr_bn_gather<KP><<<B.g2, BN_P2_BLOCK, 0, s>>>(
                                q, B.lanes.get(), t->bn_dst.get(), B.term.get(), t->bn_item2.get(),
                                (uint32_t)t->bn_n2, t->pr_outdeg.get(), B.seedbits.get(), B.r32.get(), c32n,
                                B.r64.get(), c64n, B.part.get(), B.bctr.get());
                        }
                    });
                    check_launch();
                    timed("ppr_reduce", 16.0 * KP * (double)nparts, s, [&] {
                        k_ppr_reduce<KP><<<1, PB, 0, s>>>(B.part.get(), nparts, B.lanes.get(), B.ctl.get());
                    });
                    check_launch();
                    continue;
                }
                timed("ppr_rows", edge_bytes * (double)t->m + (16.0 + 8.0 * KP) * (double)n, s, [&] {
                    k_ppr_rows<KP><<<light_blocks + seg_blocks, PB, 0, s>>>(
                        n, t->pr_off.get(), t->pr_tgt.get(), t->pr_outdeg.get(), B.seedbits.get(), B.lanes.get(), c32,
                        c32n, B.r32.get(), c64, c64n, B.r64.get(), B.part.get(), light_blocks, t->ppr_S,
                        t->ppr_sbeg.get(), t->ppr_send.get(), B.spart.get());
                });
                check_launch();
                if (t->ppr_H) {
                    timed("ppr_heavy", 8.0 * KP * (double)t->ppr_S, s, [&] {
                        k_ppr_heavy<KP><<<heavy_blocks, PB, 0, s>>>(
                            t->ppr_H, t->ppr_hrow.get(), t->ppr_hsoff.get(), B.spart.get(), t->pr_outdeg.get(),
                            B.seedbits.get(), B.lanes.get(), c32n, B.r32.get(), c64n, B.r64.get(),
                            B.part.get() + (uint64_t)light_blocks * KP * 2);
                    });
                    check_launch();
                }
                timed("ppr_reduce", 16.0 * KP * (double)nparts, s, [&] {
                    k_ppr_reduce<KP><<<1, PB, 0, s>>>(B.part.get(), nparts, B.lanes.get(), B.ctl.get());
                });
                check_launch();
            }
        };
        if (end - it0 == 8)
            graph.run(s, launches);
        else
            launches();
        it = end;
        PprCtl* h = static_cast<PprCtl*>(pinned_scratch(sizeof(PprCtl)));
        HIP_OK(hipMemcpyAsync(h, B.ctl.get(), sizeof(PprCtl), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        if (h->active == 0) break;
    }
}

template <int KP>
void ppr_write_out(const ursa_hip_topo* t, uint64_t n, PprBufs& B, int k, double* const* dst, hipStream_t s) {
    for (int q = 0; q < k; ++q) {
        k_ppr_out<KP><<<blocks_for(n, 256, 4096), 256, 0, s>>>(n, q, t->pr_new_of_old.get(), B.r64.get(), dst[q]);
        check_launch();
    }
}

}  // namespace

static void ppr_prepare_rows(ursa_hip_topo* t, hipStream_t s) {
    std::lock_guard<std::mutex> lk(t->cache_mu);
    if (t->ppr_ready) return;
    const uint64_t n = t->n;
    std::vector<uint32_t> off(n + 1), od(n);
    HIP_OK(hipMemcpyAsync(off.data(), t->pr_off.get(), (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    if (n) HIP_OK(hipMemcpyAsync(od.data(), t->pr_outdeg.get(), n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    std::vector<uint32_t> hrow, hsoff{0}, sbeg, send;
    uint64_t dangling = 0;
    for (uint64_t u = 0; u < n; ++u) {
        if (od[u] == 0) ++dangling;
        const uint32_t b = off[u], e = off[u + 1], d = e - b;
        if (d <= PPR_HEAVY) continue;
        const uint32_t len = std::max<uint32_t>(PPR_HEAVY, (d + PPR_MAXSEG - 1) / PPR_MAXSEG);
        for (uint32_t x = b; x < e; x += len) {
            sbeg.push_back(x);
            send.push_back(std::min<uint32_t>(e, x + len));
        }
        hrow.push_back((uint32_t)u);
        hsoff.push_back((uint32_t)sbeg.size());
    }
    auto up = [&](DevBuf<uint32_t>& d, const std::vector<uint32_t>& h) {
        d.alloc(h.size());
        if (!h.empty())
            HIP_OK(hipMemcpyAsync(d.get(), h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    };
    up(t->ppr_hrow, hrow);
    up(t->ppr_hsoff, hsoff);
    up(t->ppr_sbeg, sbeg);
    up(t->ppr_send, send);
    HIP_OK(hipStreamSynchronize(s));
    t->ppr_H = hrow.size();
    t->ppr_S = sbeg.size();
    t->ppr_dangling = dangling;
    t->ppr_ready = true;
}

void ppr_prepare(ursa_hip_topo* t, hipStream_t s) {
    NoArena keep_out_of_arena;
    pagerank_prepare(t, s);
    ppr_prepare_rows(t, s);
    // Automatic schedule: binned from BN_MIN_BLOCKS source blocks up on graphs
    // without heavy rows and with at least PPR_BN_MIN_DEGREE in-edges per row
    // on average (decided once; the layout is built here, outside any query's
    // workspace).
    bool bn = false;
    {
        std::lock_guard<std::mutex> lk(t->cache_mu);
        bn = !t->ppr_bn_auto && t->ppr_H == 0 && (t->n + BN_SSPAN - 1) / BN_SSPAN >= BN_MIN_BLOCKS &&
             t->m >= PPR_BN_MIN_DEGREE * t->n;
    }
    if (bn && bn_fits(t, s)) {
        bn_prepare(t, s);
        std::lock_guard<std::mutex> lk(t->cache_mu);
        t->ppr_bn_auto = true;
    }
}

// Whether solves on t sweep binned under the current schedule.
static bool ppr_binned(ursa_hip_topo* t, hipStream_t s) {
    const int sched = g_ppr_schedule.load();
    if (sched == 1) return false;
    if (sched == 2) return bn_fits(t, s);
    return t->ppr_bn_auto;
}

uint64_t ppr_workspace_bytes(const ursa_hip_topo* t, int k, bool host_out) {
    const uint64_t n = t->n, kp = (uint64_t)lanes_padded(std::max(k, 1));
    const uint64_t light_blocks = (n + PPR_ROWS - 1) / PPR_ROWS, heavy_blocks = (t->ppr_H + PB - 1) / PB;
    uint64_t b = n * kp * (4 + 8 + 8 + 16 + 16)  // r32, c32 x2, r64, c64 x2
                 + n * 4                         // seedbits
                 + (light_blocks + heavy_blocks + 1) * kp * 16 + (t->ppr_S + 1) * kp * 8 + 16 * 64 * 8;
    const int sched = g_ppr_schedule.load();
    if (sched == 2 || (sched == 0 && t->ppr_bn_auto)) {
        // Term array (the layout's slots; a nonempty tile pads to at most 8x
        // its slots when the layout is not built yet), claim counters, per-bin
        // partials (at most n / BN_RSPAN + m / BN_T2 + 1 bins).
        const uint64_t slots = t->bn_ready ? t->bn_slots : 8 * t->m;
        b += (std::max<uint64_t>(slots, BN_CH) + 1) * 8 + 64 + (n / BN_RSPAN + t->m / BN_T2 + 2) * kp * 16;
    }
    if (host_out) b += (uint64_t)k * n * 8;
    return b + (b >> 4) + (64u << 10) * 16;
}

// The lane of query q on t: damping, tol, max_iter, the teleport share of
// each of its distinct in-range seeds (returned in u, ascending) and its first
// phase; phase 3 (nothing to run, x0 is the result) when no seed is in range
// or max_iter is 0.
static PprLane ppr_make_lane(const ursa_hip_topo* t, const PprQuery& q, std::vector<uint32_t>& u) {
    const uint64_t n = t->n;
    PprLane L{};
    L.phase = 3;
    u.clear();
    for (uint32_t i = 0; i < q.nseeds; ++i)
        if (q.seeds[i] < n) u.push_back(q.seeds[i]);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    L.d = q.damping;
    L.tol = q.tol;
    L.max_iter = q.max_iter;
    if (u.empty() || q.max_iter == 0) {
        u.clear();
        return L;
    }
    L.inv_s = 1.0 / (double)u.size();
    L.pb = (1.0 - L.d) + L.d * ((double)t->ppr_dangling / (double)n);
    L.phase = (L.tol < PPR_F32_TOL && L.max_iter > PPR_F32_MIN_ITER) ? 0 : 2;
    return L;
}

void ppr_solve(ursa_hip_topo* t, const PprQuery* qs, int k, double* const* outs, uint32_t* iters_out) {
    URSA_REQUIRE(k >= 1 && k <= PPR_MAX_LANES, URSA_HIP_EINVAL, "a PPR solve takes 1..16 queries");
    const uint64_t n = t->n;
    for (int q = 0; q < k; ++q) {
        if (iters_out) iters_out[q] = 0;
        URSA_REQUIRE(qs[q].nseeds == 0 || qs[q].seeds != nullptr, URSA_HIP_EINVAL, "seeds is NULL");
        URSA_REQUIRE(qs[q].damping >= 0.0 && qs[q].damping <= 1.0, URSA_HIP_EINVAL, "damping must be in [0, 1]");
        URSA_REQUIRE(n == 0 || outs[q] != nullptr, URSA_HIP_EINVAL, "out is NULL");
    }
    if (n == 0) return;
    hipStream_t s = stream();
    ppr_prepare(t, s);
    const int KP = lanes_padded(k);

    // Host-side lane set-up: distinct in-range seeds per lane.
    std::vector<PprLane> lanes(KP);
    std::vector<uint32_t> seeds, lane_of, u;
    uint32_t max_iter_all = 0;
    for (int q = 0; q < KP; ++q) {
        PprLane& L = lanes[q];
        L = PprLane{};
        L.phase = 3;
        if (q >= k) continue;
        L = ppr_make_lane(t, qs[q], u);
        if (L.phase == 3) continue;
        max_iter_all = std::max(max_iter_all, L.max_iter);
        for (uint32_t v : u) {
            seeds.push_back(v);
            lane_of.push_back((uint32_t)q);
        }
    }

    PprBufs B;
    B.k = k;
    B.binned = ppr_binned(t, s);
    if (B.binned) {
        bn_prepare(t, s);
        B.term.alloc(std::max<uint64_t>(t->bn_slots, BN_CH));
        B.bctr.alloc(2);
        HIP_OK(hipMemsetAsync(B.bctr.get(), 0, 8, s));
        B.g1 = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(t->bn_n1, 4ull * cu_count()));
        B.g2 = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(t->bn_n2, 2ull * cu_count()));
    }
    B.r32.alloc(n * KP);
    B.c32a.alloc(n * KP);
    B.c32b.alloc(n * KP);
    B.r64.alloc(n * KP);
    B.c64a.alloc(n * KP);
    B.c64b.alloc(n * KP);
    const uint64_t light_blocks = (n + PPR_ROWS - 1) / PPR_ROWS, heavy_blocks = (t->ppr_H + PB - 1) / PB;
    B.part.alloc(std::max<uint64_t>(light_blocks + heavy_blocks, B.binned ? t->bn_bins : 0) * KP * 2);
    B.spart.alloc(std::max<uint64_t>(t->ppr_S, 1) * KP);
    B.seedbits.alloc(n);
    B.lanes.alloc(KP);
    B.ctl.alloc(1);
    HIP_OK(hipMemsetAsync(B.seedbits.get(), 0, n * sizeof(uint32_t), s));
    HIP_OK(hipMemcpyAsync(B.lanes.get(), lanes.data(), KP * sizeof(PprLane), hipMemcpyHostToDevice, s));
    if (!seeds.empty()) {
        B.seeds.alloc(seeds.size());
        B.lane_of.alloc(seeds.size());
        HIP_OK(hipMemcpyAsync(B.seeds.get(), seeds.data(), seeds.size() * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(B.lane_of.get(), lane_of.data(), seeds.size() * 4, hipMemcpyHostToDevice, s));
        k_ppr_seedbits<<<blocks_for(seeds.size(), 256), 256, 0, s>>>(B.seeds.get(), B.lane_of.get(),
                                                                     (uint32_t)seeds.size(), t->pr_new_of_old.get(),
                                                                     B.seedbits.get());
        check_launch();
    }

    std::vector<DevOut<double>> outs_dev;
    outs_dev.reserve(k);
    std::vector<double*> dst(k);
    for (int q = 0; q < k; ++q) {
        outs_dev.emplace_back(outs[q], n);
        dst[q] = outs_dev.back().p;
    }

    auto run = [&](auto kp_tag) {
        constexpr int KPc = decltype(kp_tag)::value;
        k_ppr_init<KPc><<<blocks_for(n, 256, 4096), 256, 0, s>>>(n, t->pr_outdeg.get(), B.r32.get(), B.c32a.get(),
                                                                 B.r64.get(), B.c64a.get());
        check_launch();
        ppr_iterate<KPc>(t, B, max_iter_all, s);
        ppr_write_out<KPc>(t, n, B, k, dst.data(), s);
    };
    switch (KP) {
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, 2>{}); break;
        case 4: run(std::integral_constant<int, 4>{}); break;
        case 8: run(std::integral_constant<int, 8>{}); break;
        default: run(std::integral_constant<int, 16>{}); break;
    }
    HIP_OK(hipMemcpyAsync(lanes.data(), B.lanes.get(), KP * sizeof(PprLane), hipMemcpyDeviceToHost, s));
    for (auto& o : outs_dev) o.finish(s);
    HIP_OK(hipStreamSynchronize(s));
    if (iters_out)
        for (int q = 0; q < k; ++q) iters_out[q] = lanes[q].iters;
}

// ---- cross-graph passes (kernels above: k_pprm_*) ----

namespace {

uint32_t pprm_vec_blocks(uint64_t n) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 256));
}

// Per-problem device layout inside one allocation (byte offsets, each array
// 256-byte aligned).
struct PprmLayout {
    uint64_t seedbits, r32, c32a, c32b, r64, c64a, c64b, part, spart, bytes;
    uint32_t light_blocks, seg_blocks, heavy_blocks;
};

PprmLayout pprm_layout(const ursa_hip_topo* t) {
    PprmLayout L{};
    const uint64_t n = t->n;
    L.light_blocks = (uint32_t)((n + PPR_ROWS - 1) / PPR_ROWS);
    L.seg_blocks = (uint32_t)((t->ppr_S + PPR_GROUPS - 1) / PPR_GROUPS);
    L.heavy_blocks = (uint32_t)((t->ppr_H + PB - 1) / PB);
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at += (bytes + 255) & ~(uint64_t)255;
        return o;
    };
    L.seedbits = take(n * 4);
    L.r32 = take(n * 4);
    L.c32a = take(n * 4);
    L.c32b = take(n * 4);
    L.r64 = take(n * 8);
    L.c64a = take(n * 8);
    L.c64b = take(n * 8);
    L.part = take((uint64_t)(L.light_blocks + L.heavy_blocks) * 16);
    L.spart = take(std::max<uint64_t>(t->ppr_S, 1) * 8);
    L.bytes = at;
    return L;
}

uint64_t align256(uint64_t b) { return (b + 255) & ~(uint64_t)255; }

}  // namespace

bool ppr_multi_eligible(ursa_hip_topo* t) {
    hipStream_t s = stream();
    ppr_prepare(t, s);
    return t->n > 0 && t->n <= PPR_MULTI_MAX_N && !ppr_binned(t, s);
}

uint64_t ppr_multi_workspace_bytes(const ursa_hip_topo* const* ts, int count) {
    uint64_t b = 0;
    for (int i = 0; i < count; ++i) b += pprm_layout(ts[i]).bytes + align256(ts[i]->n * 8) + 256;
    // Upload image: table, three prefix tables, lanes, controls, seeds.
    b += align256((uint64_t)count * (sizeof(PprProb) + sizeof(PprLane) + sizeof(PprCtl) + 12 + 2 * 4 * 16) + 64) +
         4 * 256;
    return b + (b >> 6) + (64u << 10);
}

void ppr_solve_multi(ursa_hip_topo* const* ts, const PprQuery* qs, int count, double* const* outs,
                     uint32_t* iters_out) {
    if (count <= 0) return;
    hipStream_t s = stream();
    const uint32_t np = (uint32_t)count;
    std::vector<PprmLayout> lay(np);
    std::vector<PprLane> lanes(np);
    std::vector<uint32_t> sprob, sval, u;
    uint64_t state_bytes = 0, out_elems = 0;
    uint32_t max_iter_all = 0;
    for (uint32_t i = 0; i < np; ++i) {
        ursa_hip_topo* t = ts[i];
        URSA_REQUIRE(ppr_multi_eligible(t), URSA_HIP_EINVAL, "graph not eligible for a cross-graph pass");
        URSA_REQUIRE(qs[i].nseeds == 0 || qs[i].seeds != nullptr, URSA_HIP_EINVAL, "seeds is NULL");
        URSA_REQUIRE(qs[i].damping >= 0.0 && qs[i].damping <= 1.0, URSA_HIP_EINVAL, "damping must be in [0, 1]");
        URSA_REQUIRE(outs[i] != nullptr, URSA_HIP_EINVAL, "out is NULL");
        if (iters_out) iters_out[i] = 0;
        lay[i] = pprm_layout(t);
        state_bytes += lay[i].bytes;
        out_elems += align256(t->n * 8) / 8;
        lanes[i] = ppr_make_lane(t, qs[i], u);
        if (lanes[i].phase != 3) max_iter_all = std::max(max_iter_all, lanes[i].max_iter);
        for (uint32_t v : u) {
            sprob.push_back(i);
            sval.push_back(v);
        }
    }
    // Upload image: [table | rows start | heavy start | vec start | lanes | ctl | seed problem | seed vertex].
    const uint64_t o_tab = 0, o_rs = align256(np * sizeof(PprProb)), o_hs = o_rs + align256((np + 1) * 4),
                   o_vs = o_hs + align256((np + 1) * 4), o_ln = o_vs + align256((np + 1) * 4),
                   o_ct = o_ln + align256(np * sizeof(PprLane)), o_sp = o_ct + align256(np * sizeof(PprCtl)),
                   o_sv = o_sp + align256(sprob.size() * 4), o_end = o_sv + align256(sval.size() * 4);
    DevBuf<char> blob(o_end), state(std::max<uint64_t>(state_bytes, 256));
    DevBuf<double> outbuf(std::max<uint64_t>(out_elems, 1));
    std::vector<char> img(o_end, 0);
    PprProb* tab = reinterpret_cast<PprProb*>(img.data() + o_tab);
    uint32_t* rs = reinterpret_cast<uint32_t*>(img.data() + o_rs);
    uint32_t* hs = reinterpret_cast<uint32_t*>(img.data() + o_hs);
    uint32_t* vs = reinterpret_cast<uint32_t*>(img.data() + o_vs);
    std::memcpy(img.data() + o_ln, lanes.data(), np * sizeof(PprLane));
    if (!sprob.empty()) {
        std::memcpy(img.data() + o_sp, sprob.data(), sprob.size() * 4);
        std::memcpy(img.data() + o_sv, sval.data(), sval.size() * 4);
    }
    char* sb = state.get();
    PprLane* d_lanes = reinterpret_cast<PprLane*>(blob.get() + o_ln);
    PprCtl* d_ctl = reinterpret_cast<PprCtl*>(blob.get() + o_ct);
    std::vector<bool> host_out(np);
    std::vector<uint64_t> out_at(np);
    uint64_t soff = 0, oat = 0;
    rs[0] = hs[0] = vs[0] = 0;
    for (uint32_t i = 0; i < np; ++i) {
        const ursa_hip_topo* t = ts[i];
        const PprmLayout& L = lay[i];
        PprProb& G = tab[i];
        G.off = t->pr_off.get();
        G.tgt = t->pr_tgt.get();
        G.od = t->pr_outdeg.get();
        G.sbeg = t->ppr_sbeg.get();
        G.send = t->ppr_send.get();
        G.hrow = t->ppr_hrow.get();
        G.hsoff = t->ppr_hsoff.get();
        G.new_of_old = t->pr_new_of_old.get();
        G.seedbits = reinterpret_cast<uint32_t*>(sb + soff + L.seedbits);
        G.r32 = reinterpret_cast<float*>(sb + soff + L.r32);
        G.c32a = reinterpret_cast<float*>(sb + soff + L.c32a);
        G.c32b = reinterpret_cast<float*>(sb + soff + L.c32b);
        G.r64 = reinterpret_cast<double*>(sb + soff + L.r64);
        G.c64a = reinterpret_cast<double*>(sb + soff + L.c64a);
        G.c64b = reinterpret_cast<double*>(sb + soff + L.c64b);
        G.part = reinterpret_cast<double*>(sb + soff + L.part);
        G.spart = reinterpret_cast<double*>(sb + soff + L.spart);
        G.lanes = d_lanes + i;
        G.ctl = d_ctl + i;
        host_out[i] = !is_device_ptr(outs[i]);
        out_at[i] = oat;
        G.out = host_out[i] ? outbuf.get() + oat : outs[i];
        oat += align256(t->n * 8) / 8;
        G.n = t->n;
        G.nseg = t->ppr_S;
        G.H = t->ppr_H;
        G.light_blocks = L.light_blocks;
        G.nparts = L.light_blocks + L.heavy_blocks;
        soff += L.bytes;
        rs[i + 1] = rs[i] + L.light_blocks + L.seg_blocks;
        hs[i + 1] = hs[i] + L.heavy_blocks;
        vs[i + 1] = vs[i] + pprm_vec_blocks(t->n);
    }
    const uint32_t rows_total = rs[np], heavy_total = hs[np], vec_total = vs[np];
    HIP_OK(hipMemcpyAsync(blob.get(), img.data(), o_end, hipMemcpyHostToDevice, s));
    const PprProb* dP = reinterpret_cast<const PprProb*>(blob.get() + o_tab);
    const uint32_t* d_rs = reinterpret_cast<const uint32_t*>(blob.get() + o_rs);
    const uint32_t* d_hs = reinterpret_cast<const uint32_t*>(blob.get() + o_hs);
    const uint32_t* d_vs = reinterpret_cast<const uint32_t*>(blob.get() + o_vs);
    k_pprm_init<<<vec_total, 256, 0, s>>>(dP, d_vs, np);
    check_launch();
    if (!sprob.empty()) {
        k_pprm_seeds<<<blocks_for(sprob.size(), 256), 256, 0, s>>>(
            dP, reinterpret_cast<const uint32_t*>(blob.get() + o_sp), reinterpret_cast<const uint32_t*>(blob.get() + o_sv),
            (uint32_t)sprob.size());
        check_launch();
    }
    uint32_t it = 0;
    while (it < max_iter_all) {
        const uint32_t end = std::min<uint32_t>(max_iter_all, it + 8);
        for (uint32_t i = it; i < end; ++i) {
            const int par = (int)(i & 1);
            timed("pprm_rows", 0.0, s, [&] { k_pprm_rows<<<rows_total, PB, 0, s>>>(dP, d_rs, np, par); });
            if (heavy_total) k_pprm_heavy<<<heavy_total, PB, 0, s>>>(dP, d_hs, np, par);
            k_pprm_reduce<<<np, PB, 0, s>>>(dP);
            check_launch();
        }
        it = end;
        PprCtl* h = static_cast<PprCtl*>(pinned_scratch(np * sizeof(PprCtl)));
        HIP_OK(hipMemcpyAsync(h, d_ctl, np * sizeof(PprCtl), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        bool active = false;
        for (uint32_t i = 0; i < np && !active; ++i) active = h[i].active != 0;
        if (!active) break;
    }
    k_pprm_out<<<vec_total, 256, 0, s>>>(dP, d_vs, np);
    check_launch();
    HIP_OK(hipMemcpyAsync(lanes.data(), d_lanes, np * sizeof(PprLane), hipMemcpyDeviceToHost, s));
    bool any_host = false;
    for (uint32_t i = 0; i < np; ++i) any_host |= host_out[i];
    if (any_host) {
        const uint64_t bytes = oat * 8;
        char* h = static_cast<char*>(pinned_scratch(bytes));
        HIP_OK(hipMemcpyAsync(h, outbuf.get(), bytes, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < np; ++i)
            if (host_out[i]) std::memcpy(outs[i], h + out_at[i] * 8, ts[i]->n * 8);
    } else {
        HIP_OK(hipStreamSynchronize(s));
    }
    if (iters_out)
        for (uint32_t i = 0; i < np; ++i) iters_out[i] = lanes[i].iters;
}

bool ppr_multi_candidate(const ursa_hip_topo* t) {
    return t->n > 0 && t->n <= PPR_MULTI_MAX_N && t->ppr_ready && g_ppr_schedule.load() != 2 &&
           (g_ppr_schedule.load() == 1 || !t->ppr_bn_auto);
}

void ppr_solve_many(ursa_hip_topo* const* ts, const PprQuery* qs, int count, double* const* outs,
                    uint32_t* iters_out) {
    std::vector<ursa_hip_topo*> mt;
    std::vector<PprQuery> mq;
    std::vector<double*> mo;
    std::vector<int> mi;
    uint64_t mn = 0;
    auto flush = [&] {
        if (mt.empty()) return;
        std::vector<uint32_t> it(mt.size(), 0);
        ppr_solve_multi(mt.data(), mq.data(), (int)mt.size(), mo.data(), it.data());
        if (iters_out)
            for (size_t j = 0; j < mi.size(); ++j) iters_out[mi[j]] = it[j];
        mt.clear();
        mq.clear();
        mo.clear();
        mi.clear();
        mn = 0;
    };
    for (int i = 0; i < count; ++i) {
        URSA_REQUIRE(ts[i] != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        ursa_hip_topo* t = ts[i];
        if (t->n == 0 || !ppr_multi_eligible(t)) {
            double* o[1] = {outs[i]};
            ppr_solve(t, &qs[i], 1, o, iters_out ? iters_out + i : nullptr);
            continue;
        }
        if (mt.size() == PPR_MULTI_MAX_PASS || mn + t->n > PPR_MULTI_MAX_NODES) flush();
        mt.push_back(t);
        mq.push_back(qs[i]);
        mo.push_back(outs[i]);
        mi.push_back(i);
        mn += t->n;
    }
    flush();
}

}  // namespace ursa

using namespace ursa;

extern "C" {

int ursa_hip_ppr_schedule(int mode) {
    return guard([&] {
        URSA_REQUIRE(mode >= 0 && mode <= 2, URSA_HIP_EINVAL, "schedule must be 0 (automatic), 1 (pull) or 2 (binned)");
        g_ppr_schedule.store(mode);
    });
}

int ursa_hip_ppr(ursa_hip_topo* t, const uint32_t* seeds, uint64_t nseeds, double damping, uint32_t max_iter,
                 double tol, double* out, uint32_t* iters_out) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(nseeds <= 0xFFFFFFFFull, URSA_HIP_EINVAL, "too many seeds");
        std::vector<uint32_t> hs(nseeds);
        if (nseeds) {
            URSA_REQUIRE(seeds != nullptr, URSA_HIP_EINVAL, "seeds is NULL");
            if (is_device_ptr(seeds))
                HIP_OK(hipMemcpy(hs.data(), seeds, nseeds * sizeof(uint32_t), hipMemcpyDeviceToHost));
            else
                std::copy(seeds, seeds + nseeds, hs.begin());
        }
        PprQuery q{hs.data(), (uint32_t)nseeds, damping, max_iter, tol};
        double* o[1] = {out};
        ursa::ppr_solve(t, &q, 1, o, iters_out);
    });
}

int ursa_hip_ppr_multi(uint32_t count, ursa_hip_topo* const* topos, const uint64_t* seed_offsets,
                       const uint32_t* seeds, double damping, uint32_t max_iter, double tol, double* const* outs,
                       uint32_t* iters_out) {
    return guard([&] {
        URSA_REQUIRE(count == 0 || (topos != nullptr && seed_offsets != nullptr && outs != nullptr), URSA_HIP_EINVAL,
                     "topos/seed_offsets/outs is NULL");
        std::vector<uint32_t> hs;
        if (count && seeds && is_device_ptr(seeds)) {
            hs.resize(seed_offsets[count]);
            if (!hs.empty())
                HIP_OK(hipMemcpy(hs.data(), seeds, hs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
            seeds = hs.data();
        }
        std::vector<PprQuery> qs(count);
        for (uint32_t i = 0; i < count; ++i) {
            const uint64_t a = seed_offsets[i], b = seed_offsets[i + 1];
            URSA_REQUIRE(b >= a && b - a <= 0xFFFFFFFFull, URSA_HIP_EINVAL, "seed_offsets must not decrease");
            URSA_REQUIRE(b == a || seeds != nullptr, URSA_HIP_EINVAL, "seeds is NULL");
            qs[i] = PprQuery{seeds ? seeds + a : nullptr, (uint32_t)(b - a), damping, max_iter, tol};
        }
        ursa::ppr_solve_many(topos, qs.data(), (int)count, outs, iters_out);
    });
}

int ursa_hip_ppr_batch(ursa_hip_topo* t, uint32_t k, const uint64_t* seed_offsets, const uint32_t* seeds,
                       double damping, uint32_t max_iter, double tol, double* out, uint32_t* iters_out) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(k == 0 || seed_offsets != nullptr, URSA_HIP_EINVAL, "seed_offsets is NULL");
        const uint64_t n = t->n;
        std::vector<uint32_t> hs;
        if (k && seeds && is_device_ptr(seeds)) {
            hs.resize(seed_offsets[k]);
            if (!hs.empty())
                HIP_OK(hipMemcpy(hs.data(), seeds, hs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
            seeds = hs.data();
        }
        for (uint32_t base = 0; base < k; base += PPR_MAX_LANES) {
            const int kk = (int)std::min<uint32_t>(PPR_MAX_LANES, k - base);
            std::vector<PprQuery> qs(kk);
            std::vector<double*> o(kk);
            for (int q = 0; q < kk; ++q) {
                const uint64_t a = seed_offsets[base + q], b = seed_offsets[base + q + 1];
                URSA_REQUIRE(b >= a && b - a <= 0xFFFFFFFFull, URSA_HIP_EINVAL, "seed_offsets must not decrease");
                URSA_REQUIRE(b == a || seeds != nullptr, URSA_HIP_EINVAL, "seeds is NULL");
                qs[q] = PprQuery{seeds ? seeds + a : nullptr, (uint32_t)(b - a), damping, max_iter, tol};
                o[q] = out ? out + (uint64_t)(base + q) * n : nullptr;
            }
            ursa::ppr_solve(t, qs.data(), kk, o.data(), iters_out ? iters_out + base : nullptr);
        }
    });
}

}  // extern "C"
