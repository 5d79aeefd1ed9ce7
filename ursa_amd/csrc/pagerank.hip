// pagerank.hip — pull PageRank over the device in-CSR (pagerank.rs:63-169).
//
// Working copy of the in-CSR (pr_prepare, cached on the topology): vertices
// renumbered so that rows with in-edges come first and, within each part, by
// decreasing out-degree (ties by id); sources relabelled and each row listing
// them in ascending new id, so the most gathered contributions sit together
// at the front of the contribution vector. Rows with in-edges are [0, nz).
// Beside it: one head bit per in-slot (set on the first slot of each row)
// and, per chunk of 512 consecutive slots, the row holding the chunk's first
// slot and three flags (the first slot is a head / some slot is a head /
// some row ends in the chunk).
//
// One f64 iteration = three launches:
//   k_pr_sweep   one wave per chunk, 8 consecutive slots per thread: two
//                16-byte column loads and 8 independent contribution gathers
//                (the first HOT contributions come from a copy in LDS made
//                once per workgroup), then the slots are summed row by row
//                along the head bits — no row offsets are read and nothing is
//                searched. Pieces of a row split across threads are joined by
//                a segmented scan over the wave. A row whose first and last
//                slot lie in the chunk is finished at once (fused epilogue):
//                   next[u]    = base + d * acc                (pagerank.rs:84)
//                   contrib[u] = next[u] / outdeg[u]           (pagerank.rs:116-120)
//                   delta     += |next[u] - rank[u]|           (pagerank.rs:44-52)
//                A row begun in an earlier chunk leaves the chunk's leading
//                piece; a row that continues past the chunk its trailing piece.
//   k_pr_fixup   rows spanning chunks (the trailing piece of the chunk holding
//                their head, then the leading pieces in chunk order) and rows
//                without in-edges (acc = 0) get the same epilogue; the delta
//                partials are folded in a fixed order.
//   k_pr_reduce  one workgroup finishes the delta sum in a fixed order,
//                records it and raises `done` once delta < tol
//                (pagerank.rs:87-90) or max_iter is reached;
// then the ordered dangling-mass sum (k_pr_dm_build, k_pr_dm_walk) sets
// base = (1-d)/n + d*dangling/n for the next iteration (pagerank.rs:76-79).
// Every reduction has a fixed order, so results are bit-identical run to run.
// `done` lives on the device and every launch returns at once when it is set,
// so the host queues iterations in batches of 8 (recorded once as a HIP graph
// and replayed) and reads the flag between batches.
//
// HBM traffic per f64 iteration (algorithmic): 12 B per in-edge (column +
// contribution gather) and 32 B per node — 12m + 32n. The weighted form adds
// the edge weight in slot order (f32 when every weight is exact in f32) and
// reads an f64 out-strength instead of the u32 out-degree. Weighted terms are
// w * (rank / strength), the contribution being computed once per node, where
// the reference computes (w * rank) / strength per edge (pagerank.rs:161-166):
// the same terms, each within one rounding of the reference's.
//
// Mixed precision (ursa_hip_pagerank_mixed) is iterative refinement around
// the same sweep. f32 sweeps (f32 ranks and contributions, compensated row
// sums) run from the uniform start until delta < 1e-6 or stops shrinking;
// an f64 sweep x1 = F(x0) of the widened iterate yields the residual
// r = x1 - x0, computed exactly as the f64 solve computes it; f32 sweeps then
// solve the correction e = r + d * (A e + dangling(e) / n) until their own
// delta is below tol / 4; the corrected iterate x0 + e gets the next f64
// sweep, and the solve ends after the first f64 sweep whose delta is below
// tol, returning that sweep's output as the reference returns its newest
// iterate. Only f64 sweeps decide convergence. A coarse tol (>= 1e-6) or a
// max_iter of 16 or less runs the f64 solve itself.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <type_traits>
#include <vector>

#include "binned.hpp"
#include "fixed_sum.hpp"
#include "ordered_sum.hpp"
#include "repeated_sum.hpp"
#include "prim.hpp"
#include "topology.hpp"

namespace ursa {

namespace {

constexpr int PR_ITEMS = 8;                  // in-slots per thread
constexpr int PR_CHUNK = 64 * PR_ITEMS;      // in-slots per wave
constexpr int PR_WORDS = PR_CHUNK / 32;      // head-bit words per chunk
constexpr int PR_FIX = 256;                  // fix-up workgroup
constexpr int PR_HOT_BYTES = 64 << 10;       // LDS copy of the hottest contributions
constexpr int PR_HOT_BLOCK = 512;            // workgroup of the LDS-copy sweep (two per CU)
constexpr int PR_PLAIN_BLOCK = 256;          // workgroup of the sweep without the copy
constexpr double PR_F32_TOL = 1e-6;          // the first f32 phase ends below this delta
constexpr uint32_t PR_F32_MIN_ITER = 16;     // max_iter at or below this: f64 solve
constexpr uint32_t PR_MAX_CORR = 40;         // correction sweeps per refinement round at most

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct PrCtl {
    uint32_t done;
    uint32_t iters;
    double delta;
    double dm;  // tree sum of the dangling values of the last sweep (mixed solve)
    double base[2];
    int32_t dm_e0;  // lowest candidate binade of the ordered dangling sum
};

struct PrPart {
    double delta, dm;
};

// Deterministic sum over a workgroup; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += lds[w];
    __syncthreads();
    return tot;
}

__device__ __forceinline__ uint32_t wave_excl_scan_u32(uint32_t v) {
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(x, off, 64);
        if (lane >= off) x += t;
    }
    return x - v;
}

// Per-thread running sum of a row piece. f64 contributions add plainly; f32
// contributions use Neumaier's compensated two-sum (the rounding error of each
// add is carried in c), so an f32 row piece is accurate to about one f32 ulp of
// its magnitude before it is widened to f64.
template <class CT>
struct RowAcc;

template <>
struct RowAcc<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ void reset() { s = 0.0; }
};

template <>
struct RowAcc<float> {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
#pragma clang fp contract(off)
        const float t = s + x;
        if (fabsf(s) >= fabsf(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s + (double)c; }
    __device__ __forceinline__ void reset() {
        s = 0.f;
        c = 0.f;
    }
};

template <bool WEIGHTED>
__device__ __forceinline__ double node_scale(const uint32_t* out_deg, const double* out_str, uint64_t u) {
    if constexpr (WEIGHTED)
        return __builtin_nontemporal_load(&out_str[u]);
    else
        return (double)__builtin_nontemporal_load(&out_deg[u]);
}

// Row epilogue shared by the sweep and fix-up kernels: u's new value from its
// row sum, its next contribution, and its share of delta (and, with DM, of the
// dangling mass). `scale` > 0 means u has out-mass. `b` is the constant term
// (base, plus the residual in a correction sweep). RT is the rank type, CT the
// contribution type of the following sweep.
template <class RT, class CT, bool DM>
__device__ __forceinline__ void pr_finish(uint64_t u, double acc, double b, double damping, double scale, double cur,
                                          RT* next, CT* cnext, double& delta, double& dm) {
    double v = __dadd_rn(b, __dmul_rn(damping, acc));  // no contraction: pagerank.rs:84
    if constexpr (sizeof(RT) == 4) v = (double)(float)v;
    __builtin_nontemporal_store((RT)v, &next[u]);
    const bool has_out = scale > 0.0;
    cnext[u] = (CT)(has_out ? v / scale : 0.0);
    delta += fabs(v - cur);
    if constexpr (DM)
        if (!has_out) dm += v;
}

// One sweep: `rank`/`contrib` are the current iterate, `next`/`cnext` the new
// one. base = ctl->base[parity] (device-driven solve, parity >= 0) or base_arg.
// With CORR the constant term of row u is base + corr[u] (correction sweeps of
// the mixed solve). HOT > 0: contributions [0, HOT) are read from an LDS copy.
template <class CT, class RT, class W, bool WEIGHTED, bool CORR, bool DM, int BLOCK, int HOT>
__global__ __launch_bounds__(BLOCK) void k_pr_sweep(
    const uint32_t* __restrict__ tgt, const W* __restrict__ w_in, const uint32_t* __restrict__ head,
    const uint32_t* __restrict__ crow, uint64_t m, uint64_t nchunks, const uint32_t* __restrict__ out_deg,
    const double* __restrict__ out_str, const RT* __restrict__ rank, const CT* __restrict__ contrib,
    RT* __restrict__ next, CT* __restrict__ cnext, const float* __restrict__ corr, double* __restrict__ lead_val,
    double* __restrict__ tail_val, PrPart* __restrict__ part, const PrCtl* ctl, int parity, double base_arg,
    double damping) {
    constexpr int WAVES = BLOCK / 64;
    __shared__ double s_red[2][WAVES];
    __shared__ CT s_hot[HOT > 0 ? HOT : 1];
    if (ctl->done) return;
    const double base = parity >= 0 ? ctl->base[parity] : base_arg;
    if constexpr (HOT > 0) {
        for (uint32_t i = threadIdx.x; i < (uint32_t)HOT; i += BLOCK) s_hot[i] = contrib[i];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double delta = 0.0, dm = 0.0;
    const uint64_t ngroups = (nchunks + WAVES - 1) / WAVES;
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint64_t c = g * WAVES + wave;
        if (c >= nchunks) break;  // wave-uniform; later groups lie further out
        const uint64_t i0 = c * PR_CHUNK + (uint64_t)lane * PR_ITEMS;
        const int cnt = i0 >= m ? 0 : (int)(m - i0 < PR_ITEMS ? m - i0 : PR_ITEMS);

        uint32_t col[PR_ITEMS];
        if (cnt == PR_ITEMS) {
            const u32x4* p = reinterpret_cast<const u32x4*>(tgt + i0);
            const u32x4 a = __builtin_nontemporal_load(p), b = __builtin_nontemporal_load(p + 1);
            col[0] = a.x; col[1] = a.y; col[2] = a.z; col[3] = a.w;
            col[4] = b.x; col[5] = b.y; col[6] = b.z; col[7] = b.w;
        } else {
#pragma unroll
            for (int j = 0; j < PR_ITEMS; ++j) col[j] = j < cnt ? tgt[i0 + j] : 0u;
        }
        CT v[PR_ITEMS];
#pragma unroll
        for (int j = 0; j < PR_ITEMS; ++j) {
            CT x = (CT)0;
            if (j < cnt) {
                if constexpr (HOT > 0) {
                    if (col[j] < (uint32_t)HOT)
                        x = s_hot[col[j]];
                    else
                        x = contrib[col[j]];
                } else {
                    x = contrib[col[j]];
                }
            }
            v[j] = x;
        }
        if constexpr (WEIGHTED) {
#pragma unroll
            for (int j = 0; j < PR_ITEMS; ++j)
                if (j < cnt) v[j] = (CT)__builtin_nontemporal_load(&w_in[i0 + j]) * v[j];
        }

        // Head bits of this thread's slots; a thread past the last slot acts
        // as the start of a row (the end of the slot array).
        uint32_t hb = 1u;
        if (cnt > 0) hb = (head[i0 >> 5] >> (uint32_t)(i0 & 31)) & 0xFFu;
        uint32_t nxt = __shfl_down(hb, 1, 64) & 1u;  // is slot i0 + 8 a row start?
        if (lane == 63) {
            const uint64_t i8 = (c + 1) * PR_CHUNK;
            nxt = i8 >= m ? 1u : (head[i8 >> 5] & 1u);
        }
        // end bit j: slot j is the last slot of its row
        uint32_t E = (hb >> 1) | (nxt << 7);
        if (cnt < PR_ITEMS) E = cnt ? ((E & ((1u << cnt) - 1u)) | (1u << (cnt - 1))) : 0u;
        // row(j) = rbase + popcount(hb & ((2 << j) - 1))
        const uint32_t h0c = __shfl(hb, 0, 64) & 1u;
        const uint32_t rbase = crow[c] - h0c + wave_excl_scan_u32((uint32_t)__popc(hb));

        RowAcc<CT> acc;
        double ev[PR_ITEMS];
        uint32_t emask = 0;
        double lead = 0.0;
        int lead_j = -1;
        bool open_first = (hb & 1u) == 0u;  // slot 0 continues a row begun before this thread
#pragma unroll
        for (int j = 0; j < PR_ITEMS; ++j) {
            ev[j] = 0.0;
            if (j < cnt) {
                acc.add(v[j]);
                if ((E >> j) & 1u) {
                    const double sv = acc.value();
                    acc.reset();
                    if (open_first) {
                        lead = sv;
                        lead_j = j;
                        open_first = false;
                    } else {
                        ev[j] = sv;
                        emask |= 1u << j;
                    }
                }
            }
        }

        // Segmented scan over the wave of (open trailing piece, has a row start):
        // a thread's carry is the piece of its first row summed by the preceding
        // threads since that row's start.
        double sv = acc.value();
        uint32_t f = hb != 0u;
        for (int off = 1; off < 64; off <<= 1) {
            const double vu = __shfl_up(sv, off, 64);
            const uint32_t fu = __shfl_up(f, off, 64);
            if (lane >= off) {
                if (!f) sv = vu + sv;
                f |= fu;
            }
        }
        double cin = __shfl_up(sv, 1, 64);
        uint32_t fin = __shfl_up(f, 1, 64);
        if (lane == 0) {
            cin = 0.0;
            fin = 0u;
        }
        if (lead_j >= 0) {
            const double val = cin + lead;
            if (fin) {
                ev[lead_j] = val;
                emask |= 1u << lead_j;
            } else {
                lead_val[c] = val;  // the row began in an earlier chunk: k_pr_fixup finishes it
            }
        }
        if (lane == 63) {
            if (!f) {
                if (lead_j < 0) lead_val[c] = sv;  // the whole chunk lies inside one row
            } else if (!nxt) {
                tail_val[c] = sv;  // the last row continues into the next chunk
            }
        }

        // Epilogue of the rows finished here: their loads first, then the stores.
        double sc[PR_ITEMS], cur[PR_ITEMS], bb[PR_ITEMS];
        uint32_t er[PR_ITEMS];
#pragma unroll
        for (int j = 0; j < PR_ITEMS; ++j) {
            er[j] = rbase + (uint32_t)__popc(hb & ((2u << j) - 1u));
            if ((emask >> j) & 1u) {
                sc[j] = node_scale<WEIGHTED>(out_deg, out_str, er[j]);
                cur[j] = (double)__builtin_nontemporal_load(&rank[er[j]]);
                bb[j] = base;
                if constexpr (CORR) bb[j] = base + (double)corr[er[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < PR_ITEMS; ++j)
            if ((emask >> j) & 1u) pr_finish<RT, CT, DM>(er[j], ev[j], bb[j], damping, sc[j], cur[j], next, cnext, delta, dm);
    }
    for (int off = 32; off > 0; off >>= 1) {
        delta += __shfl_down(delta, off, 64);
        dm += __shfl_down(dm, off, 64);
    }
    if (lane == 0) {
        s_red[0][wave] = delta;
        s_red[1][wave] = dm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        PrPart r{0.0, 0.0};
        for (int w = 0; w < WAVES; ++w) {
            r.delta += s_red[0][w];
            r.dm += s_red[1][w];
        }
        part[blockIdx.x] = r;
    }
}

// Rows spanning chunks and rows without in-edges, then the sweep's partials
// folded per workgroup in index order.
template <class CT, class RT, bool WEIGHTED, bool CORR, bool DM>
__global__ __launch_bounds__(PR_FIX) void k_pr_fixup(
    const uint8_t* __restrict__ cflag, const uint32_t* __restrict__ crow, const uint32_t* __restrict__ off,
    uint64_t nchunks, uint64_t nz, uint64_t n,
    const uint32_t* __restrict__ out_deg, const double* __restrict__ out_str, const RT* __restrict__ rank,
    RT* __restrict__ next, CT* __restrict__ cnext, const float* __restrict__ corr,
    const double* __restrict__ lead_val, const double* __restrict__ tail_val, const PrPart* __restrict__ part,
    uint64_t nparts, PrPart* __restrict__ fpart, const PrCtl* ctl, int parity, double base_arg, double damping) {
    __shared__ double s_red[PR_FIX / 64];
    if (ctl->done) return;
    const double base = parity >= 0 ? ctl->base[parity] : base_arg;
    const uint64_t i = blockIdx.x * (uint64_t)PR_FIX + threadIdx.x;
    double delta = 0.0, dm = 0.0;
    if (i < nparts) {
        delta = part[i].delta;
        dm = part[i].dm;
    }
    uint64_t u = n;
    double acc = 0.0;
    if (i < nchunks) {
        const uint32_t f = cflag[i];
        if (!(f & 1u) && (f & 4u)) {
            // The row u holding this chunk's first slot ends here; it began in
            // chunk j, the one holding its head slot, and the chunks between lie
            // inside it. Its pieces are added left to right; the loads are
            // issued eight at a time (a hub's row spans hundreds of chunks).
            u = crow[i];
            const uint64_t j = off[u] / PR_CHUNK;
            acc = tail_val[j];
            uint64_t q = j + 1;
            for (; q + 8 <= i + 1; q += 8) {
                double v[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) v[r] = lead_val[q + r];
#pragma unroll
                for (int r = 0; r < 8; ++r) acc += v[r];
            }
            for (; q <= i; ++q) acc += lead_val[q];
        }
    } else if (i < nchunks + (n - nz)) {
        u = nz + (i - nchunks);  // no in-edges: acc = 0 (pagerank.rs:113-123)
    }
    if (u < n) {
        double b = base;
        if constexpr (CORR) b = base + (double)corr[u];
        pr_finish<RT, CT, DM>(u, acc, b, damping, node_scale<WEIGHTED>(out_deg, out_str, u), (double)rank[u], next, cnext,
                              delta, dm);
    }
    const double bd = block_sum(delta, s_red);
    const double bm = block_sum(dm, s_red);
    if (threadIdx.x == 0) fpart[blockIdx.x] = PrPart{bd, bm};
}

__global__ __launch_bounds__(1024) void k_pr_reduce(const PrPart* __restrict__ fpart, uint64_t count, PrCtl* ctl,
                                                    uint32_t max_iter, double tol) {
    __shared__ double s_red[1024 / 64];
    if (ctl->done) return;
    double d = 0.0, x = 0.0;
    for (uint64_t i = threadIdx.x; i < count; i += blockDim.x) {
        d += fpart[i].delta;
        x += fpart[i].dm;
    }
    const double delta = block_sum(d, s_red);
    const double dmt = block_sum(x, s_red);
    if (threadIdx.x == 0) {
        const uint32_t iters = ctl->iters + 1;  // the iteration count lives on the device
        ctl->delta = delta;
        ctl->dm = dmt;
        ctl->iters = iters;
        if (delta < tol || iters >= max_iter) ctl->done = 1;
    }
}

// Dangling mass of the iterate `rank` summed over the dangling nodes in
// ascending node order, rounded as the reference's sequential loop rounds
// (pagerank.rs:76-78): the gather and the chunk / group tables here, the
// ordered walk in k_pr_dm_walk.
__global__ __launch_bounds__(OS_BUILD_THREADS) void k_pr_dm_build(const double* __restrict__ rank,
                                                                  const uint32_t* __restrict__ dang, uint64_t d,
                                                                  const PrCtl* ctl, double* __restrict__ xd,
                                                                  uint64_t* __restrict__ cq, uint32_t* __restrict__ cm,
                                                                  double* __restrict__ cs) {
    if (ctl->done) return;
    os_build(rank, dang, d, ctl->dm_e0, xd, cq, cm, cs);
}

// base = (1-d)/n + d*dangling/n for the iteration of parity `parity`
// (pagerank.rs:79), from the ordered dangling sum.
__global__ __launch_bounds__(OS_WALK_THREADS) void k_pr_dm_walk(const double* __restrict__ xd, uint64_t d,
                                                                const uint64_t* __restrict__ cq,
                                                                const uint32_t* __restrict__ cm,
                                                                const double* __restrict__ cs, PrCtl* ctl, int parity,
                                                                double damping, double nf) {
    __shared__ OsLds lds;
    if (ctl->done) return;
    const double dm = os_walk(xd, d, cq, cm, cs, ctl->dm_e0, lds);
    if (threadIdx.x == 0) {
        ctl->base[parity] = __dadd_rn(__ddiv_rn(1.0 - damping, nf), __ddiv_rn(__dmul_rn(damping, dm), nf));
        const OsState st = os_state(dm);
        if (st.fast) ctl->dm_e0 = st.e - 13;
    }
}

// rank0 = 1/n, contrib0 = rank0 / scale (in the precision of RT / CT).
template <class RT, class CT, bool WEIGHTED>
__global__ void k_pr_init(uint64_t n, double nf, const uint32_t* __restrict__ out_deg,
                          const double* __restrict__ out_str, RT* __restrict__ rank, CT* __restrict__ contrib) {
    const uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (u < n) {
        const RT r = (RT)(1.0 / nf);
        rank[u] = r;
        const double sc = node_scale<WEIGHTED>(out_deg, out_str, u);
        contrib[u] = (CT)(sc > 0.0 ? (double)r / sc : 0.0);
    }
}

// Dangling flags in the original node order: u is dangling when its
// out-degree is 0 or its out-strength is <= 0, read at its renumbered slot
// (pagerank.rs:109-111, 152-153).
template <bool WEIGHTED>
__global__ void k_pr_dang_flags(uint64_t n, const uint32_t* __restrict__ new_of_old, const uint32_t* __restrict__ out_deg,
                                const double* __restrict__ out_str, uint8_t* __restrict__ flags) {
    const uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (u < n) flags[u] = node_scale<WEIGHTED>(out_deg, out_str, new_of_old[u]) <= 0.0 ? 1 : 0;
}

__global__ void k_pr_ctl_init(PrCtl* ctl, int32_t dm_e0, double base0) {
    ctl->done = 0;
    ctl->iters = 0;
    ctl->delta = 0.0;
    ctl->dm = 0.0;
    ctl->base[0] = base0;
    ctl->base[1] = 0.0;
    ctl->dm_e0 = dm_e0;
}

// out_str[v] = sum of v's out-edge weights (pagerank.rs:148-150): in out-CSR
// order for vertices with at most OS_LIGHT out-edges; a wave per heavier
// vertex (k_out_strength_heavy) sums lane-strided pieces into eight partial
// sums per lane and joins them in a fixed order and butterfly, so a hub's sum
// is not one thread's serial walk. The order differs from the reference's left-to-right sum only for
// heavy vertices (same terms, within a few ulps).
constexpr uint32_t OS_LIGHT = 64;

__global__ void k_out_strength(const uint32_t* __restrict__ out_off, const uint32_t* __restrict__ out_eid,
                               const double* __restrict__ w, uint64_t n, double* __restrict__ out_str) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = out_off[v], b = out_off[v + 1];
        if (b - a > OS_LIGHT) continue;
        double s = -0.0;
        for (uint32_t k = a; k < b; ++k) s += w[out_eid[k]];
        out_str[v] = s;
    }
}

__global__ __launch_bounds__(256) void k_out_strength_heavy(const uint32_t* __restrict__ out_off,
                                                            const uint32_t* __restrict__ out_eid,
                                                            const double* __restrict__ w, uint64_t n,
                                                            double* __restrict__ out_str) {
    const int lane = threadIdx.x & 63;
    const uint64_t gw = blockIdx.x * (uint64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t v0 = gw * 64; v0 < n; v0 += nw * 64) {
        const uint64_t mine = v0 + lane;
        const bool heavy = mine < n && out_off[mine + 1] - out_off[mine] > OS_LIGHT;
        uint64_t mask = __ballot(heavy);
        while (mask) {
            const int j = __ffsll((unsigned long long)mask) - 1;
            mask &= mask - 1;
            const uint64_t v = v0 + j;
            const uint32_t a = out_off[v], b = out_off[v + 1];
            // Eight partial sums per lane (slots a + lane + 64 (8 i + j) go to
            // partial j) keep eight weight gathers in flight; the partials are
            // then joined in a fixed order.
            double p[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) p[j] = -0.0;
            uint32_t k = a + lane;
            for (; k + 7 * 64 < b; k += 8 * 64) {
                uint32_t e[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) e[j] = out_eid[k + j * 64];
#pragma unroll
                for (int j = 0; j < 8; ++j) p[j] += w[e[j]];
            }
            for (int j = 0; k < b; k += 64, ++j) p[j] += w[out_eid[k]];
            double s = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if (lane == 0) out_str[v] = s;
        }
    }
}

// Weights permuted into slot order; flags any weight not exact in f32.
__global__ void k_permute_weights(const uint32_t* __restrict__ in_eid, const double* __restrict__ w, uint64_t m,
                                  double* __restrict__ w64, float* __restrict__ w32, uint32_t* inexact) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const double x = w[in_eid[k]];
        const float f = (float)x;
        w64[k] = x;
        w32[k] = f;
        if ((double)f != x) atomicOr(inexact, 1u);
    }
}

// Mixed solve: x64 = widened f32 iterate, with its f64 contributions.
template <bool WEIGHTED>
__global__ void k_pr_widen(uint64_t n, const uint32_t* __restrict__ out_deg, const double* __restrict__ out_str,
                           const float* __restrict__ x32, double* __restrict__ x64, double* __restrict__ c64) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const double x = (double)x32[u];
        const double sc = node_scale<WEIGHTED>(out_deg, out_str, u);
        x64[u] = x;
        c64[u] = sc > 0.0 ? x / sc : 0.0;
    }
}

// Mixed solve: residual r = x1 - x0 of an f64 sweep (rounded to f32), which is
// also the correction solve's start e0 = r; per-workgroup dangling sums of r.
template <bool WEIGHTED>
__global__ __launch_bounds__(PR_FIX) void k_pr_residual(uint64_t n, const uint32_t* __restrict__ out_deg,
                                                        const double* __restrict__ out_str,
                                                        const double* __restrict__ x1, const double* __restrict__ x0,
                                                        float* __restrict__ r32, float* __restrict__ e32,
                                                        float* __restrict__ ce32, PrPart* __restrict__ fpart) {
    __shared__ double s_red[PR_FIX / 64];
    const uint64_t u = blockIdx.x * (uint64_t)PR_FIX + threadIdx.x;
    double dm = 0.0;
    if (u < n) {
        const float r = (float)(x1[u] - x0[u]);
        const double sc = node_scale<WEIGHTED>(out_deg, out_str, u);
        r32[u] = r;
        e32[u] = r;
        ce32[u] = sc > 0.0 ? (float)((double)r / sc) : 0.f;
        if (!(sc > 0.0)) dm = (double)r;
    }
    const double bm = block_sum(dm, s_red);
    if (threadIdx.x == 0) fpart[blockIdx.x] = PrPart{0.0, bm};
}

// Mixed solve: x0 += e, with the f64 contributions of the corrected iterate.
template <bool WEIGHTED>
__global__ void k_pr_apply(uint64_t n, const uint32_t* __restrict__ out_deg, const double* __restrict__ out_str,
                           const float* __restrict__ e32, double* __restrict__ x64, double* __restrict__ c64) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const double x = x64[u] + (double)e32[u];
        const double sc = node_scale<WEIGHTED>(out_deg, out_str, u);
        x64[u] = x;
        c64[u] = sc > 0.0 ? x / sc : 0.0;
    }
}

// ---- renumbering: rows with in-edges first, then by decreasing out-degree ----
__global__ void k_rl_keys(const uint32_t* __restrict__ out_deg, const uint32_t* __restrict__ in_off, uint64_t n,
                          uint64_t* keys, uint32_t* iota) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t no_in = in_off[v + 1] == in_off[v] ? 1u : 0u;
        keys[v] = no_in << 32 | (uint32_t)~out_deg[v];
        iota[v] = (uint32_t)v;
    }
}

__global__ void k_rl_invert(const uint32_t* __restrict__ old_of_new, uint64_t n, const uint32_t* __restrict__ in_off,
                            const uint32_t* __restrict__ out_deg, uint32_t* __restrict__ new_of_old,
                            uint32_t* __restrict__ rdeg, uint32_t* __restrict__ routdeg) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = old_of_new[i];
        new_of_old[o] = (uint32_t)i;
        rdeg[i] = in_off[o + 1] - in_off[o];
        routdeg[i] = out_deg[o];
    }
}

// nz = number of rows with in-edges (they come first in the new numbering).
__global__ void k_rl_nz(const uint32_t* __restrict__ rdeg, uint64_t n, uint64_t* nz) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (rdeg[i] > 0 && (i + 1 == n || rdeg[i + 1] == 0)) *nz = i + 1;
}

// Slot k of row u becomes the key (new(u), new(source)) with its edge id
// beside it; a stable sort of the keys then lists every row's sources in
// ascending new id. Lanes that gather the same contribution line sit next to
// each other, so the wave asks L2 for that line once.
__global__ void k_rl_slot_keys(const uint32_t* __restrict__ owner, const uint32_t* __restrict__ in_tgt,
                               const uint32_t* __restrict__ new_of_old, uint64_t m, uint64_t* __restrict__ key) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x)
        key[k] = (uint64_t)new_of_old[owner[k]] << 32 | new_of_old[in_tgt[k]];
}

__global__ void k_rl_low_words(const uint64_t* __restrict__ key, uint64_t m, uint32_t* __restrict__ r_tgt) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x)
        r_tgt[k] = (uint32_t)key[k];
}

// Head bit of the first slot of every row with in-edges.
__global__ void k_pr_heads(const uint32_t* __restrict__ off, uint64_t nz, uint32_t* __restrict__ head) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nz; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = off[r];
        atomicOr(&head[k >> 5], 1u << (k & 31));
    }
}

// Per chunk: the row holding its first slot (largest r < nz with off[r] <= slot)
// and its flags (bit 0 first slot starts a row, bit 1 some slot starts a row,
// bit 2 some row ends in the chunk, i.e. a row starts in (first, first + 512]
// with the end of the slot array counting as a start).
__global__ void k_pr_chunks(const uint32_t* __restrict__ off, uint64_t nz, const uint32_t* __restrict__ head,
                            uint64_t m, uint64_t nchunks, uint32_t* __restrict__ crow, uint8_t* __restrict__ cflag) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nchunks;
         c += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t slot = c * PR_CHUNK;
        uint64_t lo = 0, hi = nz;  // off[lo] <= slot < off[hi]
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (off[mid] <= slot)
                lo = mid;
            else
                hi = mid;
        }
        crow[c] = (uint32_t)lo;
        uint32_t any = 0, later = 0;
        for (int w = 0; w < PR_WORDS; ++w) {
            const uint32_t x = head[c * PR_WORDS + w];
            any |= x;
            later |= w == 0 ? (x & ~1u) : x;
        }
        const uint64_t nxt = slot + PR_CHUNK;
        const uint32_t next_start = nxt >= m ? 1u : (head[nxt >> 5] & 1u);
        cflag[c] = (uint8_t)((head[c * PR_WORDS] & 1u) | (any ? 2u : 0u) | ((later || next_start) ? 4u : 0u));
    }
}

__global__ void k_gather_f64(const double* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                             double* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

__global__ void k_fill(double* p, uint64_t n, double v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Sweep schedule (ursa_hip_pagerank_schedule): 0 automatic, 1 pull, 2 binned.
std::atomic<int> g_pr_schedule{0};

// ---- binned sweep (binned.hpp) ----

// Phase 1 of a PageRank sweep (bn_scatter_items); also clears phase 2's claim
// counter ctr[1] for this sweep.
template <class CT, class W, bool WEIGHTED>
__global__ __launch_bounds__(BN_P1_BLOCK) void k_bn_scatter(const uint16_t* __restrict__ bsrc,
                                                            const uint32_t* __restrict__ place,
                                                            const W* __restrict__ w,
                                                            const uint32_t* __restrict__ items, uint32_t nitems,
                                                            uint64_t n, const CT* __restrict__ contrib,
                                                            CT* __restrict__ term, uint32_t* __restrict__ ctr,
                                                            const PrCtl* ctl) {
    __shared__ CT s_c[BN_SSPAN];
    __shared__ uint32_t s_it;
    if (ctl->done) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[1] = 0;
    bn_scatter_items<CT, CT, W, WEIGHTED>(bsrc, place, w, items, nitems, n, contrib, 1u, term, ctr, s_c, &s_it);
}

// Phase 2: item = (bin b, bin-major chunks [q0, q1), first row r0, rows),
// claimed through ctr[1] in the order of the item list (largest first); also
// clears phase 1's counter ctr[0] for the next sweep. K: the f32 sweeps'
// fixed-point scale (unused by f64 sweeps).
template <class CT, class RT, bool WEIGHTED, bool CORR, bool DM>
__global__ __launch_bounds__(BN_P2_BLOCK, 4) void k_bn_gather(
    const uint16_t* __restrict__ bdst, const CT* __restrict__ term,
    const uint32_t* __restrict__ items, uint32_t nitems, uint64_t n, const uint32_t* __restrict__ out_deg,
    const double* __restrict__ out_str, const RT* __restrict__ rank, RT* __restrict__ next, CT* __restrict__ cnext,
    const float* __restrict__ corr, PrPart* __restrict__ part, uint32_t* __restrict__ ctr, const PrCtl* ctl,
    int parity, double base_arg, double damping, int K) {
    constexpr int WAVES = BN_P2_BLOCK / 64;
    constexpr int PARTS = BnAcc<CT>::PARTS;
    __shared__ unsigned long long s_acc[PARTS * BN_RSPAN];
    __shared__ double s_red[WAVES];
    __shared__ uint32_t s_it;
    if (ctl->done) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[0] = 0;
    const double base = parity >= 0 ? ctl->base[parity] : base_arg;
    const float scale = sizeof(CT) == 4 ? __builtin_ldexpf(1.0f, K - 32) : 0.0f;
    const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (;;) {
        if (threadIdx.x == 0) s_it = atomicAdd(&ctr[1], 1u);
        for (uint32_t i = threadIdx.x; i < PARTS * BN_RSPAN; i += BN_P2_BLOCK) s_acc[i] = 0ull;
        __syncthreads();
        const uint32_t it = s_it;
        if (it >= nitems) break;
        const uint32_t* I = items + 8 * (uint64_t)it;
        const uint32_t b = I[0], q0 = I[1], q1 = I[2], rows = I[4];
        const uint64_t r0 = I[3];
        bn_gather_windows<CT, CT>(bdst, term, (uint64_t)q0 * BN_CH, (uint64_t)q1 * BN_CH, wave, lane, scale, s_acc);
        __syncthreads();
        {
            // The epilogue's loads for four of the thread's rows first, then the stores.
            double delta = 0.0, dm = 0.0;
            for (int k0r = 0; k0r < BN_EP; k0r += 4) {
                double cur[4], sc[4], bb[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t i = threadIdx.x + (k0r + k) * BN_P2_BLOCK;
                    const uint64_t u = r0 + (i < rows ? i : 0);  // no branch: row r0 always exists
                    cur[k] = (double)rank[u];
                    sc[k] = node_scale<WEIGHTED>(out_deg, out_str, u);
                    bb[k] = base;
                    if constexpr (CORR) bb[k] = base + (double)corr[u];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t i = threadIdx.x + (k0r + k) * BN_P2_BLOCK;
                    if (i < rows) {
                        double acc;
                        if constexpr (PARTS == 2)
                            acc = fx_value(s_acc[i], (long long)s_acc[BN_RSPAN + i]);
                        else
                            acc = __builtin_ldexp((double)(long long)s_acc[i], -K);
                        pr_finish<RT, CT, DM>(r0 + i, acc, bb[k], damping, sc[k], cur[k], next, cnext, delta, dm);
                    }
                }
            }
            const double bd = block_sum(delta, s_red);
            const double bm = block_sum(dm, s_red);
            if (threadIdx.x == 0) part[b] = PrPart{bd, bm};
        }
        __syncthreads();
    }
}

// Free-standing exact sum (ursa_hip_exact_sum): one workgroup adds the fixed
// point parts of every x[i] into one accumulator.
__global__ __launch_bounds__(256) void k_fx_sum(const double* __restrict__ x, uint64_t n, double* out) {
    __shared__ unsigned long long s_lo, s_hi;
    if (threadIdx.x == 0) {
        s_lo = 0ull;
        s_hi = 0ull;
    }
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        unsigned long long lo;
        long long hi;
        fx_split(x[i], lo, hi);
        atomicAdd(&s_lo, lo);
        atomicAdd(&s_hi, (unsigned long long)hi);
    }
    __syncthreads();
    if (threadIdx.x == 0) *out = fx_value(s_lo, (long long)s_hi);
}

// ---- binned layout (bn_prepare) ----
__global__ void k_bn_keys(const uint32_t* __restrict__ tgt, uint64_t m, uint32_t* __restrict__ key,
                          uint32_t* __restrict__ iota) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
        key[k] = tgt[k] >> BN_SBITS;
        iota[k] = (uint32_t)k;
    }
}

// row_bin[r] = the bin holding row r (bstart: first row of each bin, nrb + 1 values).
__global__ void k_bn_rowbin(const uint32_t* __restrict__ bstart, uint64_t nrb, uint64_t n,
                            uint32_t* __restrict__ row_bin) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = nrb;  // bstart[lo] <= r < bstart[hi]
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (bstart[mid] <= r)
                lo = mid;
            else
                hi = mid;
        }
        row_bin[r] = (uint32_t)lo;
    }
}

// Unpadded binned slot j holds in-slot perm[j]; tile id = block * nrb + bin.
__global__ void k_bn_tile(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ owner,
                          const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ row_bin, uint64_t m,
                          uint64_t nrb, uint32_t* __restrict__ tile) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = perm[j];
        tile[j] = (uint32_t)((uint64_t)(tgt[k] >> BN_SBITS) * nrb + row_bin[owner[k]]);
    }
}

// toff[t] = first unpadded binned slot whose tile id is >= t, for t in [0, T];
// the tile ids of the binned slots never decrease.
__global__ void k_bn_toff(const uint32_t* __restrict__ tile, uint64_t m, uint64_t T, uint32_t* __restrict__ toff) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j <= m; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lo = j == 0 ? 0 : (uint64_t)tile[j - 1] + 1;
        const uint64_t hi = j == m ? T : (uint64_t)tile[j];
        for (uint64_t t = lo; t <= hi; ++t) toff[t] = (uint32_t)j;
    }
}

// Padded tile lengths (whole chunks); plen[T] = 0 so the scan's last value is the total.
__global__ void k_bn_plen(const uint32_t* __restrict__ toff, uint64_t T, uint32_t* __restrict__ plen) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t <= T; t += (uint64_t)gridDim.x * blockDim.x)
        plen[t] = t < T ? (toff[t + 1] - toff[t] + BN_CH - 1) / BN_CH * BN_CH : 0u;
}

// Padding slots: source 0, row BN_PAD, edge id 0 (both orders have mp slots).
__global__ void k_bn_pad_fill(uint64_t mp, uint16_t* __restrict__ bsrc, uint16_t* __restrict__ bdst,
                              uint32_t* __restrict__ beid) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < mp; p += (uint64_t)gridDim.x * blockDim.x) {
        bsrc[p] = 0;
        bdst[p] = (uint16_t)BN_PAD;
        beid[p] = 0;
    }
}

// Unpadded slot j (tile t) goes to poff[t] + (j - toff[t]) in the
// source-major order (source index, edge id) and to bpos[b * nsb + c] +
// (j - toff[t]) in the bin-major order (row index), t = c * nrb + b.
__global__ void k_bn_place(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ owner,
                           const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ eid,
                           const uint32_t* __restrict__ tile, const uint32_t* __restrict__ toff,
                           const uint64_t* __restrict__ poff, const uint64_t* __restrict__ bpos,
                           const uint32_t* __restrict__ bstart, uint64_t nsb, uint64_t nrb, uint64_t m,
                           uint16_t* __restrict__ bsrc, uint16_t* __restrict__ bdst, uint32_t* __restrict__ beid) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = perm[j], t = tile[j];
        const uint64_t off = j - toff[t], c = t / nrb, b = t % nrb;
        bsrc[poff[t] + off] = (uint16_t)(tgt[k] & (BN_SSPAN - 1));
        beid[poff[t] + off] = eid[k];
        bdst[bpos[b * nsb + c] + off] = (uint16_t)(owner[k] - bstart[b]);
    }
}

// Padded tile lengths in bin-major order (i = b * nsb + c); blen[T] = 0.
__global__ void k_bn_blen(const uint32_t* __restrict__ plen, uint64_t nsb, uint64_t nrb, uint32_t* __restrict__ blen) {
    const uint64_t T = nsb * nrb;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= T; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = i / nsb, c = i % nsb;
        blen[i] = i < T ? plen[c * nrb + b] : 0u;
    }
}

// Source-major chunk q -> its bin-major chunk: q lies in tile t (poff[t] <=
// q * BN_CH < poff[t + 1]), whose bin-major run starts at bpos[b * nsb + c].
__global__ void k_bn_places(const uint64_t* __restrict__ poff, const uint64_t* __restrict__ bpos, uint64_t nsb,
                            uint64_t nrb, uint64_t Q, uint32_t* __restrict__ place) {
    const uint64_t T = nsb * nrb;
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < Q; q += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t p = q * BN_CH;
        uint64_t lo = 0, hi = T;  // poff[lo] <= p < poff[hi]
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (poff[mid] <= p)
                lo = mid;
            else
                hi = mid;
        }
        const uint64_t c = lo / nrb, b = lo % nrb;
        place[q] = (uint32_t)((bpos[b * nsb + c] + (p - poff[lo])) / BN_CH);
    }
}

}  // namespace

// PageRank's working copy of the in-CSR, built once per topology (layout at
// the top of this file). A row's sum runs in ascending new source id, not in
// the in-CSR's order: the same terms, rounded in another order.
static void pr_prepare(ursa_hip_topo* t, const DevCsr& in, hipStream_t s) {
    NoArena keep_out_of_arena;
    std::lock_guard<std::mutex> lk(t->cache_mu);
    if (t->pr_ready) return;
    const uint64_t n = t->n, m = t->m;
    t->pr_old_of_new.alloc(n);
    t->pr_new_of_old.alloc(n);
    t->pr_off.alloc(n + 1);
    t->pr_tgt.alloc(m ? m : 1);
    t->pr_eid.alloc(m ? m : 1);
    t->pr_outdeg.alloc(n);
    uint64_t nz = 0;
    {
        DevBuf<uint64_t> keys(n), skeys(n), dnz(1);
        DevBuf<uint32_t> iota(n), rdeg(n);
        k_rl_keys<<<blocks_for(n, 256), 256, 0, s>>>(t->out_deg.get(), in.off.get(), n, keys.get(), iota.get());
        check_launch();
        prim::sort_pairs_u64_u32(keys.get(), skeys.get(), iota.get(), t->pr_old_of_new.get(), n, 33, s);
        k_rl_invert<<<blocks_for(n, 256), 256, 0, s>>>(t->pr_old_of_new.get(), n, in.off.get(), t->out_deg.get(),
                                                       t->pr_new_of_old.get(), rdeg.get(), t->pr_outdeg.get());
        check_launch();
        prim::exclusive_sum_u32(rdeg.get(), t->pr_off.get(), n, s);
        HIP_OK(hipMemsetAsync(dnz.get(), 0, sizeof(uint64_t), s));
        k_rl_nz<<<blocks_for(n, 256), 256, 0, s>>>(rdeg.get(), n, dnz.get());
        check_launch();
        HIP_OK(hipMemcpyAsync(&nz, dnz.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    if (m) {
        DevBuf<uint64_t> key(m), skey(m);
        {
            DevBuf<uint32_t> owner(m);
            slot_owners(n, in, owner.get(), s);
            k_rl_slot_keys<<<blocks_for(m, 256), 256, 0, s>>>(owner.get(), in.tgt.get(), t->pr_new_of_old.get(), m,
                                                              key.get());
            check_launch();
        }
        int bits = 1;
        while (bits < 32 && (n - 1) >> bits) ++bits;
        prim::sort_pairs_u64_u32(key.get(), skey.get(), in.eid.get(), t->pr_eid.get(), m, 32 + bits, s);
        k_rl_low_words<<<blocks_for(m, 256), 256, 0, s>>>(skey.get(), m, t->pr_tgt.get());
        check_launch();
    }
    const uint64_t nchunks = (m + PR_CHUNK - 1) / PR_CHUNK;
    t->pr_head.alloc(nchunks * PR_WORDS + 1);
    HIP_OK(hipMemsetAsync(t->pr_head.get(), 0, (nchunks * PR_WORDS + 1) * sizeof(uint32_t), s));
    t->pr_crow.alloc(nchunks ? nchunks : 1);
    t->pr_cflag.alloc(nchunks ? nchunks : 1);
    if (nz) {
        k_pr_heads<<<blocks_for(nz, 256), 256, 0, s>>>(t->pr_off.get(), nz, t->pr_head.get());
        check_launch();
    }
    if (nchunks) {
        k_pr_chunks<<<blocks_for(nchunks, 256), 256, 0, s>>>(t->pr_off.get(), nz, t->pr_head.get(), m, nchunks,
                                                             t->pr_crow.get(), t->pr_cflag.get());
        check_launch();
    }
    HIP_OK(hipStreamSynchronize(s));
    t->pr_nz = nz;
    t->pr_chunks = nchunks;
    t->pr_ready = true;
}

uint64_t pr_max_in_degree(ursa_hip_topo* t, hipStream_t s) {
    std::lock_guard<std::mutex> lk(t->cache_mu);
    if (t->pr_max_in == ~0ull) {
        std::vector<uint32_t> h(t->n + 1);
        HIP_OK(hipMemcpyAsync(h.data(), t->pr_off.get(), (t->n + 1) * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        uint64_t mx = 0;
        for (uint64_t r = 0; r < t->n; ++r) mx = std::max<uint64_t>(mx, h[r + 1] - h[r]);
        t->pr_max_in = mx;
    }
    return t->pr_max_in;
}

// Bins hold at most BN_RSPAN rows and, on the heaviest rows, fewer: with m
// in-slots there are at most n / BN_RSPAN + m / BN_T2 + 1 of them.
bool bn_fits(ursa_hip_topo* t, hipStream_t s) {
    const uint64_t n = t->n, m = t->m;
    const uint64_t blocks = (n + BN_SSPAN - 1) / BN_SSPAN, bins_max = n / BN_RSPAN + m / BN_T2 + 1;
    return blocks * bins_max <= BN_MAX_TILES && pr_max_in_degree(t, s) < BN_MAX_ROW;
}

// The binned layout (binned.hpp) is built once per topology after pr_prepare.
// Work items are cut on the host from the tile tables.
void bn_prepare(ursa_hip_topo* t, hipStream_t s) {
    NoArena keep_out_of_arena;
    std::lock_guard<std::mutex> lk(t->cache_mu);
    if (t->bn_ready) return;
    const uint64_t n = t->n, m = t->m;
    // Bins: consecutive rows, at most BN_RSPAN of them and at most BN_T2
    // in-slots (a longer row alone), from the renumbered row offsets.
    std::vector<uint32_t> h_off(n + 1), h_bstart;
    HIP_OK(hipMemcpyAsync(h_off.data(), t->pr_off.get(), (n + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (uint64_t r = 0; r < n;) {
        h_bstart.push_back((uint32_t)r);
        uint64_t e = r + 1;
        while (e < n && e - r < BN_RSPAN && (uint64_t)h_off[e + 1] - h_off[r] <= BN_T2) ++e;
        r = e;
    }
    h_bstart.push_back((uint32_t)n);
    const uint64_t nsb = (n + BN_SSPAN - 1) / BN_SSPAN, nrb = h_bstart.size() - 1, T = nsb * nrb;
    t->bn_bstart.alloc(nrb + 1);
    HIP_OK(hipMemcpyAsync(t->bn_bstart.get(), h_bstart.data(), (nrb + 1) * 4, hipMemcpyHostToDevice, s));
    DevBuf<uint32_t> toff(T + 1), perm, tile, owner;
    if (m) {
        perm.alloc(m);
        tile.alloc(m);
        owner.alloc(m);
        {
            DevBuf<uint32_t> key(m), skey(m), iota(m);
            k_bn_keys<<<blocks_for(m, 256), 256, 0, s>>>(t->pr_tgt.get(), m, key.get(), iota.get());
            check_launch();
            prim::sort_pairs_u32(key.get(), skey.get(), iota.get(), perm.get(), m, bits_for(nsb), s);
        }
        slot_owners(n, m, t->pr_off.get(), owner.get(), s);
        DevBuf<uint32_t> row_bin(n);
        k_bn_rowbin<<<blocks_for(n, 256), 256, 0, s>>>(t->bn_bstart.get(), nrb, n, row_bin.get());
        k_bn_tile<<<blocks_for(m, 256), 256, 0, s>>>(perm.get(), owner.get(), t->pr_tgt.get(), row_bin.get(), m, nrb,
                                                     tile.get());
        k_bn_toff<<<blocks_for(m + 1, 256), 256, 0, s>>>(tile.get(), m, T, toff.get());
    } else {
        k_bn_toff<<<1, 256, 0, s>>>(nullptr, 0, T, toff.get());
    }
    check_launch();
    // Padded tile lengths; their running sums in source-major order (poff,
    // indexed by t = c * nrb + b) and in bin-major order (bpos, by b * nsb + c).
    DevBuf<uint32_t> plen(T + 1);
    DevBuf<uint64_t> poff(T + 1), bpos(T + 1);
    k_bn_plen<<<blocks_for(T + 1, 256), 256, 0, s>>>(toff.get(), T, plen.get());
    check_launch();
    prim::exclusive_sum_u32_to_u64(plen.get(), poff.get(), T + 1, s);
    {
        DevBuf<uint32_t> blen(T + 1);
        k_bn_blen<<<blocks_for(T + 1, 256), 256, 0, s>>>(plen.get(), nsb, nrb, blen.get());
        check_launch();
        prim::exclusive_sum_u32_to_u64(blen.get(), bpos.get(), T + 1, s);
    }
    std::vector<uint64_t> h_poff(T + 1), h_bpos(T + 1);
    HIP_OK(hipMemcpyAsync(h_poff.data(), poff.get(), (T + 1) * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(h_bpos.data(), bpos.get(), (T + 1) * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const uint64_t mp = h_poff[T], Q = mp / BN_CH;
    URSA_REQUIRE(Q <= 0xFFFFFFFFull, URSA_HIP_ERANGE, "binned PageRank layout exceeds 2^32 chunks");
    t->bn_src.alloc(std::max<uint64_t>(mp, BN_CH));
    t->bn_dst.alloc(std::max<uint64_t>(mp, BN_CH));
    t->bn_eid.alloc(std::max<uint64_t>(mp, BN_CH));
    t->bn_place.alloc(std::max<uint64_t>(Q, 1));
    if (mp) {
        k_bn_pad_fill<<<blocks_for(mp, 256), 256, 0, s>>>(mp, t->bn_src.get(), t->bn_dst.get(), t->bn_eid.get());
        k_bn_place<<<blocks_for(m, 256), 256, 0, s>>>(perm.get(), owner.get(), t->pr_tgt.get(), t->pr_eid.get(),
                                                      tile.get(), toff.get(), poff.get(), bpos.get(),
                                                      t->bn_bstart.get(), nsb, nrb, m, t->bn_src.get(),
                                                      t->bn_dst.get(), t->bn_eid.get());
        k_bn_places<<<blocks_for(Q, 256), 256, 0, s>>>(poff.get(), bpos.get(), nsb, nrb, Q, t->bn_place.get());
        check_launch();
    }
    HIP_OK(hipStreamSynchronize(s));

    // Phase 1: each source block's chunks in pieces of at most BN_T1 slots.
    std::vector<uint32_t> it1;
    const uint64_t c1 = BN_T1 / BN_CH;
    for (uint64_t c = 0; c < nsb; ++c) {
        const uint64_t a = h_poff[c * nrb] / BN_CH, e = h_poff[(c + 1) * nrb] / BN_CH;
        for (uint64_t q = a; q < e; q += c1)
            it1.insert(it1.end(), {(uint32_t)c, (uint32_t)q, (uint32_t)std::min<uint64_t>(e, q + c1), 0u});
    }
    // Phase 2: one item per bin (bin, chunks, first row, rows), largest first.
    struct It2 {
        uint32_t b, q0, q1, r0, rows;
    };
    std::vector<It2> items;
    for (uint64_t b = 0; b < nrb; ++b)
        items.push_back({(uint32_t)b, (uint32_t)(h_bpos[b * nsb] / BN_CH), (uint32_t)(h_bpos[(b + 1) * nsb] / BN_CH),
                         h_bstart[b], h_bstart[b + 1] - h_bstart[b]});
    std::stable_sort(items.begin(), items.end(),
                     [](const It2& x, const It2& y) { return x.q1 - x.q0 > y.q1 - y.q0; });
    std::vector<uint32_t> it2;
    for (const It2& x : items) it2.insert(it2.end(), {x.b, x.q0, x.q1, x.r0, x.rows, 0u, 0u, 0u});
    t->bn_item1.alloc(std::max<size_t>(it1.size(), 4));
    t->bn_item2.alloc(it2.size());
    if (!it1.empty())
        HIP_OK(hipMemcpyAsync(t->bn_item1.get(), it1.data(), it1.size() * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(t->bn_item2.get(), it2.data(), it2.size() * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));
    t->bn_slots = mp;
    t->bn_n1 = it1.size() / 4;
    t->bn_n2 = items.size();
    t->bn_bins = nrb;
    t->bn_ready = true;
}

void pagerank_prepare(ursa_hip_topo* t, hipStream_t s) { pr_prepare(t, t->incoming(s), s); }

// Solve PageRank (f64 solve, or with `mixed` the refinement solve described at
// the top of this file). Every sweep of either kind counts toward max_iter,
// and the answer is always the output of an f64 sweep (or the uniform start
// when max_iter == 0).
void pagerank_run(ursa_hip_topo* t, const double* weights_user, double damping, uint32_t max_iter, double tol,
                  double* out_user, uint32_t* iters_out, bool mixed, uint32_t* f32_iters_out) {
    hipStream_t s = stream();
    const uint64_t n = t->n, m = t->m;
    if (iters_out) *iters_out = 0;
    if (f32_iters_out) *f32_iters_out = 0;
    if (n == 0) return;
    URSA_REQUIRE(out_user != nullptr, URSA_HIP_EINVAL, "out is NULL");
    const double nf = (double)n;
    DevOut<double> out(out_user, n);
    if (max_iter == 0) {  // pagerank.rs:73-75: the loop body never runs
        k_fill<<<blocks_for(n, 256), 256, 0, s>>>(out.p, n, 1.0 / nf);
        check_launch();
        out.finish(s);
        return;
    }
    if (!(tol < PR_F32_TOL) || max_iter <= PR_F32_MIN_ITER) mixed = false;
    pr_prepare(t, t->incoming(s), s);
    // Sweep schedule: binned (streamed terms, exact row sums) or the pull sweep.
    const uint64_t bn_blocks = (n + BN_SSPAN - 1) / BN_SSPAN;
    const int sched = g_pr_schedule.load();
    const bool binned = (sched == 2 || (sched == 0 && bn_blocks >= BN_MIN_BLOCKS)) && bn_fits(t, s);
    if (binned) bn_prepare(t, s);
    const uint64_t nchunks = t->pr_chunks, nz = t->pr_nz;
    const uint32_t* tgt = t->pr_tgt.get();
    const uint32_t* head = t->pr_head.get();
    const uint32_t* crow = t->pr_crow.get();
    const uint8_t* cflag = t->pr_cflag.get();

    const bool weighted = weights_user != nullptr;
    DevIn<double> w(weights_user, m, s);
    DevBuf<double> out_str, w64;
    DevBuf<float> w32;
    bool use_f32w = false;
    if (weighted) {
        DevBuf<double> os_old(n);
        out_str.alloc(n);
        k_out_strength<<<blocks_for(n, 256), 256, 0, s>>>(t->out.off.get(), t->out.eid.get(), w.p, n, os_old.get());
        k_out_strength_heavy<<<blocks_for((n + 63) / 64, 4, 8192), 256, 0, s>>>(t->out.off.get(), t->out.eid.get(), w.p,
                                                                               n, os_old.get());
        k_gather_f64<<<blocks_for(n, 256), 256, 0, s>>>(os_old.get(), t->pr_old_of_new.get(), n, out_str.get());
        check_launch();
        HIP_OK(hipStreamSynchronize(s));
        // Weights in the sweep's slot order (the binned order includes padding
        // slots, whose edge id 0 is valid whenever there are edges).
        const uint64_t wn = binned ? std::max<uint64_t>(t->bn_slots, BN_CH) : m;
        w64.alloc(wn);
        w32.alloc(wn);
        DevBuf<uint32_t> inexact(1);
        HIP_OK(hipMemsetAsync(inexact.get(), 0, sizeof(uint32_t), s));
        const uint64_t wcount = binned ? t->bn_slots : m;
        if (m && wcount) {
            k_permute_weights<<<blocks_for(wcount, 256), 256, 0, s>>>(binned ? t->bn_eid.get() : t->pr_eid.get(), w.p,
                                                                      wcount, w64.get(), w32.get(), inexact.get());
            check_launch();
        }
        uint32_t h = 0;
        HIP_OK(hipMemcpyAsync(&h, inexact.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        use_f32w = (h == 0);
        // The f32 sweeps always read f32 weights; keep both when f64 ones are needed.
        if (use_f32w) w64.reset();
        else if (!mixed) w32.reset();
    }

    // Launch shape: with the LDS copy, two persistent workgroups per CU;
    // without it (graphs too small to fill the copy) one workgroup per group
    // of four chunks.
    const bool hot64 = n >= (uint64_t)(PR_HOT_BYTES / 8), hot32 = n >= (uint64_t)(PR_HOT_BYTES / 4);
    auto grid_for = [&](bool hot) -> unsigned {
        const uint64_t waves = hot ? PR_HOT_BLOCK / 64 : PR_PLAIN_BLOCK / 64;
        const uint64_t groups = std::max<uint64_t>((nchunks + waves - 1) / waves, 1);
        return (unsigned)(hot ? std::min<uint64_t>(groups, 2ull * cu_count()) : groups);
    };
    const unsigned g64 = grid_for(hot64), g32 = grid_for(hot32);
    const uint64_t nparts = std::max(g64, g32);
    const uint64_t fix_items = std::max<uint64_t>(nchunks + (n - nz), nparts);
    const unsigned fix_blocks = (unsigned)((fix_items + PR_FIX - 1) / PR_FIX);

    DevBuf<double> rank[2] = {DevBuf<double>(n), DevBuf<double>(n)};
    DevBuf<double> contrib[2] = {DevBuf<double>(n), DevBuf<double>(n)};
    DevBuf<float> x32[2], c32[2], r32;
    if (mixed) {
        for (int i = 0; i < 2; ++i) {
            x32[i].alloc(n);
            c32[i].alloc(n);
        }
        r32.alloc(n);
    }
    DevBuf<double> lead(std::max<uint64_t>(nchunks, 1)), tail(std::max<uint64_t>(nchunks, 1));
    DevBuf<PrPart> part(nparts), fpart(std::max<uint64_t>(fix_blocks, (n + PR_FIX - 1) / PR_FIX));
    DevBuf<PrCtl> ctl(1);
    const unsigned init_blocks = blocks_for(n, 256, 1u << 30);
    // Binned sweep state: the term arrays, per-bin partials and the two
    // work-claim counters (cleared here, then by the kernels themselves, so
    // recorded sweeps replay without a memset).
    const uint64_t nrb = binned ? t->bn_bins : 0;
    DevBuf<double> term64;
    DevBuf<float> term32;
    DevBuf<PrPart> bpart;
    DevBuf<uint32_t> bctr;
    unsigned g1 = 1, g2 = 1;
    if (binned) {
        term64.alloc(std::max<uint64_t>(t->bn_slots, BN_CH));
        if (mixed) term32.alloc(std::max<uint64_t>(t->bn_slots, BN_CH));
        bpart.alloc(nrb);
        bctr.alloc(2);
        HIP_OK(hipMemsetAsync(bctr.get(), 0, 8, s));
        g1 = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(t->bn_n1, 4ull * cu_count()));
        g2 = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(t->bn_n2, 2ull * cu_count()));
    }
    PrPart* sweep_part = binned ? bpart.get() : fpart.get();
    const uint64_t sweep_parts = binned ? nrb : fix_blocks;
    const uint32_t* od = t->pr_outdeg.get();
    const double* os = weighted ? out_str.get() : nullptr;

    // The dangling nodes in ascending original order, as renumbered slots: the
    // order of the dangling-mass sum (pagerank.rs:76-78, 109-111).
    uint64_t nd = 0;
    DevBuf<uint32_t> dang(n);
    {
        DevBuf<uint8_t> flags(n);
        DevBuf<uint64_t> cnt(1);
        if (weighted)
            k_pr_dang_flags<true><<<init_blocks, 256, 0, s>>>(n, t->pr_new_of_old.get(), od, os, flags.get());
        else
            k_pr_dang_flags<false><<<init_blocks, 256, 0, s>>>(n, t->pr_new_of_old.get(), od, os, flags.get());
        check_launch();
        timed("pagerank_dm_list", 9.0 * (double)n, s, [&] {
            prim::select_flagged_u32(t->pr_new_of_old.get(), flags.get(), dang.get(), cnt.get(), n, s);
        });
        HIP_OK(hipMemcpyAsync(&nd, cnt.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    const uint64_t nch = (nd + OS_CHUNK - 1) / OS_CHUNK, nch1 = nch ? nch : 1;
    const unsigned dm_blocks = (unsigned)((nch + OS_BUILD_THREADS / 64 - 1) / (OS_BUILD_THREADS / 64));
    DevBuf<double> xd(nd ? nd : 1), cs(nch1);
    DevBuf<uint64_t> cq(nch1 * OS_BINS);
    DevBuf<uint32_t> cm(nch1);
    // base[parity] from the ordered dangling sum of iterate r.
    auto dangling_base = [&](const double* r, int parity) {
        if (nch)
            timed("pagerank_dm_build", 20.0 * (double)nd, s, [&] {
                k_pr_dm_build<<<dm_blocks, OS_BUILD_THREADS, 0, s>>>(r, dang.get(), nd, ctl.get(), xd.get(), cq.get(),
                                                                      cm.get(), cs.get());
            });
        timed("pagerank_dm_walk", 0.0, s, [&] {
            k_pr_dm_walk<<<1, OS_WALK_THREADS, 0, s>>>(xd.get(), nd, cq.get(), cm.get(), cs.get(), ctl.get(), parity,
                                                       damping, nf);
        });
        check_launch();
    };
    // The first iterate is 1/n everywhere, so its dangling mass is 1/n added nd
    // times in order: closed form on the host (every later one on the device).
    const double dm0 = repeated_sum(1.0 / nf, nd);
    const double base0 = (1.0 - damping) / nf + damping * dm0 / nf;  // pagerank.rs:79
    const int32_t e0 = dm0 > 0.0 ? (int32_t)std::ilogb(dm0) - 13 : 0;
    k_pr_ctl_init<<<1, 1, 0, s>>>(ctl.get(), e0, base0);
    check_launch();

    const double node_bytes = weighted ? 36.0 : 32.0;
    const double f64_edge = weighted ? (use_f32w ? 16.0 : 20.0) : 12.0;
    const double f32_edge = weighted ? 12.0 : 8.0;
    const double f32_node = weighted ? 24.0 : 20.0;

    // f64 sweep of rank[p] (device-driven base ctl->base[p] when dev, else base_arg).
    auto sweep64 = [&](int p, bool dev, double base_arg) {
        const int par = dev ? p : -1;
        const double* rk = rank[p].get();
        double* nx = rank[p ^ 1].get();
        const double* cb = contrib[p].get();
        double* cn = contrib[p ^ 1].get();
        auto go = [&](auto hot_tag, auto w_tag, auto wt_tag) {
            constexpr int HOT = decltype(hot_tag)::value;
            constexpr bool WT = decltype(wt_tag)::value;
            using W = typename decltype(w_tag)::type;
            constexpr int BLOCK = HOT ? PR_HOT_BLOCK : PR_PLAIN_BLOCK;
            const W* wp = nullptr;
            if constexpr (WT) {
                if constexpr (sizeof(W) == 4) wp = w32.get();
                else wp = w64.get();
            }
            if (binned) {
                if constexpr (HOT == 0) {
                    k_bn_scatter<double, W, WT><<<g1, BN_P1_BLOCK, 0, s>>>(
                        t->bn_src.get(), t->bn_place.get(), wp, t->bn_item1.get(), (uint32_t)t->bn_n1, n, cb,
                        term64.get(), bctr.get(), ctl.get());
                    k_bn_gather<double, double, WT, false, false><<<g2, BN_P2_BLOCK, 0, s>>>(
                        t->bn_dst.get(), term64.get(), t->bn_item2.get(), (uint32_t)t->bn_n2, n,
                        od, os, rk, nx, cn, nullptr, bpart.get(), bctr.get(), ctl.get(), par, base_arg, damping, 0);
                }
                return;
            }
            k_pr_sweep<double, double, W, WT, false, false, BLOCK, HOT><<<g64, BLOCK, 0, s>>>(
                tgt, wp, head, crow, m, nchunks, od, os, rk, cb, nx, cn, nullptr, lead.get(), tail.get(), part.get(),
                ctl.get(), par, base_arg, damping);
        };
        struct F { using type = float; };
        struct D { using type = double; };
        timed("pagerank_spmv", f64_edge * (double)m + node_bytes * (double)n, s, [&] {
            using H0 = std::integral_constant<int, 0>;
            using H1 = std::integral_constant<int, PR_HOT_BYTES / 8>;
            using T = std::true_type;
            using Fl = std::false_type;
            if (!weighted) {
                if (hot64 && !binned) go(H1{}, D{}, Fl{}); else go(H0{}, D{}, Fl{});
            } else if (use_f32w) {
                if (hot64 && !binned) go(H1{}, F{}, T{}); else go(H0{}, F{}, T{});
            } else {
                if (hot64 && !binned) go(H1{}, D{}, T{}); else go(H0{}, D{}, T{});
            }
        });
        check_launch();
        if (binned) return;
        timed("pagerank_fixup", 0.0, s, [&] {
            if (weighted)
                k_pr_fixup<double, double, true, false, false><<<fix_blocks, PR_FIX, 0, s>>>(
                    cflag, crow, t->pr_off.get(), nchunks, nz, n, od, os, rk, nx, cn, nullptr, lead.get(), tail.get(),
                    part.get(), g64,
                    fpart.get(), ctl.get(), par, base_arg, damping);
            else
                k_pr_fixup<double, double, false, false, false><<<fix_blocks, PR_FIX, 0, s>>>(
                    cflag, crow, t->pr_off.get(), nchunks, nz, n, od, os, rk, nx, cn, nullptr, lead.get(), tail.get(),
                    part.get(), g64,
                    fpart.get(), ctl.get(), par, base_arg, damping);
        });
        check_launch();
    };
    // f32 sweep xin -> xout (host-driven base; corr = residual of a correction sweep).
    auto sweep32 = [&](const float* xin, const float* cin_, float* xout, float* cout, const float* corr,
                       double base_arg, double l1_bound) {
        // Binned f32 sweeps: fixed-point scale 2^K with K = 60 - ceil(log2 B), B a
        // bound on the L1 norm of the iterate (so on every sum of its terms).
        int fx_k = 149;
        if (l1_bound > 0.0) fx_k = std::min(149, std::max(1, 60 - (int)std::ceil(std::log2(l1_bound))));
        auto go = [&](auto hot_tag, auto wt_tag, auto corr_tag) {
            constexpr int HOT = decltype(hot_tag)::value;
            constexpr bool WT = decltype(wt_tag)::value;
            constexpr bool CORR = decltype(corr_tag)::value;
            constexpr int BLOCK = HOT ? PR_HOT_BLOCK : PR_PLAIN_BLOCK;
            if (binned) {
                if constexpr (HOT == 0) {
                    k_bn_scatter<float, float, WT><<<g1, BN_P1_BLOCK, 0, s>>>(
                        t->bn_src.get(), t->bn_place.get(), WT ? w32.get() : nullptr, t->bn_item1.get(),
                        (uint32_t)t->bn_n1, n, cin_, term32.get(), bctr.get(), ctl.get());
                    k_bn_gather<float, float, WT, CORR, true><<<g2, BN_P2_BLOCK, 0, s>>>(
                        t->bn_dst.get(), term32.get(), t->bn_item2.get(), (uint32_t)t->bn_n2, n,
                        od, os, xin, xout, cout, corr, bpart.get(), bctr.get(), ctl.get(), -1, base_arg, damping,
                        fx_k);
                }
                return;
            }
            k_pr_sweep<float, float, float, WT, CORR, true, BLOCK, HOT><<<g32, BLOCK, 0, s>>>(
                tgt, WT ? w32.get() : nullptr, head, crow, m, nchunks, od, os, xin, cin_, xout, cout, corr,
                lead.get(), tail.get(), part.get(), ctl.get(), -1, base_arg, damping);
            k_pr_fixup<float, float, WT, CORR, true><<<fix_blocks, PR_FIX, 0, s>>>(
                cflag, crow, t->pr_off.get(), nchunks, nz, n, od, os, xin, xout, cout, corr, lead.get(), tail.get(),
                part.get(), g32,
                fpart.get(), ctl.get(), -1, base_arg, damping);
        };
        using H0 = std::integral_constant<int, 0>;
        using H1 = std::integral_constant<int, PR_HOT_BYTES / 4>;
        using T = std::true_type;
        using Fl = std::false_type;
        const bool c = corr != nullptr;
        timed(c ? "pagerank_spmv_f32c" : "pagerank_spmv_f32",
              f32_edge * (double)m + (f32_node + (c ? 4.0 : 0.0)) * (double)n, s, [&] {
                  if (hot32 && !binned) {
                      if (weighted) { if (c) go(H1{}, T{}, T{}); else go(H1{}, T{}, Fl{}); }
                      else { if (c) go(H1{}, Fl{}, T{}); else go(H1{}, Fl{}, Fl{}); }
                  } else {
                      if (weighted) { if (c) go(H0{}, T{}, T{}); else go(H0{}, T{}, Fl{}); }
                      else { if (c) go(H0{}, Fl{}, T{}); else go(H0{}, Fl{}, Fl{}); }
                  }
              });
        check_launch();
    };
    auto reduce = [&](uint32_t mi, double tol_it) {
        k_pr_reduce<<<1, 1024, 0, s>>>(sweep_part, sweep_parts, ctl.get(), mi, tol_it);
        check_launch();
    };
    auto read_ctl = [&]() {
        PrCtl h{};
        HIP_OK(hipMemcpyAsync(&h, ctl.get(), sizeof(PrCtl), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return h;
    };

    uint32_t iters = 0, f32_iters = 0;
    int result = 0;  // rank[result] holds the answer
    if (!mixed) {
        if (weighted)
            k_pr_init<double, double, true><<<init_blocks, 256, 0, s>>>(n, nf, od, os, rank[0].get(), contrib[0].get());
        else
            k_pr_init<double, double, false><<<init_blocks, 256, 0, s>>>(n, nf, od, os, rank[0].get(), contrib[0].get());
        check_launch();
        // Iterations are queued in batches; with per-kernel timing on, one at a
        // time so no launch after convergence is booked. A full batch of 8 is
        // recorded once per start parity as a HIP graph and replayed.
        const uint32_t batch = prof_enabled() ? 1 : 8;
        CapturedLaunches graphs[2];
        uint32_t it = 0;
        PrCtl h = read_ctl();
        while (it < max_iter && !h.done) {
            const uint32_t it_end = std::min<uint32_t>(max_iter, it + batch);
            const uint32_t it0 = it;
            auto launches = [&] {
                for (uint32_t i = it0; i < it_end; ++i) {
                    const int p = (int)(i & 1);
                    sweep64(p, true, 0.0);
                    reduce(max_iter, tol);
                    dangling_base(rank[p ^ 1].get(), p ^ 1);
                }
            };
            if (it_end - it0 == 8)
                graphs[it0 & 1].run(s, launches);
            else
                launches();
            it = it_end;
            h = read_ctl();
        }
        iters = h.iters;
        result = (int)(iters & 1);
    } else {
        constexpr uint32_t NO_LIMIT = 0xFFFFFFFFu;
        // f32 sweeps from the uniform start.
        if (weighted)
            k_pr_init<float, float, true><<<init_blocks, 256, 0, s>>>(n, nf, od, os, x32[0].get(), c32[0].get());
        else
            k_pr_init<float, float, false><<<init_blocks, 256, 0, s>>>(n, nf, od, os, x32[0].get(), c32[0].get());
        check_launch();
        double dm = (double)nd * (double)(float)(1.0 / nf);
        double prev = 0.0;
        int p = 0;
        for (;;) {
            // The iterate is a probability vector: its L1 norm is 1 up to f32 rounding.
            sweep32(x32[p].get(), c32[p].get(), x32[p ^ 1].get(), c32[p ^ 1].get(), nullptr,
                    (1.0 - damping) / nf + damping * dm / nf, 1.01);
            reduce(NO_LIMIT, -1.0);
            const PrCtl h = read_ctl();
            ++iters;
            ++f32_iters;
            p ^= 1;
            dm = h.dm;
            const bool stalled = f32_iters >= 3 && h.delta > 0.7 * prev;
            prev = h.delta;
            if (h.delta < PR_F32_TOL || stalled || iters + 2 >= max_iter) break;
        }
        if (weighted)
            k_pr_widen<true><<<init_blocks, 256, 0, s>>>(n, od, os, x32[p].get(), rank[0].get(), contrib[0].get());
        else
            k_pr_widen<false><<<init_blocks, 256, 0, s>>>(n, od, os, x32[p].get(), rank[0].get(), contrib[0].get());
        check_launch();
        for (;;) {
            // f64 sweep rank[0] -> rank[1]: the residual and the stopping test.
            dangling_base(rank[0].get(), 0);
            sweep64(0, true, 0.0);
            reduce(NO_LIMIT, -1.0);
            PrCtl h = read_ctl();
            ++iters;
            if (h.delta < tol || iters >= max_iter) {
                result = 1;
                break;
            }
            // Correction: e = r + d (A e + dangling(e) / n), r = rank[1] - rank[0], from e0 = r.
            const unsigned rb = (unsigned)((n + PR_FIX - 1) / PR_FIX);
            if (weighted)
                k_pr_residual<true><<<rb, PR_FIX, 0, s>>>(n, od, os, rank[1].get(), rank[0].get(), r32.get(),
                                                          x32[0].get(), c32[0].get(), fpart.get());
            else
                k_pr_residual<false><<<rb, PR_FIX, 0, s>>>(n, od, os, rank[1].get(), rank[0].get(), r32.get(),
                                                           x32[0].get(), c32[0].get(), fpart.get());
            check_launch();
            k_pr_reduce<<<1, 1024, 0, s>>>(fpart.get(), rb, ctl.get(), NO_LIMIT, -1.0);
            check_launch();
            double dme = read_ctl().dm;
            int pe = 0;
            // Every correction iterate has L1 norm at most |r|_1 / (1 - d), |r|_1 =
            // the f64 sweep's delta (A and the dangling spread do not grow L1).
            const double e_bound = 1.01 * h.delta / (1.0 - damping);
            for (uint32_t j = 0; j < PR_MAX_CORR && iters + 1 < max_iter; ++j) {
                sweep32(x32[pe].get(), c32[pe].get(), x32[pe ^ 1].get(), c32[pe ^ 1].get(), r32.get(),
                        damping * dme / nf, e_bound);
                reduce(NO_LIMIT, -1.0);
                const PrCtl hc = read_ctl();
                ++iters;
                ++f32_iters;
                pe ^= 1;
                dme = hc.dm;
                if (hc.delta < 0.25 * tol) break;
            }
            if (weighted)
                k_pr_apply<true><<<init_blocks, 256, 0, s>>>(n, od, os, x32[pe].get(), rank[0].get(), contrib[0].get());
            else
                k_pr_apply<false><<<init_blocks, 256, 0, s>>>(n, od, os, x32[pe].get(), rank[0].get(), contrib[0].get());
            check_launch();
        }
    }
    k_gather_f64<<<blocks_for(n, 256), 256, 0, s>>>(rank[result].get(), t->pr_new_of_old.get(), n, out.p);
    check_launch();
    out.finish(s);
    if (iters_out) *iters_out = iters;
    if (f32_iters_out) *f32_iters_out = f32_iters;
}

}  // namespace ursa

extern "C" int ursa_hip_pagerank(ursa_hip_topo* t, const double* weights, double damping, uint32_t max_iter,
                                 double tol, double* out, uint32_t* iters_out) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        ursa::pagerank_run(t, weights, damping, max_iter, tol, out, iters_out, false, nullptr);
    });
}

extern "C" int ursa_hip_pagerank_schedule(int mode) {
    if (mode < 0 || mode > 2) {
        ursa::set_last_error("schedule must be 0 (automatic), 1 (pull) or 2 (binned)");
        return URSA_HIP_EINVAL;
    }
    ursa::g_pr_schedule.store(mode);
    return URSA_HIP_OK;
}

extern "C" int ursa_hip_exact_sum(const double* x, uint64_t n, double* out) {
    return ursa::guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        URSA_REQUIRE(n == 0 || x != nullptr, URSA_HIP_EINVAL, "x is NULL");
        for (uint64_t i = 0; i < n && !ursa::is_device_ptr(x); ++i)
            URSA_REQUIRE(std::fabs(x[i]) < 4.0, URSA_HIP_ERANGE, "exact_sum takes values of magnitude below 4");
        hipStream_t s = ursa::stream();
        ursa::DevIn<double> xd(x, n, s);
        ursa::DevBuf<double> res(1);
        ursa::k_fx_sum<<<1, 256, 0, s>>>(xd.p, n, res.get());
        ursa::check_launch();
        HIP_OK(hipMemcpyAsync(out, res.get(), sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    });
}

extern "C" int ursa_hip_pagerank_mixed(ursa_hip_topo* t, const double* weights, double damping, uint32_t max_iter,
                                       double tol, double* out, uint32_t* iters_out, uint32_t* f32_iters_out) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        ursa::pagerank_run(t, weights, damping, max_iter, tol, out, iters_out, true, f32_iters_out);
    });
}


namespace ursa {
namespace {

__global__ void k_os_iota(uint32_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

// Candidate window for a free-standing ordered sum: from the sum of the finite
// magnitudes (any window gives the same result; a close one is faster).
__global__ __launch_bounds__(1024) void k_os_window(const double* __restrict__ x, uint64_t n, PrCtl* ctl) {
    __shared__ double s_red[1024 / 64];
    double a = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = fabs(x[i]);
        if (v <= 1.7e308) a += v;
    }
    const double tot = block_sum(a, s_red);
    if (threadIdx.x == 0) {
        const OsState st = os_state(tot);
        ctl->done = 0;
        ctl->dm_e0 = st.fast ? st.e - 13 : -1000;
    }
}

__global__ __launch_bounds__(OS_WALK_THREADS) void k_os_walk_out(const double* __restrict__ xd, uint64_t d,
                                                                 const uint64_t* __restrict__ cq,
                                                                 const uint32_t* __restrict__ cm,
                                                                 const double* __restrict__ cs, const PrCtl* ctl,
                                                                 double* out) {
    __shared__ OsLds lds;
    const double v = os_walk(xd, d, cq, cm, cs, ctl->dm_e0, lds);
    if (threadIdx.x == 0) *out = v;
}

}  // namespace

double ordered_sum(const double* x_user, uint64_t n) {
    hipStream_t s = stream();
    if (n == 0) return 0.0;
    URSA_REQUIRE(n <= 0xFFFFFFFFull, URSA_HIP_ERANGE, "ordered_sum takes fewer than 2^32 values");
    DevIn<double> x(x_user, n, s);
    const uint64_t nch = (n + OS_CHUNK - 1) / OS_CHUNK;
    DevBuf<uint32_t> idx(n);
    DevBuf<double> xd(n), res(1), cs(nch);
    DevBuf<uint64_t> cq(nch * OS_BINS);
    DevBuf<uint32_t> cm(nch);
    DevBuf<PrCtl> ctl(1);
    k_os_iota<<<blocks_for(n, 256, 4096), 256, 0, s>>>(idx.get(), n);
    k_os_window<<<1, 1024, 0, s>>>(x.p, n, ctl.get());
    const unsigned blocks = (unsigned)((nch + OS_BUILD_THREADS / 64 - 1) / (OS_BUILD_THREADS / 64));
    k_pr_dm_build<<<blocks, OS_BUILD_THREADS, 0, s>>>(x.p, idx.get(), n, ctl.get(), xd.get(), cq.get(), cm.get(),
                                                       cs.get());
    k_os_walk_out<<<1, OS_WALK_THREADS, 0, s>>>(xd.get(), n, cq.get(), cm.get(), cs.get(), ctl.get(), res.get());
    check_launch();
    double h = 0.0;
    HIP_OK(hipMemcpyAsync(&h, res.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return h;
}

}  // namespace ursa

extern "C" int ursa_hip_ordered_sum(const double* x, uint64_t n, double* out) {
    return ursa::guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        URSA_REQUIRE(n == 0 || x != nullptr, URSA_HIP_EINVAL, "x is NULL");
        *out = ursa::ordered_sum(x, n);
    });
}
