// pagerank.hip — pull PageRank over the device in-CSR (pagerank.rs:63-169).
//
// One iteration = three launches:
//   k_pr_tiles   merge-path SpMV: tile t owns 2048 consecutive items of the
//                merged (row-ends, in-edges) path, so every workgroup does the
//                same work whatever the degree skew. In-edge columns are read
//                coalesced, their contributions gathered into LDS, each thread
//                walks 8 path items, and a wave-level segmented scan joins row
//                pieces that straddle threads. Rows that end inside the tile
//                get the fused epilogue at once:
//                   next[u]    = base + d * acc                (pagerank.rs:84)
//                   contrib[u] = next[u] / outdeg[u]           (pagerank.rs:116-120)
//                   delta     += |next[u] - rank[u]|           (pagerank.rs:44-52)
//                   dangling  += next[u] if outdeg[u] == 0     (pagerank.rs:78)
//   k_pr_fixup   rows that began in an earlier tile: their per-tile pieces are
//                summed in tile order and finished with the same epilogue.
//   k_pr_reduce  one workgroup sums the per-tile partials in a fixed order,
//                records delta, derives base = (1-d)/n + d*dangling/n for the
//                next iteration (pagerank.rs:79) and raises `done` once
//                delta < tol (pagerank.rs:87-90) or max_iter is reached.
// Every reduction has a fixed order, so results are bit-identical run to run.
// Because `done` lives on the device and every launch returns at once when
// it is set, the host queues iterations in batches and reads the flag only
// between batches.
//
// HBM traffic per iteration (algorithmic): 12 B per in-edge (u32 column + f64
// contribution gather) and 32 B per node (row end 4, rank 8, out-degree 4,
// next 8, contribution 8) — 12m + 32n. The weighted form adds the edge weight
// (f32 when every weight is exactly representable, else f64) and reads an f64
// out-strength instead of the u32 out-degree.
#include <cmath>
#include <vector>

#include "ordered_sum.hpp"
#include "repeated_sum.hpp"
#include "prim.hpp"
#include "topology.hpp"

namespace ursa {

namespace {

constexpr int PR_BLOCK = 256;
constexpr int PR_ITEMS = 8;
constexpr int PR_TILE = PR_BLOCK * PR_ITEMS;

struct PrCtl {
    uint32_t done;
    uint32_t iters;
    double delta;
    double base[2];
    int32_t dm_e0;  // lowest candidate binade of the ordered dangling sum
};

// Split point of the merge path at diagonal `diag`: x row-ends and y in-edges
// precede it. row_end[i] = in_off[i + 1]; a row end is taken before in-edge j
// exactly when row_end <= j.
__device__ __forceinline__ void merge_path_search(uint64_t diag, const uint32_t* row_end, uint64_t n, uint64_t m,
                                                  uint64_t& x, uint64_t& y) {
    uint64_t lo = diag > m ? diag - m : 0;
    uint64_t hi = diag < n ? diag : n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if ((uint64_t)row_end[mid] + mid + 1 <= diag)
            lo = mid + 1;
        else
            hi = mid;
    }
    x = lo;
    y = diag - lo;
}

__global__ void k_pr_partition(const uint32_t* in_off, uint64_t n, uint64_t m, uint64_t tiles, uint32_t* tx,
                               uint32_t* ty) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t > tiles) return;
    uint64_t diag = t * (uint64_t)PR_TILE;
    if (diag > n + m) diag = n + m;
    uint64_t x, y;
    merge_path_search(diag, in_off + 1, n, m, x, y);
    tx[t] = (uint32_t)x;
    ty[t] = (uint32_t)y;
}

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

// Row epilogue shared by the tile and fix-up kernels. `scale` is the
// out-degree (unweighted) or out-strength (weighted) of u; scale > 0 means u
// has out-mass, otherwise u is dangling. The next contribution is stored in
// the contribution type CT of the following sweep.
template <class CT>
__device__ __forceinline__ void pr_finish_row(uint64_t u, double acc, double base, double damping, double scale,
                                              const double* rank, double* next, CT* contrib_next,
                                              double& delta) {
    const double v = __dadd_rn(base, __dmul_rn(damping, acc));  // no contraction: pagerank.rs:84
    __builtin_nontemporal_store(v, &next[u]);
    const bool has_out = scale > 0.0;
    contrib_next[u] = (CT)(has_out ? v / scale : 0.0);
    delta += fabs(v - __builtin_nontemporal_load(&rank[u]));
}

template <bool WEIGHTED>
__device__ __forceinline__ double node_scale(const uint32_t* out_deg, const double* out_str, uint64_t u) {
    if constexpr (WEIGHTED)
        return __builtin_nontemporal_load(&out_str[u]);
    else
        return (double)__builtin_nontemporal_load(&out_deg[u]);
}

// LDS slot of tile item i: one pad word per 32 words, so the 8-item stride of
// the per-thread walk falls into distinct banks.
template <class CT>
__device__ __forceinline__ uint32_t lds_slot(uint32_t i) {
    return sizeof(CT) == 8 ? i + (i >> 4) : i + (i >> 5);
}

template <class CT, class W, bool WEIGHTED>
__global__ __launch_bounds__(PR_BLOCK) void k_pr_tiles(
    const uint32_t* __restrict__ in_off, const uint32_t* __restrict__ in_tgt, const W* __restrict__ w_in,
    const uint32_t* __restrict__ tile_x, const uint32_t* __restrict__ tile_y, uint64_t n,
    const uint32_t* __restrict__ out_deg, const double* __restrict__ out_str, const double* __restrict__ rank,
    const CT* __restrict__ contrib, double* __restrict__ next, CT* __restrict__ contrib_next,
    double* __restrict__ carry_val, double* __restrict__ tail_val, double* __restrict__ tile_delta,
    PrCtl* ctl, int parity, double damping) {
    __shared__ CT s_val[PR_TILE + PR_TILE / 16];
    __shared__ uint32_t s_end[PR_TILE + 1];
    __shared__ double s_wv[PR_BLOCK / 64];
    __shared__ int s_wf[PR_BLOCK / 64];
    __shared__ double s_red[PR_BLOCK / 64];

    if (ctl->done) return;
    const uint64_t t = blockIdx.x;
    const uint64_t x0 = tile_x[t], y0 = tile_y[t];
    const uint64_t x1 = tile_x[t + 1], y1 = tile_y[t + 1];
    const uint64_t rows = x1 - x0, nnz = y1 - y0, len = rows + nnz;
    const uint64_t n_end = (x1 < n ? x1 : n - 1) - x0 + 1;  // row ends staged (incl. trailing row)
    const double base = ctl->base[parity];

    {
        // All of this thread's columns first, then all its gathers: PR_ITEMS
        // independent gathers in flight per thread instead of one at a time.
        uint32_t col[PR_ITEMS];
#pragma unroll
        for (int g = 0; g < PR_ITEMS; ++g) {
            const uint64_t i = threadIdx.x + (uint64_t)g * PR_BLOCK;
            col[g] = i < nnz ? __builtin_nontemporal_load(&in_tgt[y0 + i]) : 0u;
        }
        CT cv[PR_ITEMS];
#pragma unroll
        for (int g = 0; g < PR_ITEMS; ++g) {
            const uint64_t i = threadIdx.x + (uint64_t)g * PR_BLOCK;
            if (i < nnz) cv[g] = contrib[col[g]];
        }
#pragma unroll
        for (int g = 0; g < PR_ITEMS; ++g) {
            const uint64_t i = threadIdx.x + (uint64_t)g * PR_BLOCK;
            if (i < nnz) {
                CT c = cv[g];
                if constexpr (WEIGHTED) c = (CT)__builtin_nontemporal_load(&w_in[y0 + i]) * c;
                s_val[lds_slot<CT>((uint32_t)i)] = c;
            }
        }
    }
    if (x0 < n)
        for (uint64_t i = threadIdx.x; i < n_end; i += PR_BLOCK) s_end[i] = __builtin_nontemporal_load(&in_off[x0 + 1 + i]);
    const bool starts_mid = (x0 < n) && ((uint64_t)in_off[x0] < y0);
    __syncthreads();

    // Thread's start on the tile-local path.
    const uint64_t d0 = (uint64_t)threadIdx.x * PR_ITEMS;
    uint64_t x = x0, y = y0;
    if (d0 < len) {
        uint64_t lo = d0 > nnz ? d0 - nnz : 0, hi = d0 < rows ? d0 : rows;
        while (lo < hi) {
            uint64_t mid = (lo + hi) >> 1;
            if ((uint64_t)s_end[mid] + mid + 1 <= y0 + d0)
                lo = mid + 1;
            else
                hi = mid;
        }
        x = x0 + lo;
        y = y0 + d0 - lo;
    }

    RowAcc<CT> sum;
    double ev[PR_ITEMS];
    uint32_t er[PR_ITEMS];
    uint32_t emask = 0;
#pragma unroll
    for (int it = 0; it < PR_ITEMS; ++it) {
        ev[it] = 0.0;
        er[it] = 0;
        if (d0 + it < len) {
            if (x < n && y < (uint64_t)s_end[x - x0]) {
                sum.add(s_val[lds_slot<CT>((uint32_t)(y - y0))]);
                ++y;
            } else {
                er[it] = (uint32_t)x;
                ev[it] = sum.value();
                emask |= 1u << it;
                sum.reset();
                ++x;
            }
        }
    }

    // Segmented inclusive scan of (sum, emitted) across the workgroup: a
    // thread's incoming carry is the piece of its first row accumulated by
    // the preceding threads since their last row end.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double v = sum.value();
    int f = emask != 0;
    for (int off = 1; off < 64; off <<= 1) {
        double vu = __shfl_up(v, off, 64);
        int fu = __shfl_up(f, off, 64);
        if (lane >= off) {
            if (!f) v = vu + v;
            f |= fu;
        }
    }
    if (lane == 63) {
        s_wv[wave] = v;
        s_wf[wave] = f;
    }
    __syncthreads();
    double pv = 0.0;
    for (int w = 0; w < wave; ++w) pv = s_wf[w] ? s_wv[w] : pv + s_wv[w];
    const double v_incl = f ? v : pv + v;
    double c_in = __shfl_up(v_incl, 1, 64);
    if (lane == 0) c_in = pv;
    if (threadIdx.x == 0) c_in = 0.0;
    if (threadIdx.x == PR_BLOCK - 1) carry_val[t] = (x1 < n) ? v_incl : 0.0;

    double delta = 0.0;
    bool first = true;
#pragma unroll
    for (int it = 0; it < PR_ITEMS; ++it) {
        if (emask & (1u << it)) {
            double val = ev[it];
            if (first) {
                val = c_in + val;
                first = false;
            }
            const uint64_t u = er[it];
            if (u == x0 && starts_mid) {
                tail_val[t] = val;
            } else {
                pr_finish_row(u, val, base, damping, node_scale<WEIGHTED>(out_deg, out_str, u), rank, next,
                              contrib_next, delta);
            }
        }
    }
    const double bd = block_sum(delta, s_red);
    if (threadIdx.x == 0) tile_delta[t] = bd;
}

// Finishes rows that began in an earlier tile, then folds the per-tile delta
// partials of 256 consecutive tiles (in tile order) into one partial per
// workgroup, so the final reduce reads few values.
template <class CT, bool WEIGHTED>
__global__ __launch_bounds__(256) void k_pr_fixup(
    const uint32_t* __restrict__ in_off, const uint32_t* __restrict__ tile_x, const uint32_t* __restrict__ tile_y,
    uint64_t tiles, uint64_t n, const uint32_t* __restrict__ out_deg, const double* __restrict__ out_str,
    const double* __restrict__ rank, double* __restrict__ next, CT* __restrict__ contrib_next,
    const double* __restrict__ carry_val, const double* __restrict__ tail_val, const double* __restrict__ tile_delta,
    double* __restrict__ blk_delta, const PrCtl* ctl, int parity, double damping) {
    __shared__ double s_red[256 / 64];
    if (ctl->done) return;
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    double delta = 0.0;
    if (t < tiles) {
        const uint64_t x0 = tile_x[t], y0 = tile_y[t], x1 = tile_x[t + 1];
        if (x0 < n && x1 > x0 && (uint64_t)in_off[x0] < y0) {
            // Tiles j..t-1 all end inside row x0; tile t holds its end.
            uint64_t j = t - 1;
            while (j > 0 && tile_x[j] == x0) --j;
            double acc = 0.0;
            for (uint64_t q = j; q < t; ++q) acc += carry_val[q];
            acc += tail_val[t];
            pr_finish_row(x0, acc, ctl->base[parity], damping, node_scale<WEIGHTED>(out_deg, out_str, x0), rank,
                          next, contrib_next, delta);
        }
        delta = tile_delta[t] + delta;
    }
    const double bd = block_sum(delta, s_red);
    if (threadIdx.x == 0) blk_delta[blockIdx.x] = bd;
}

__global__ __launch_bounds__(1024) void k_pr_reduce(const double* __restrict__ blk_delta, uint64_t count, PrCtl* ctl,
                                                    uint32_t max_iter, double tol) {
    __shared__ double s_red[1024 / 64];
    if (ctl->done) return;
    double d = 0.0;
    for (uint64_t i = threadIdx.x; i < count; i += blockDim.x) d += blk_delta[i];
    const double delta = block_sum(d, s_red);
    if (threadIdx.x == 0) {
        const uint32_t iters = ctl->iters + 1;  // the iteration count lives on the device
        ctl->delta = delta;
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

// rank0 = 1/n, contrib0 = rank0 / scale.
template <class CT, bool WEIGHTED>
__global__ void k_pr_init(uint64_t n, double nf, const uint32_t* __restrict__ out_deg,
                          const double* __restrict__ out_str, double* __restrict__ rank,
                          CT* __restrict__ contrib) {
    const uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (u < n) {
        const double r = 1.0 / nf;
        rank[u] = r;
        const double sc = node_scale<WEIGHTED>(out_deg, out_str, u);
        contrib[u] = (CT)(sc > 0.0 ? r / sc : 0.0);
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
    ctl->base[0] = base0;
    ctl->base[1] = 0.0;
    ctl->dm_e0 = dm_e0;
}

// out_str[v] = sum of v's out-edge weights in out-CSR order (pagerank.rs:148-150).
__global__ void k_out_strength(const uint32_t* __restrict__ out_off, const uint32_t* __restrict__ out_eid,
                               const double* __restrict__ w, uint64_t n, double* __restrict__ out_str) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        double s = -0.0;
        for (uint32_t k = out_off[v]; k < out_off[v + 1]; ++k) s += w[out_eid[k]];
        out_str[v] = s;
    }
}

// Weights permuted into in-CSR slot order; flags any weight not exact in f32.
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

// contrib[u] = rank[u] / scale (f64), used when the mixed-precision solve
// hands over from its f32 sweeps to its f64 sweeps.
template <bool WEIGHTED>
__global__ void k_pr_contrib64(uint64_t n, const uint32_t* __restrict__ out_deg, const double* __restrict__ out_str,
                               const double* __restrict__ rank, double* __restrict__ contrib) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const double sc = node_scale<WEIGHTED>(out_deg, out_str, u);
        contrib[u] = sc > 0.0 ? rank[u] / sc : 0.0;
    }
}

// ---- renumbering by decreasing out-degree ----
__global__ void k_rl_keys(const uint32_t* __restrict__ out_deg, uint64_t n, uint32_t* keys, uint32_t* iota) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        keys[v] = ~out_deg[v];
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

__global__ void k_gather_f64(const double* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                             double* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

__global__ void k_fill(double* p, uint64_t n, double v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

// PageRank's working copy of the in-CSR, built once per topology: vertices
// renumbered by decreasing out-degree (ties by id), sources relabelled, and
// each row's sources listed in ascending new id (parallel edges keep their
// slot order). The most-gathered contributions (high out-degree sources) then
// sit in a few megabytes at the front of the contribution vector and stay in
// L2, instead of being spread over it. A row's sum runs in this order, not the
// in-CSR's: the same terms, rounded in another order.
// Also caches the merge-path tile partition of that CSR.
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
    {
        DevBuf<uint32_t> keys(n), iota(n), skeys(n), rdeg(n);
        k_rl_keys<<<blocks_for(n, 256), 256, 0, s>>>(t->out_deg.get(), n, keys.get(), iota.get());
        check_launch();
        prim::sort_pairs_u32(keys.get(), skeys.get(), iota.get(), t->pr_old_of_new.get(), n, 32, s);
        k_rl_invert<<<blocks_for(n, 256), 256, 0, s>>>(t->pr_old_of_new.get(), n, in.off.get(), t->out_deg.get(),
                                                       t->pr_new_of_old.get(), rdeg.get(), t->pr_outdeg.get());
        check_launch();
        prim::exclusive_sum_u32(rdeg.get(), t->pr_off.get(), n, s);
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
    const uint64_t tiles = (n + m + PR_TILE - 1) / PR_TILE;
    t->pr_tile_x.alloc(tiles + 1);
    t->pr_tile_y.alloc(tiles + 1);
    k_pr_partition<<<blocks_for(tiles + 1, 256), 256, 0, s>>>(t->pr_off.get(), n, m, tiles, t->pr_tile_x.get(),
                                                              t->pr_tile_y.get());
    check_launch();
    HIP_OK(hipStreamSynchronize(s));
    t->pr_tiles = tiles;
    t->pr_ready = true;
}

// Solve PageRank. With `mixed`, sweeps first run with f32 contributions and
// compensated f32 row sums (8 B per edge instead of 12) until the L1 delta
// stops shrinking quickly or falls below 2e-6, then hand over to f64 sweeps,
// which refine the same iterate until delta < tol. Every iteration of either
// kind counts toward max_iter, and the answer is always the last f64 iterate
// (or the uniform start when max_iter == 0).
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
    pr_prepare(t, t->incoming(s), s);
    const uint64_t tiles = t->pr_tiles;
    // All solve-time arrays below are in the renumbered vertex order.
    struct {
        const uint32_t* off;
        const uint32_t* tgt;
        const uint32_t* eid;
    } in{t->pr_off.get(), t->pr_tgt.get(), t->pr_eid.get()};

    const bool weighted = weights_user != nullptr;
    DevIn<double> w(weights_user, m, s);
    DevBuf<double> out_str, w64;
    DevBuf<float> w32;
    bool use_f32w = false;
    if (weighted) {
        DevBuf<double> os_old(n);
        out_str.alloc(n);
        k_out_strength<<<blocks_for(n, 256), 256, 0, s>>>(t->out.off.get(), t->out.eid.get(), w.p, n, os_old.get());
        k_gather_f64<<<blocks_for(n, 256), 256, 0, s>>>(os_old.get(), t->pr_old_of_new.get(), n, out_str.get());
        check_launch();
        HIP_OK(hipStreamSynchronize(s));
        w64.alloc(m);
        w32.alloc(m);
        DevBuf<uint32_t> inexact(1);
        HIP_OK(hipMemsetAsync(inexact.get(), 0, sizeof(uint32_t), s));
        if (m) {
            k_permute_weights<<<blocks_for(m, 256), 256, 0, s>>>(in.eid, w.p, m, w64.get(), w32.get(),
                                                                 inexact.get());
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

    DevBuf<double> rank[2] = {DevBuf<double>(n), DevBuf<double>(n)};
    DevBuf<double> contrib[2] = {DevBuf<double>(n), DevBuf<double>(n)};
    DevBuf<float> contrib32[2];
    if (mixed) {
        contrib32[0].alloc(n);
        contrib32[1].alloc(n);
    }
    DevBuf<double> carry(tiles), tail(tiles), td(tiles), fd(tiles);
    DevBuf<PrCtl> ctl(1);
    const unsigned init_blocks = blocks_for(n, 256, 1u << 30);
    const uint32_t* od = t->pr_outdeg.get();
    const double* os = weighted ? out_str.get() : nullptr;
    if (mixed) {
        if (weighted)
            k_pr_init<float, true><<<init_blocks, 256, 0, s>>>(n, nf, od, os, rank[0].get(), contrib32[0].get());
        else
            k_pr_init<float, false><<<init_blocks, 256, 0, s>>>(n, nf, od, os, rank[0].get(), contrib32[0].get());
    } else {
        if (weighted)
            k_pr_init<double, true><<<init_blocks, 256, 0, s>>>(n, nf, od, os, rank[0].get(), contrib[0].get());
        else
            k_pr_init<double, false><<<init_blocks, 256, 0, s>>>(n, nf, od, os, rank[0].get(), contrib[0].get());
    }
    check_launch();

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
    const double base_a = (1.0 - damping) / nf, base_b = damping * dm0;
    const double base0 = base_a + base_b / nf;  // pagerank.rs:79
    const int32_t e0 = dm0 > 0.0 ? (int32_t)std::ilogb(dm0) - 13 : 0;
    k_pr_ctl_init<<<1, 1, 0, s>>>(ctl.get(), e0, base0);
    check_launch();

    const double node_bytes = weighted ? 36.0 : 32.0;
    const double f64_edge = weighted ? (use_f32w ? 16.0 : 20.0) : 12.0;
    const double f32_edge = weighted ? 12.0 : 8.0;
    const unsigned fix_blocks = blocks_for(tiles, 256, 1u << 30);
    const uint32_t* tx = t->pr_tile_x.get();
    const uint32_t* ty = t->pr_tile_y.get();

    // One iteration `it` (parity it & 1) with contribution type CT.
    auto iteration = [&](uint32_t it, bool f32) {
        const int p = (int)(it & 1);
        const double* rk = rank[p].get();
        double* nx = rank[p ^ 1].get();
        if (f32) {
            const float* cb = contrib32[p].get();
            float* cn = contrib32[p ^ 1].get();
            timed("pagerank_spmv_f32", f32_edge * (double)m + (node_bytes - 4.0) * (double)n, s, [&] {
                if (weighted)
                    k_pr_tiles<float, float, true><<<(unsigned)tiles, PR_BLOCK, 0, s>>>(
                        in.off, in.tgt, w32.get(), tx, ty, n, od, os, rk, cb, nx, cn, carry.get(),
                        tail.get(), td.get(), ctl.get(), p, damping);
                else
                    k_pr_tiles<float, float, false><<<(unsigned)tiles, PR_BLOCK, 0, s>>>(
                        in.off, in.tgt, nullptr, tx, ty, n, od, os, rk, cb, nx, cn, carry.get(),
                        tail.get(), td.get(), ctl.get(), p, damping);
            });
            if (weighted)
                k_pr_fixup<float, true><<<fix_blocks, 256, 0, s>>>(in.off, tx, ty, tiles, n, od, os, rk, nx, cn,
                                                                   carry.get(), tail.get(), td.get(), fd.get(),
                                                                   ctl.get(), p, damping);
            else
                k_pr_fixup<float, false><<<fix_blocks, 256, 0, s>>>(in.off, tx, ty, tiles, n, od, os, rk, nx, cn,
                                                                    carry.get(), tail.get(), td.get(), fd.get(),
                                                                    ctl.get(), p, damping);
        } else {
            const double* cb = contrib[p].get();
            double* cn = contrib[p ^ 1].get();
            timed("pagerank_spmv", f64_edge * (double)m + node_bytes * (double)n, s, [&] {
                if (!weighted)
                    k_pr_tiles<double, double, false><<<(unsigned)tiles, PR_BLOCK, 0, s>>>(
                        in.off, in.tgt, nullptr, tx, ty, n, od, os, rk, cb, nx, cn, carry.get(),
                        tail.get(), td.get(), ctl.get(), p, damping);
                else if (use_f32w)
                    k_pr_tiles<double, float, true><<<(unsigned)tiles, PR_BLOCK, 0, s>>>(
                        in.off, in.tgt, w32.get(), tx, ty, n, od, os, rk, cb, nx, cn, carry.get(),
                        tail.get(), td.get(), ctl.get(), p, damping);
                else
                    k_pr_tiles<double, double, true><<<(unsigned)tiles, PR_BLOCK, 0, s>>>(
                        in.off, in.tgt, w64.get(), tx, ty, n, od, os, rk, cb, nx, cn, carry.get(),
                        tail.get(), td.get(), ctl.get(), p, damping);
            });
            if (weighted)
                k_pr_fixup<double, true><<<fix_blocks, 256, 0, s>>>(in.off, tx, ty, tiles, n, od, os, rk, nx,
                                                                    cn, carry.get(), tail.get(), td.get(), fd.get(),
                                                                    ctl.get(), p, damping);
            else
                k_pr_fixup<double, false><<<fix_blocks, 256, 0, s>>>(in.off, tx, ty, tiles, n, od, os, rk, nx,
                                                                     cn, carry.get(), tail.get(), td.get(), fd.get(),
                                                                     ctl.get(), p, damping);
        }
        check_launch();
    };
    auto reduce = [&](uint32_t it, double tol_it) {
        k_pr_reduce<<<1, 1024, 0, s>>>(fd.get(), fix_blocks, ctl.get(), max_iter, tol_it);
        check_launch();
        const int q = (int)((it & 1) ^ 1);
        dangling_base(rank[q].get(), q);
    };
    auto read_ctl = [&]() {
        PrCtl h{};
        HIP_OK(hipMemcpyAsync(&h, ctl.get(), sizeof(PrCtl), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return h;
    };

    uint32_t it = 0;
    uint32_t f32_iters = 0;
    if (mixed) {
        // f32 sweeps, one at a time (the host watches the delta). The reduce
        // kernel's own convergence test is disabled (tol -1) so it only raises
        // `done` at max_iter.
        double prev = 0.0;
        while (it < max_iter) {
            iteration(it, true);
            reduce(it, -1.0);
            const PrCtl h = read_ctl();
            ++it;
            ++f32_iters;
            if (h.done) break;
            const bool stalled = f32_iters >= 3 && h.delta > 0.7 * prev;
            prev = h.delta;
            if (h.delta < 2e-6 || stalled) break;
        }
        // Hand over: f64 contributions of the current iterate.
        const int p = (int)(it & 1);
        if (weighted)
            k_pr_contrib64<true><<<blocks_for(n, 256), 256, 0, s>>>(n, od, os, rank[p].get(), contrib[p].get());
        else
            k_pr_contrib64<false><<<blocks_for(n, 256), 256, 0, s>>>(n, od, os, rank[p].get(), contrib[p].get());
        check_launch();
        const PrCtl h = read_ctl();
        if (h.done) {
            // max_iter was spent in f32 sweeps; one more f64 sweep is not allowed,
            // so the f32 result stands.
            it = max_iter;
        }
    }
    PrCtl h = read_ctl();
    // Iterations are queued in batches; with per-kernel timing on, one at a
    // time so no launch after convergence is booked. A full batch of 8 is
    // recorded once per start parity as a HIP graph and replayed.
    const uint32_t batch = prof_enabled() ? 1 : 8;
    CapturedLaunches graphs[2];
    while (it < max_iter && !h.done) {
        const uint32_t it_end = std::min<uint32_t>(max_iter, it + batch);
        const uint32_t it0 = it;
        auto launches = [&] {
            for (uint32_t i = it0; i < it_end; ++i) {
                iteration(i, false);
                reduce(i, tol);
            }
        };
        if (it_end - it0 == 8)
            graphs[it0 & 1].run(s, launches);
        else
            launches();
        it = it_end;
        h = read_ctl();
    }
    h = read_ctl();
    const uint32_t iters = h.iters;
    k_gather_f64<<<blocks_for(n, 256), 256, 0, s>>>(rank[iters & 1].get(), t->pr_new_of_old.get(), n, out.p);
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
