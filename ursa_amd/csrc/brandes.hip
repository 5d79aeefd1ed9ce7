// brandes.hip — Brandes betweenness (betweenness, ursa-core/src/algo/
// betweenness.rs:45-230, 349-362) over out-edges, sigma and delta in f64.
//
// Sources are processed in batches of up to 64, which is exactly the
// reference's accumulation chunk (SRC_CHUNK, betweenness.rs:46): a batch's
// per-source dependencies are added into a chunk partial in source order, and
// the chunk partials are added into bc in chunk order, as accumulate_sources
// does (betweenness.rs:63-68). Within a batch all sources advance level by level
// together, so one launch per level serves up to 64 traversals.
//
// Forward (one launch per level): the level's (source, vertex) entries are
// expanded wave-cooperatively (degree prefix-sum over 64 lanes); an unvisited
// target is claimed with a CAS on its distance, and every out-slot from level
// L-1 into level L adds sigma[u] to sigma[v] (betweenness.rs:204-214: per CSR
// slot, so parallel edges count as distinct paths). Path counts are integers
// held in f64; their sums are exact below 2^53 and so independent of the order
// of the atomic adds. An add whose result reaches 2^53 (the only way one can
// round) raises a flag, and that level's path counts are then recomputed by an
// ordered pull: sigma[w] = sum of sigma[u] over w's in-slots u at level L-1, in
// in-CSR slot order. Results are therefore bit-identical run to run for any
// path count.
// Backward (one launch per level, deepest first): each vertex v at level L-1
// pulls delta[v] = sum over out-slots w at level L of
// sigma[v] * ((1 + delta[w]) / sigma[w]) — the same per-term arithmetic as
// reverse_accumulate (betweenness.rs:89-97), summed in out-CSR slot order.
// Vertices with more than 128 out-slots are summed by a whole wave with a
// fixed-shape reduction, so results are bit-identical run to run.
//
// Algorithmic bytes per source: 32 B per reached out-edge (forward column +
// distance, backward column + distance + sigma + delta) + 44 B per node.
#include <algorithm>
#include <vector>

#include "frontier.hpp"
#include "prim.hpp"
#include "rng.hpp"

namespace ursa {

namespace {

constexpr int BB = 256;
constexpr int BW = BB / 64;
constexpr uint32_t HEAVY_OUT = 128;

struct BrCtr {
    uint32_t len;
    uint32_t heavy;
    unsigned long long edges;
    uint32_t big;  // a path-count sum reached 2^53 this level
    uint32_t pad;
};

constexpr double SIGMA_EXACT = 9007199254740992.0;  // 2^53: integer sums below it are exact

// Adds sv to a path count; true when the sum reached 2^53.
__device__ __forceinline__ bool sigma_add(double* p, double sv) {
    return unsafeAtomicAdd(p, sv) + sv >= SIGMA_EXACT;
}

__global__ void k_br_init(int32_t* dist, double* sigma, double* delta, uint64_t total) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        dist[i] = -1;
        sigma[i] = 0.0;
        delta[i] = 0.0;
    }
}

__global__ void k_br_seed(const uint32_t* src, uint32_t nb, uint64_t n, int32_t* dist, double* sigma,
                          unsigned long long* q, BrCtr* ctr) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t len = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t s = src[b];
        if (s >= n) continue;
        dist[b * n + s] = 0;
        sigma[b * n + s] = 1.0;
        q[len++] = ((unsigned long long)b << 32) | s;
    }
    ctr->len = len;
}

__global__ __launch_bounds__(BB) void k_br_forward(const unsigned long long* __restrict__ q, uint32_t qlen,
                                                   const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt,
                                                   uint64_t n, int32_t* __restrict__ dist, double* __restrict__ sigma,
                                                   int32_t L, unsigned long long* __restrict__ nq, BrCtr* ctr) {
    const int lane = lane_id();
    const uint64_t gw = blockIdx.x * (uint64_t)BW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * BW;
    unsigned long long edges = 0;
    bool big = false;
    for (uint64_t base = gw * 64; base < qlen; base += nwaves * 64) {
        const uint64_t idx = base + lane;
        uint32_t d = 0, st = 0, b = 0;
        double su = 0.0;
        if (idx < qlen) {
            const unsigned long long e = q[idx];
            b = (uint32_t)(e >> 32);
            const uint32_t u = (uint32_t)e;
            st = off[u];
            d = off[u + 1] - st;
            su = sigma[b * n + u];
        }
        edges += d;
        const uint32_t incl = wave_incl_scan(d);
        const uint32_t excl = incl - d;
        const uint32_t total = __shfl(incl, 63, 64);
        for (uint32_t t0 = 0; t0 < total; t0 += 64) {
            const uint32_t t = t0 + lane;
            const int j = wave_owner(excl, t);
            const uint32_t e = t - __shfl(excl, j, 64);
            const uint32_t ob = __shfl(b, j, 64);
            const uint32_t ost = __shfl(st, j, 64);
            const double osu = __shfl(su, j, 64);
            bool claimed = false;
            uint32_t v = 0;
            if (t < total) {
                v = tgt[ost + e];
                int32_t* dv = &dist[ob * n + v];
                int32_t cur = *dv;
                if (cur < 0) {
                    const int32_t old = atomicCAS(dv, -1, L);
                    claimed = old == -1;
                    cur = claimed ? L : old;
                }
                if (cur == L) big |= sigma_add(&sigma[ob * n + v], osu);
            }
            const uint32_t slot = wave_append_slot(claimed, &ctr->len);
            if (claimed) nq[slot] = ((unsigned long long)ob << 32) | v;
        }
    }
    edges = wave_sum_u64(edges);
    if (lane == 0 && edges) atomicAdd(&ctr->edges, edges);
    if (__ballot(big) && lane == 0) atomicOr(&ctr->big, 1u);
}

// Ordered recount of a level's path counts (source-major layout): for every
// (source b, vertex w) entry of level L, sigma = sum over w's in-slots u with
// dist[b][u] == L - 1 of sigma[b][u], in in-CSR slot order.
__global__ void k_br_sigma_pull(const unsigned long long* __restrict__ q, uint32_t qlen,
                                const uint32_t* __restrict__ in_off, const uint32_t* __restrict__ in_tgt, uint64_t n,
                                const int32_t* __restrict__ dist, double* __restrict__ sigma, int32_t L) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < qlen; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long e = q[i];
        const uint64_t bo = (e >> 32) * n;
        const uint32_t w = (uint32_t)e;
        double acc = 0.0;
        for (uint32_t k = in_off[w]; k < in_off[w + 1]; ++k) {
            const uint32_t u = in_tgt[k];
            if (dist[bo + u] == L - 1) acc += sigma[bo + u];
        }
        sigma[bo + w] = acc;
    }
}

// delta for the entries of one level; heavy entries are deferred to a wave each.
__global__ void k_br_backward(const unsigned long long* __restrict__ q, uint32_t qlen,
                              const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt, uint64_t n,
                              const int32_t* __restrict__ dist, const double* __restrict__ sigma,
                              double* __restrict__ delta, int32_t L, uint32_t* __restrict__ heavy, BrCtr* ctr) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < qlen; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long e = q[i];
        const uint64_t b = e >> 32;
        const uint32_t v = (uint32_t)e;
        const uint32_t a = off[v], z = off[v + 1];
        if (z - a > HEAVY_OUT) {
            heavy[atomicAdd(&ctr->heavy, 1u)] = (uint32_t)i;
            continue;
        }
        const uint64_t bo = b * n;
        const double sv = sigma[bo + v];
        double acc = 0.0;
        for (uint32_t k = a; k < z; ++k) {
            const uint32_t w = tgt[k];
            if (dist[bo + w] == L) acc += sv * ((1.0 + delta[bo + w]) / sigma[bo + w]);
        }
        delta[bo + v] = acc;
    }
}

__global__ __launch_bounds__(BB) void k_br_backward_heavy(const unsigned long long* __restrict__ q,
                                                          const uint32_t* __restrict__ heavy, const BrCtr* ctr,
                                                          const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ tgt, uint64_t n,
                                                          const int32_t* __restrict__ dist,
                                                          const double* __restrict__ sigma,
                                                          double* __restrict__ delta, int32_t L) {
    const uint64_t gw = blockIdx.x * (uint64_t)BW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * BW;
    const uint32_t cnt = ctr->heavy;
    for (uint64_t h = gw; h < cnt; h += nwaves) {
        const unsigned long long e = q[heavy[h]];
        const uint64_t bo = (e >> 32) * n;
        const uint32_t v = (uint32_t)e;
        const double sv = sigma[bo + v];
        double acc = 0.0;
        for (uint32_t k = off[v] + lane_id(); k < off[v + 1]; k += 64) {
            const uint32_t w = tgt[k];
            if (dist[bo + w] == L) acc += sv * ((1.0 + delta[bo + w]) / sigma[bo + w]);
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if (lane_id() == 0) delta[bo + v] = acc;
    }
}

// chunk partial in source order (betweenness.rs:57-60, :94-96), then
// bc += partial in chunk order (betweenness.rs:63-68).
__global__ void k_br_accumulate(const double* __restrict__ delta, const int32_t* __restrict__ dist, uint32_t nb,
                                uint64_t n, double* __restrict__ bc) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (uint32_t b = 0; b < nb; ++b)
            if (dist[b * n + v] > 0) acc += delta[b * n + v];
        bc[v] += acc;
    }
}

__global__ void k_add(const double* __restrict__ a, uint64_t n, double* __restrict__ bc) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        bc[v] += a[v];
}

__global__ void k_scale(double* p, uint64_t n, double s) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = p[i] * s;
}


// ---- vertex-major batches: the 64 sources of a batch are the 64 lanes of a
// wavefront. State for (vertex v, source b) lives at v * 64 + b, so one edge
// scan of v serves every source whose frontier holds v, and the per-source
// reads and atomics for a neighbour w are one contiguous 256 / 512 B access.
constexpr uint32_t VM_HEAVY = 256;  // backward: vertices with more out-edges are split into chunks of this size
// Forward: vertices with more out-edges are split into chunks of this size. A
// wave walks its vertex's (or chunk's) edges four at a time for all 64
// sources, so a long list is a long serial chain that ends the level last;
// the forward sums are exact integers, so the split changes no result (the
// backward chunking fixes the f64 order of its partial sums and stays).
constexpr uint32_t VM_FHEAVY = 64;

struct VmCtr {
    uint32_t len;
    uint32_t big;  // a path-count sum reached 2^53 this level
    unsigned long long edges;
};

// Device-side state of a batch's forward level loop: level L expands the
// queue entries order[a, z) (level L - 1) and appends level L at order[z, ...);
// lvl[k] is the start of level k in order[].
struct VmCtl {
    int32_t L;
    uint32_t done;
    uint32_t fresh;  // no level has run yet
    uint32_t big;    // some level's path-count sums reached 2^53
    unsigned long long a, z;
    unsigned long long edges;
    int32_t levels;  // when done: the reference loop's level count
    int32_t pad;
};

// Closes the level just expanded (its length from the counters) and opens the
// next, or marks the loop done once a level came out empty.
__global__ void k_vm_level_step(VmCtl* ctl, VmCtr* ctr, unsigned long long* lvl) {
    VmCtl c = *ctl;
    if (c.done) return;
    if (!c.fresh) {
        const unsigned long long len = ctr->len;
        c.edges += ctr->edges;
        if (ctr->big) c.big = 1;
        lvl[c.L + 1] = c.z + len;
        c.a = c.z;
        c.z = c.z + len;
        c.L += 1;
    }
    c.fresh = 0;
    if (c.z == c.a) {
        c.done = 1;
        c.levels = c.L;
    }
    *ctr = VmCtr{};
    *ctl = c;
}

__global__ void k_vm_init(int32_t* dist, double* sigma, double* delta, uint32_t* mark, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n * 64;
         i += (uint64_t)gridDim.x * blockDim.x) {
        dist[i] = -1;
        sigma[i] = 0.0;
        delta[i] = 0.0;
        if (i < n) mark[i] = 0;
    }
}

__global__ void k_vm_seed(const uint32_t* src, uint32_t nb, uint64_t n, int32_t* dist, double* sigma, uint32_t* mark,
                          uint32_t* order, VmCtr* ctr, VmCtl* ctl, unsigned long long* lvl) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t len = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t s = src[b];
        if (s >= n) continue;
        dist[(uint64_t)s * 64 + b] = 0;
        sigma[(uint64_t)s * 64 + b] = 1.0;
        if (mark[s] == 0) {
            mark[s] = 1;  // queued at level 0 (mark = level + 1)
            order[len++] = s;
        }
    }
    *ctr = VmCtr{};
    VmCtl c{};
    c.L = 1;
    c.fresh = 1;
    c.a = 0;
    c.z = len;
    *ctl = c;
    lvl[0] = 0;
    lvl[1] = len;
}

// A wave's newly queued vertices, appended 64 at a time: lane i holds the
// i-th pending vertex, cnt is wave-uniform.
struct VmAppend {
    uint32_t buf = 0;
    uint32_t cnt = 0;
};

__device__ __forceinline__ void vm_append_flush(VmAppend& ap, uint32_t* __restrict__ nq, VmCtr* ctr) {
    if (ap.cnt == 0) return;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(&ctr->len, ap.cnt);
    base = __shfl(base, 0, 64);
    if ((uint32_t)lane_id() < ap.cnt) nq[base + lane_id()] = ap.buf;
    ap.cnt = 0;
}

// Expand v's out-edges [e0, e1) for the lanes whose level at v is L-1.
__device__ __forceinline__ void vm_expand(uint32_t v, uint32_t e0, uint32_t e1, const uint32_t* __restrict__ tgt,
                                          int32_t* __restrict__ dist, double* __restrict__ sigma, int32_t L,
                                          uint32_t* __restrict__ mark, uint32_t* __restrict__ nq, VmCtr* ctr,
                                          bool count_edges, uint32_t deg, unsigned long long& edges, VmAppend& ap,
                                          bool& big) {
    const int b = lane_id();
    const bool fl = dist[(uint64_t)v * 64 + b] == L - 1;
    const uint64_t fmask = __ballot(fl);
    if (!fmask) return;
    if (count_edges && b == 0) edges += (unsigned long long)deg * __popcll(fmask);
    const double sv = fl ? sigma[(uint64_t)v * 64 + b] : 0.0;
    for (uint32_t e0q = e0; e0q < e1; e0q += 4) {
        // Four slots at a time, in phases so that their atomics overlap:
        // targets and current levels, then the distance claims, then the
        // path-count adds, then the queue marks. (Path counts are integers, exact
        // below 2^53 whatever the order of the adds, and the queue order never
        // reaches a result.)
        uint32_t wq[4];
        int32_t dq[4];
        bool claimed[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) wq[j] = e0q + j < e1 ? tgt[e0q + j] : 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) dq[j] = (fl && e0q + j < e1) ? dist[(uint64_t)wq[j] * 64 + b] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            claimed[j] = false;
            if (fl && e0q + j < e1 && dq[j] < 0) {  // a repeated target sees the claim through the CAS
                const int32_t old = atomicCAS(&dist[(uint64_t)wq[j] * 64 + b], -1, L);
                claimed[j] = old == -1;
                dq[j] = claimed[j] ? L : old;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (fl && e0q + j < e1 && dq[j] == L) big |= sigma_add(&sigma[(uint64_t)wq[j] * 64 + b], sv);
        uint64_t cm[4];
        bool fresh[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cm[j] = __ballot(claimed[j]);
            fresh[j] = false;
            if (cm[j] && b == __ffsll((unsigned long long)cm[j]) - 1)
                fresh[j] = atomicMax(&mark[wq[j]], (uint32_t)(L + 1)) < (uint32_t)(L + 1);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (cm[j] && __ballot(fresh[j])) {  // wq[j] is the same on every lane
                if ((uint32_t)b == ap.cnt) ap.buf = wq[j];
                if (++ap.cnt == 64) vm_append_flush(ap, nq, ctr);
            }
        }
    }
}

__global__ __launch_bounds__(BB) void k_vm_forward(uint32_t* __restrict__ order, const VmCtl* ctl,
                                                   const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt,
                                                   int32_t* __restrict__ dist, double* __restrict__ sigma,
                                                   uint32_t* __restrict__ mark, VmCtr* ctr) {
    if (ctl->done) return;
    const uint32_t* q = order + ctl->a;
    uint32_t* nq = order + ctl->z;
    const uint64_t qlen = ctl->z - ctl->a;
    const int32_t L = ctl->L;
    const uint64_t gw = blockIdx.x * (uint64_t)BW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * BW;
    unsigned long long edges = 0;
    bool big = false;
    VmAppend ap;
    for (uint64_t i = gw; i < qlen; i += nwaves) {
        const uint32_t v = q[i];
        const uint32_t a = off[v], z = off[v + 1];
        if (z - a > VM_FHEAVY) continue;  // chunked by k_vm_forward_heavy
        vm_expand(v, a, z, tgt, dist, sigma, L, mark, nq, ctr, true, z - a, edges, ap, big);
    }
    vm_append_flush(ap, nq, ctr);
    if (__ballot(big) && lane_id() == 0) atomicOr(&ctr->big, 1u);
    block_add<BB>(edges, &ctr->edges);
}

// Heavy vertices, one wave per chunk of VM_FHEAVY edges (forward) or
// VM_HEAVY edges (backward); chunk c of heavy vertex h lies in
// [cbase[h], cbase[h+1]).
__device__ __forceinline__ uint32_t vm_chunk_owner(const uint32_t* cbase, uint32_t nh, uint32_t c) {
    uint32_t lo = 0, hi = nh;  // largest h with cbase[h] <= c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cbase[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(BB) void k_vm_forward_heavy(const uint32_t* __restrict__ heavy,
                                                         const uint32_t* __restrict__ cbase, uint32_t nh,
                                                         uint32_t chunks, const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ tgt, int32_t* __restrict__ dist,
                                                         double* __restrict__ sigma, uint32_t* __restrict__ mark,
                                                         uint32_t* __restrict__ order, const VmCtl* ctl, VmCtr* ctr) {
    if (ctl->done) return;
    uint32_t* nq = order + ctl->z;
    const int32_t L = ctl->L;
    const uint64_t gw = blockIdx.x * (uint64_t)BW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * BW;
    unsigned long long edges = 0;
    bool big = false;
    VmAppend ap;
    for (uint64_t c = gw; c < chunks; c += nwaves) {
        const uint32_t h = vm_chunk_owner(cbase, nh, (uint32_t)c);
        const uint32_t v = heavy[h];
        const uint32_t a = off[v], z = off[v + 1];
        const uint32_t e0 = a + ((uint32_t)c - cbase[h]) * VM_FHEAVY;
        const uint32_t e1 = min(z, e0 + VM_FHEAVY);
        vm_expand(v, e0, e1, tgt, dist, sigma, L, mark, nq, ctr, c == cbase[h], z - a, edges, ap, big);
    }
    vm_append_flush(ap, nq, ctr);
    if (__ballot(big) && lane_id() == 0) atomicOr(&ctr->big, 1u);
    block_add<BB>(edges, &ctr->edges);
}

// Ordered recount of a level's path counts (vertex-major layout): one wave per
// vertex w queued at level L; lane b with dist == L sums sigma over w's
// in-slots u at level L - 1 in in-CSR slot order.
__global__ __launch_bounds__(BB) void k_vm_sigma_pull(const uint32_t* __restrict__ order, const VmCtl* ctl,
                                                      const VmCtr* ctr, const uint32_t* __restrict__ in_off,
                                                      const uint32_t* __restrict__ in_tgt,
                                                      const int32_t* __restrict__ dist, double* __restrict__ sigma) {
    if (ctl->done || !ctr->big) return;
    const uint32_t* q = order + ctl->z;
    const uint64_t qlen = ctr->len;
    const int32_t L = ctl->L;
    const uint64_t gw = blockIdx.x * (uint64_t)BW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * BW;
    const int b = lane_id();
    for (uint64_t i = gw; i < qlen; i += nwaves) {
        const uint32_t w = q[i];
        const bool on = dist[(uint64_t)w * 64 + b] == L;
        if (!__ballot(on)) continue;
        double acc = 0.0;
        for (uint32_t k = in_off[w]; k < in_off[w + 1]; ++k) {
            const uint64_t ui = (uint64_t)in_tgt[k] * 64 + b;
            if (on && dist[ui] == L - 1) acc += sigma[ui];
        }
        if (on) sigma[(uint64_t)w * 64 + b] = acc;
    }
}

// delta for lanes at level L: sum over out-slots w at level L+1 of
// sigma[v] * ((1 + delta[w]) / sigma[w]) in slot order.
__device__ __forceinline__ double vm_pull(uint32_t v, uint32_t e0, uint32_t e1, const uint32_t* __restrict__ tgt,
                                          const int32_t* __restrict__ dist, const double* __restrict__ sigma,
                                          const double* __restrict__ delta, int32_t L, bool fl) {
    const int b = lane_id();
    double acc = 0.0;
    if (!fl) return acc;
    const double sv = sigma[(uint64_t)v * 64 + b];
    uint32_t e = e0;
    for (; e + 4 <= e1; e += 4) {  // four slots' loads in flight; the sum keeps slot order
        uint64_t wi[4];
        bool on[4];
        double dl[4], sg[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) wi[j] = (uint64_t)tgt[e + j] * 64 + b;
#pragma unroll
        for (int j = 0; j < 4; ++j) on[j] = dist[wi[j]] == L + 1;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            dl[j] = on[j] ? delta[wi[j]] : 0.0;
            sg[j] = on[j] ? sigma[wi[j]] : 1.0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (on[j]) acc += sv * ((1.0 + dl[j]) / sg[j]);
    }
    for (; e < e1; ++e) {
        const uint64_t wi = (uint64_t)tgt[e] * 64 + b;
        if (dist[wi] == L + 1) acc += sv * ((1.0 + delta[wi]) / sigma[wi]);
    }
    return acc;
}

__global__ __launch_bounds__(BB) void k_vm_backward(const uint32_t* __restrict__ q, uint32_t qlen,
                                                    const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt,
                                                    const int32_t* __restrict__ dist,
                                                    const double* __restrict__ sigma, double* __restrict__ delta,
                                                    int32_t L) {
    const uint64_t gw = blockIdx.x * (uint64_t)BW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * BW;
    for (uint64_t i = gw; i < qlen; i += nwaves) {
        const uint32_t v = q[i];
        const uint32_t a = off[v], z = off[v + 1];
        if (z - a > VM_HEAVY) continue;
        const bool fl = dist[(uint64_t)v * 64 + lane_id()] == L;
        if (!__ballot(fl)) continue;
        const double acc = vm_pull(v, a, z, tgt, dist, sigma, delta, L, fl);
        if (fl) delta[(uint64_t)v * 64 + lane_id()] = acc;
    }
}

__global__ __launch_bounds__(BB) void k_vm_backward_heavy(const uint32_t* __restrict__ heavy,
                                                          const uint32_t* __restrict__ cbase, uint32_t nh,
                                                          uint32_t chunks, const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ tgt,
                                                          const int32_t* __restrict__ dist,
                                                          const double* __restrict__ sigma,
                                                          const double* __restrict__ delta, int32_t L,
                                                          double* __restrict__ part) {
    const uint64_t gw = blockIdx.x * (uint64_t)BW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * BW;
    for (uint64_t c = gw; c < chunks; c += nwaves) {
        const uint32_t h = vm_chunk_owner(cbase, nh, (uint32_t)c);
        const uint32_t v = heavy[h];
        const uint32_t a = off[v], z = off[v + 1];
        const uint32_t e0 = a + ((uint32_t)c - cbase[h]) * VM_HEAVY;
        const uint32_t e1 = min(z, e0 + VM_HEAVY);
        const bool fl = dist[(uint64_t)v * 64 + lane_id()] == L;
        part[c * 64 + lane_id()] = __ballot(fl) ? vm_pull(v, e0, e1, tgt, dist, sigma, delta, L, fl) : 0.0;
    }
}

// delta of heavy vertices at level L: chunk partials summed in chunk order.
__global__ void k_vm_backward_combine(const uint32_t* __restrict__ heavy, const uint32_t* __restrict__ cbase,
                                      uint32_t nh, const int32_t* __restrict__ dist, const double* __restrict__ part,
                                      double* __restrict__ delta, int32_t L) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= (uint64_t)nh * 64) return;
    const uint32_t h = (uint32_t)(i >> 6), b = (uint32_t)(i & 63);
    const uint64_t vi = (uint64_t)heavy[h] * 64 + b;
    if (dist[vi] != L) return;
    double acc = 0.0;
    for (uint32_t c = cbase[h]; c < cbase[h + 1]; ++c) acc += part[(uint64_t)c * 64 + b];
    delta[vi] = acc;
}

// chunk partial over the batch's sources in source order, then bc += partial.
__global__ void k_vm_accumulate(const int32_t* __restrict__ dist, const double* __restrict__ delta, uint32_t nb,
                                uint64_t n, double* __restrict__ bc) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (uint32_t b = 0; b < nb; ++b)
            if (dist[v * 64 + b] > 0) acc += delta[v * 64 + b];
        bc[v] += acc;
    }
}

__global__ void k_vm_heavy_flags(const uint32_t* __restrict__ off, uint64_t n, uint32_t thr, uint8_t* flag,
                                 uint32_t* iota, uint32_t* chunks) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = off[v + 1] - off[v];
        flag[v] = d > thr;
        iota[v] = (uint32_t)v;
        chunks[v] = (d + thr - 1) / thr;
    }
}

__global__ void k_vm_gather_chunks(const uint32_t* __restrict__ heavy, uint32_t nh, const uint32_t* __restrict__ chunks,
                                   uint32_t* __restrict__ hchunks) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < nh) hchunks[i] = chunks[heavy[i]];
}

}  // namespace

// Vertex-major Brandes over the sources, 64 per batch (one reference chunk
// per batch); bc accumulated into out (device).
static void betweenness_vm(ursa_hip_topo* t, const std::vector<uint32_t>& src, double* out, hipStream_t s,
                           uint64_t* reached_edges) {
    const uint64_t n = t->n;
    const DevCsr& o = t->out;
    // Heavy vertices and their chunk bases, for the forward (VM_FHEAVY) and the
    // backward (VM_HEAVY) kernels.
    struct HeavySet {
        DevBuf<uint32_t> heavy, cbase;
        uint32_t nh = 0, nchunks = 0;
    };
    auto heavy_set = [&](uint32_t thr) {
        HeavySet hs;
        DevBuf<uint8_t> flag(n);
        DevBuf<uint32_t> iota(n), chunks(n);
        DevBuf<uint64_t> dcnt(1);
        hs.heavy.alloc(n);
        k_vm_heavy_flags<<<blocks_for(n, 256, 1u << 16), 256, 0, s>>>(o.off.get(), n, thr, flag.get(), iota.get(),
                                                                      chunks.get());
        check_launch();
        prim::select_flagged_u32(iota.get(), flag.get(), hs.heavy.get(), dcnt.get(), n, s);
        uint64_t nh64 = 0;
        HIP_OK(hipMemcpyAsync(&nh64, dcnt.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        hs.nh = (uint32_t)nh64;
        hs.cbase.alloc(hs.nh + 1);
        if (hs.nh) {
            DevBuf<uint32_t> hch(hs.nh);
            k_vm_gather_chunks<<<blocks_for(hs.nh, 256), 256, 0, s>>>(hs.heavy.get(), hs.nh, chunks.get(), hch.get());
            check_launch();
            prim::exclusive_sum_u32(hch.get(), hs.cbase.get(), hs.nh, s);
            HIP_OK(hipMemcpyAsync(&hs.nchunks, hs.cbase.get() + hs.nh, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
        }
        return hs;
    };
    HeavySet fwd_heavy = heavy_set(VM_FHEAVY);
    HeavySet bwd_heavy = heavy_set(VM_HEAVY);
    DevBuf<uint32_t>& heavy = bwd_heavy.heavy;
    DevBuf<uint32_t>& cbase = bwd_heavy.cbase;
    const uint32_t nh = bwd_heavy.nh, nchunks = bwd_heavy.nchunks;
    DevBuf<int32_t> dist(n * 64);
    DevBuf<double> sigma(n * 64), delta(n * 64), part((uint64_t)std::max<uint32_t>(nchunks, 1) * 64);
    DevBuf<uint32_t> mark(n), order(n * 64), dsrc(64);
    DevBuf<VmCtr> ctr(1);
    DevBuf<VmCtl> ctl(1);
    DevBuf<unsigned long long> dlvl(n + 2);
    VmCtl* h = nullptr;
    HIP_OK(hipHostMalloc((void**)&h, sizeof(VmCtl), hipHostMallocDefault));
    auto read = [&]() {
        HIP_OK(hipMemcpyAsync(h, ctl.get(), sizeof(VmCtl), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return *h;
    };
    uint64_t total_edges = 0;
    const unsigned hgrid = blocks_for(std::max<uint32_t>(nchunks, 1), BW, GRID_CAP);
    const unsigned fhgrid = blocks_for(std::max<uint32_t>(fwd_heavy.nchunks, 1), BW, GRID_CAP);
    const unsigned fgrid = blocks_for(n, BW, GRID_CAP);
    // Forward levels in groups of 8, every launch reading its level, queue range
    // and counters from the control block: one recorded graph replayed for every
    // group of every batch. Each vertex enters the list of a level at most once
    // and reaches at most 64 distinct levels over the batch, so order[] (64 n)
    // cannot overflow. `recount` adds the ordered path-count pull after each
    // level (taken only when a sum reached 2^53).
    const DevCsr* in_csr = nullptr;
    auto group = [&](bool recount) {
        for (int i = 0; i < 8; ++i) {
            k_vm_level_step<<<1, 1, 0, s>>>(ctl.get(), ctr.get(), dlvl.get());
            timed("brandes_forward", 0.0, s, [&] {
                k_vm_forward<<<fgrid, BB, 0, s>>>(order.get(), ctl.get(), o.off.get(), o.tgt.get(), dist.get(),
                                                  sigma.get(), mark.get(), ctr.get());
                if (fwd_heavy.nchunks)
                    k_vm_forward_heavy<<<fhgrid, BB, 0, s>>>(fwd_heavy.heavy.get(), fwd_heavy.cbase.get(), fwd_heavy.nh,
                                                             fwd_heavy.nchunks, o.off.get(), o.tgt.get(), dist.get(),
                                                             sigma.get(), mark.get(), order.get(), ctl.get(),
                                                             ctr.get());
            });
            if (recount)
                k_vm_sigma_pull<<<fgrid, BB, 0, s>>>(order.get(), ctl.get(), ctr.get(), in_csr->off.get(),
                                                     in_csr->tgt.get(), dist.get(), sigma.get());
            check_launch();
        }
    };
    CapturedLaunches fast;
    try {
        for (uint64_t b0 = 0; b0 < src.size(); b0 += 64) {
            const uint32_t nb = (uint32_t)std::min<uint64_t>(64, src.size() - b0);
            HIP_OK(hipMemcpyAsync(dsrc.get(), src.data() + b0, nb * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            VmCtl c{};
            for (int pass = 0; pass < 2; ++pass) {
                const bool recount = pass == 1;
                k_vm_init<<<blocks_for(n * 64, 256, 1u << 16), 256, 0, s>>>(dist.get(), sigma.get(), delta.get(),
                                                                           mark.get(), n);
                k_vm_seed<<<1, 1, 0, s>>>(dsrc.get(), nb, n, dist.get(), sigma.get(), mark.get(), order.get(), ctr.get(),
                                          ctl.get(), dlvl.get());
                check_launch();
                if (recount) {
                    in_csr = &t->incoming(s);
                    do {
                        group(true);
                    } while (!read().done);
                } else {
                    do {
                        fast.run(s, [&] { group(false); });
                    } while (!read().done);
                }
                c = *h;
                if (!c.big) break;  // no sum reached 2^53: the counts are exact
            }
            total_edges += c.edges;
            std::vector<uint64_t> lvl((size_t)c.levels + 1);
            HIP_OK(hipMemcpyAsync(lvl.data(), dlvl.get(), lvl.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            const int32_t levels = (int32_t)lvl.size() - 1;
            for (int32_t L = levels - 2; L >= 1; --L) {
                const uint64_t a = lvl[L], z = lvl[L + 1];
                const uint32_t qlen = (uint32_t)(z - a);
                timed("brandes_backward", 0.0, s, [&] {
                    if (qlen)
                        k_vm_backward<<<blocks_for(qlen, BW, 1u << 16), BB, 0, s>>>(
                            order.get() + a, qlen, o.off.get(), o.tgt.get(), dist.get(), sigma.get(), delta.get(), L);
                    if (nchunks) {
                        k_vm_backward_heavy<<<hgrid, BB, 0, s>>>(heavy.get(), cbase.get(), nh, nchunks, o.off.get(),
                                                                 o.tgt.get(), dist.get(), sigma.get(), delta.get(), L,
                                                                 part.get());
                        k_vm_backward_combine<<<blocks_for((uint64_t)nh * 64, 256), 256, 0, s>>>(
                            heavy.get(), cbase.get(), nh, dist.get(), part.get(), delta.get(), L);
                    }
                });
                check_launch();
            }
            k_vm_accumulate<<<blocks_for(n, 256, 1u << 16), 256, 0, s>>>(dist.get(), delta.get(), nb, n, out);
            check_launch();
        }
    } catch (...) {
        (void)hipHostFree(h);
        throw;
    }
    (void)hipHostFree(h);
    if (reached_edges) *reached_edges += total_edges;
}

namespace {
}  // namespace

// bc over an explicit source list; `batch` sources advance together (1..64).
void betweenness_sources(ursa_hip_topo* t, const uint32_t* sources_user, uint64_t k, bool apply_scale, double scale,
                         uint32_t batch, double* out_user, uint64_t* reached_edges) {
    const uint64_t n = t->n;
    if (reached_edges) *reached_edges = 0;
    if (n == 0) return;
    hipStream_t s = stream();
    DevOut<double> out(out_user, n);
    HIP_OK(hipMemsetAsync(out.p, 0, n * sizeof(double), s));
    std::vector<uint32_t> src(k);
    if (k) {
        if (is_device_ptr(sources_user))
            HIP_OK(hipMemcpy(src.data(), sources_user, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
        else
            std::copy(sources_user, sources_user + k, src.begin());
    }
    batch = std::max<uint32_t>(1, std::min<uint32_t>(batch, 64));
    if (batch == 64 && k > 0) {
        // Batches of 64 = the reference's chunks: vertex-major lanes.
        uint64_t re = 0;
        betweenness_vm(t, src, out.p, s, &re);
        if (apply_scale) {
            k_scale<<<blocks_for(n, 256, 1u << 16), 256, 0, s>>>(out.p, n, scale);
            check_launch();
        }
        out.finish(s);
        if (reached_edges) *reached_edges = re;
        return;
    }
    const uint32_t B = (uint32_t)std::min<uint64_t>(batch, std::max<uint64_t>(k, 1));
    const DevCsr& o = t->out;
    DevBuf<int32_t> dist(B * n);
    DevBuf<double> sigma(B * n), delta(B * n);
    DevBuf<unsigned long long> order(B * n);
    DevBuf<uint32_t> heavy(B * n), dsrc(B);
    DevBuf<BrCtr> ctr(1);
    BrCtr* h = nullptr;
    HIP_OK(hipHostMalloc((void**)&h, sizeof(BrCtr), hipHostMallocDefault));
    auto read = [&]() {
        HIP_OK(hipMemcpyAsync(h, ctr.get(), sizeof(BrCtr), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return *h;
    };
    uint64_t total_edges = 0;
    try {
        // Batches never straddle a 64-source chunk, so with batch 64 one batch is
        // one reference chunk; smaller batches are summed into the same chunk
        // partial in source order.
        DevBuf<double> partial(n);
        for (uint64_t c0 = 0; c0 < k; c0 += 64) {
            const uint64_t c1 = std::min<uint64_t>(k, c0 + 64);
            const bool chunk_is_batch = (c1 - c0) <= B;
            double* acc_into = chunk_is_batch ? out.p : partial.get();
            if (!chunk_is_batch) HIP_OK(hipMemsetAsync(partial.get(), 0, n * sizeof(double), s));
            for (uint64_t b0 = c0; b0 < c1; b0 += B) {
                const uint32_t nb = (uint32_t)std::min<uint64_t>(B, c1 - b0);
                k_br_init<<<blocks_for(nb * n, 256, 1u << 16), 256, 0, s>>>(dist.get(), sigma.get(), delta.get(),
                                                                           nb * n);
                HIP_OK(hipMemcpyAsync(dsrc.get(), src.data() + b0, nb * sizeof(uint32_t), hipMemcpyHostToDevice, s));
                HIP_OK(hipMemsetAsync(ctr.get(), 0, sizeof(BrCtr), s));
                k_br_seed<<<1, 1, 0, s>>>(dsrc.get(), nb, n, dist.get(), sigma.get(), order.get(), ctr.get());
                check_launch();
                std::vector<uint64_t> lvl{0, read().len};
                for (int32_t L = 1;; ++L) {
                    const uint64_t a = lvl[L - 1], z = lvl[L];
                    if (z == a) break;
                    HIP_OK(hipMemsetAsync(ctr.get(), 0, sizeof(BrCtr), s));
                    const uint32_t qlen = (uint32_t)(z - a);
                    timed("brandes_forward", 8.0 * (double)qlen, s, [&] {
                        k_br_forward<<<blocks_for((qlen + 63) / 64, BW, 1u << 16), BB, 0, s>>>(
                            order.get() + a, qlen, o.off.get(), o.tgt.get(), n, dist.get(), sigma.get(), L,
                            order.get() + z, ctr.get());
                    });
                    check_launch();
                    const BrCtr c = read();
                    total_edges += c.edges;
                    if (c.big && c.len) {
                        const DevCsr& in = t->incoming(s);
                        k_br_sigma_pull<<<blocks_for(c.len, 256, 1u << 16), 256, 0, s>>>(
                            order.get() + z, c.len, in.off.get(), in.tgt.get(), n, dist.get(), sigma.get(), L);
                        check_launch();
                    }
                    lvl.push_back(z + c.len);
                }
                // Backward: levels deepest-1 .. 1 (level-0 sources need no delta).
                const int32_t levels = (int32_t)lvl.size() - 1;
                for (int32_t L = levels - 1; L >= 1; --L) {
                    const uint64_t a = lvl[L], z = lvl[L + 1];
                    if (z == a) continue;
                    const uint32_t qlen = (uint32_t)(z - a);
                    HIP_OK(hipMemsetAsync(ctr.get(), 0, sizeof(BrCtr), s));
                    timed("brandes_backward", 0.0, s, [&] {
                        k_br_backward<<<blocks_for(qlen, 256, 1u << 16), 256, 0, s>>>(
                            order.get() + a, qlen, o.off.get(), o.tgt.get(), n, dist.get(), sigma.get(), delta.get(),
                            L + 1, heavy.get(), ctr.get());
                        k_br_backward_heavy<<<blocks_for(qlen, BW, 1u << 14), BB, 0, s>>>(
                            order.get() + a, heavy.get(), ctr.get(), o.off.get(), o.tgt.get(), n, dist.get(),
                            sigma.get(), delta.get(), L + 1);
                    });
                    check_launch();
                }
                k_br_accumulate<<<blocks_for(n, 256, 1u << 16), 256, 0, s>>>(delta.get(), dist.get(), nb, n, acc_into);
                check_launch();
            }
            if (!chunk_is_batch) {
                // bc += partial (chunk order)
                k_add<<<blocks_for(n, 256, 1u << 16), 256, 0, s>>>(partial.get(), n, out.p);
                check_launch();
            }
        }
    } catch (...) {
        (void)hipHostFree(h);
        throw;
    }
    (void)hipHostFree(h);
    if (apply_scale) {
        k_scale<<<blocks_for(n, 256, 1u << 16), 256, 0, s>>>(out.p, n, scale);
        check_launch();
    }
    out.finish(s);
    if (reached_edges) *reached_edges = total_edges;
}

}  // namespace ursa

extern "C" {

int ursa_hip_betweenness_sources(ursa_hip_topo* t, const uint32_t* sources, uint64_t k, int apply_scale,
                                 double scale, double* out) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (t->n == 0) return;
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        URSA_REQUIRE(k == 0 || sources != nullptr, URSA_HIP_EINVAL, "sources is NULL");
        ursa::betweenness_sources(t, sources, k, apply_scale != 0, scale, 64, out, nullptr);
    });
}

// betweenness.rs:153-177: sources from sample_sources, n/k scaling when sampled.
int ursa_hip_betweenness(ursa_hip_topo* t, int has_sample, double sample, int has_seed, uint64_t seed, double* out) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        const uint64_t n = t->n;
        if (n == 0) return;  // betweenness.rs:155-157
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        std::vector<uint32_t> src = ursa::sample_sources(n, has_sample != 0, sample, has_seed != 0, seed);
        const uint64_t k = src.size();
        ursa::betweenness_sources(t, src.data(), k, has_sample != 0, (double)n / (double)k, 64, out, nullptr);
    });
}

int ursa_hip_shuffled_order(uint64_t n, uint64_t seed, uint32_t* out) {
    return ursa::guard([&] {
        if (n == 0) return;
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        std::vector<uint32_t> o = ursa::shuffled_order(n, seed);
        std::copy(o.begin(), o.end(), out);
    });
}

int ursa_hip_betweenness_sources_ex(ursa_hip_topo* t, const uint32_t* sources, uint64_t k, int apply_scale,
                                    double scale, uint32_t batch, double* out, uint64_t* reached_edges) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (reached_edges) *reached_edges = 0;
        if (t->n == 0) return;
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        URSA_REQUIRE(k == 0 || sources != nullptr, URSA_HIP_EINVAL, "sources is NULL");
        ursa::betweenness_sources(t, sources, k, apply_scale != 0, scale, batch, out, reached_edges);
    });
}

}  // extern "C"
