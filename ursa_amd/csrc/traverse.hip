// traverse.hip — order-exact frontier traversals: k_hop (ursa-core/src/algo/
// hop.rs:52-130) and unweighted shortest_path (ursa-core/src/algo/bfs.rs:48-102).
//
// Both outputs depend on the serial discovery order: k_hop emits reached nodes
// in the order they are first seen, and shortest_path follows the first
// discoverer of each node. In the serial loop, level-L frontier vertex at
// position p scans its neighbour slots j = 0, 1, ... (out then in for Both), so
// a node v first reached at level L+1 is discovered by the smallest (p, j) over
// the frontier edges that hit it. The device version computes exactly that:
// every frontier edge into a not-yet-visited v does atomicMin(key[v], p<<32|j);
// the new nodes are then sorted by key, which is their serial discovery order,
// and parent[v] = frontier[key >> 32] is their serial first discoverer.
// Outputs are therefore byte-identical to the reference's.
#include <algorithm>
#include <cstring>
#include <vector>

#include "frontier.hpp"
#include "prim.hpp"
#include "traverse.hpp"

struct ursa_hip_pairs {
    std::vector<uint32_t> src, dst;
};

namespace ursa {

void direction_views(ursa_hip_topo* t, int dir, hipStream_t s, NbrView& fwd, NbrView& rev);

namespace {

constexpr int TB = 256;
constexpr int TW = TB / 64;

struct OCtr {
    uint32_t count;
    uint32_t hit;  // target reached (shortest_path)
    // Frontier positions deferred to k_oexpand_heavy (high 32 bits) and the
    // 64-slot chunks of their lists (low 32 bits), taken by one atomic add
    // each, so deferred entry h owns the chunks from its low word on.
    unsigned long long hpack;
};

constexpr uint32_t O_HEAVY = 1024;  // frontier vertices with more edges go grid-wide

// A wave's new candidates wait in LDS (128 per wave) and go to the candidate
// list 64 or more at a time, so the list's counter sees one atomic per
// flush instead of one per 64-slot step.
struct OBuf {
    uint32_t* s;   // this wave's 128 LDS words
    uint32_t cnt;  // wave-uniform
};

__device__ __forceinline__ void o_flush(OBuf& ob, uint32_t* __restrict__ cand, OCtr* ctr) {
    if (ob.cnt == 0) return;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(&ctr->count, ob.cnt);
    base = __shfl(base, 0, 64);
    for (uint32_t i = lane_id(); i < ob.cnt; i += 64) cand[base + i] = ob.s[i];
    ob.cnt = 0;
}

// Appends the wave's fresh candidates to its LDS buffer.
__device__ __forceinline__ void o_push(bool fresh, uint32_t v, uint32_t* __restrict__ cand, OCtr* ctr, OBuf& ob) {
    const uint64_t mask = __ballot(fresh);
    if (!mask) return;
    if (fresh) ob.s[ob.cnt + (uint32_t)__popcll(mask & lanemask_lt())] = v;
    ob.cnt += (uint32_t)__popcll(mask);  // < 128: flushed below once it reaches 64
    if (ob.cnt >= 64) o_flush(ob, cand, ctr);
}

// Offers v, reached by frontier position p through its slot j: atomicMin of
// the discovery key, and the thread that replaced the initial key appends v to
// the candidate list.
__device__ __forceinline__ void o_offer(bool active, uint32_t v, uint64_t p, uint32_t j,
                                        const int32_t* __restrict__ level, unsigned long long* __restrict__ key,
                                        uint32_t* __restrict__ cand, OCtr* ctr, OBuf& ob) {
    bool fresh = false;
    if (active && level[v] < 0) fresh = atomicMin(&key[v], ((unsigned long long)p << 32) | j) == ~0ull;
    o_push(fresh, v, cand, ctr, ob);
}

// Expand frontier F[0..flen): candidates get atomicMin(key, p<<32 | j) and are
// appended to cand once. Frontier vertices with more than O_HEAVY slots are
// deferred to k_oexpand_heavy.
__global__ __launch_bounds__(TB) void k_oexpand(const uint32_t* __restrict__ F, uint32_t flen, NbrView fwd,
                                                const int32_t* __restrict__ level,
                                                unsigned long long* __restrict__ key, uint32_t* __restrict__ cand,
                                                OCtr* ctr, uint32_t* __restrict__ heavy,
                                                uint32_t* __restrict__ hbase) {
    __shared__ uint32_t s_buf[TW][128];
    const int lane = lane_id();
    const uint64_t gw = blockIdx.x * (uint64_t)TW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * TW;
    OBuf ob{s_buf[threadIdx.x >> 6], 0};
    for (uint64_t base = gw * 64; base < flen; base += nwaves * 64) {
        const uint64_t idx = base + lane;
        uint32_t u = 0, d = 0, da = 0, sa = 0, sb = 0;
        if (idx < flen) {
            u = F[idx];
            d = nv_degree(fwd, u, da);
            sa = fwd.off_a[u];
            if (fwd.off_b) sb = fwd.off_b[u];
            if (d > O_HEAVY) {
                const unsigned long long old = atomicAdd(&ctr->hpack, (1ull << 32) | (unsigned long long)((d + 63) / 64));
                heavy[old >> 32] = (uint32_t)idx;
                hbase[old >> 32] = (uint32_t)old;
                d = 0;
            }
        }
        const uint32_t incl = wave_incl_scan(d);
        const uint32_t excl = incl - d;
        const uint32_t total = __shfl(incl, 63, 64);
        for (uint32_t t0 = 0; t0 < total; t0 += 64) {
            const uint32_t t = t0 + lane;
            const int j = wave_owner(excl, t);
            const uint32_t e = t - __shfl(excl, j, 64);
            const uint32_t oa = __shfl(da, j, 64);
            const uint32_t osa = __shfl(sa, j, 64);
            const uint32_t osb = __shfl(sb, j, 64);
            uint32_t v = 0;
            if (t < total) v = e < oa ? fwd.tgt_a[osa + e] : fwd.tgt_b[osb + (e - oa)];
            o_offer(t < total, v, base + j, e, level, key, cand, ctr, ob);
        }
    }
    o_flush(ob, cand, ctr);
}

// Deferred frontier vertices. Their 64-slot chunks are numbered in one
// sequence (entry h owns chunks hbase[h] .. hbase[h+1]-1, the last entry up to
// the total) and every wave takes a contiguous run of that sequence, so a
// hub's slots spread over the grid and no wave walks the deferred list. The
// keys are atomic minima, so which wave offers a slot changes nothing.
__global__ __launch_bounds__(TB) void k_oexpand_heavy(const uint32_t* __restrict__ F, NbrView fwd,
                                                      const int32_t* __restrict__ level,
                                                      unsigned long long* __restrict__ key,
                                                      uint32_t* __restrict__ cand, OCtr* ctr,
                                                      const uint32_t* __restrict__ heavy,
                                                      const uint32_t* __restrict__ hbase) {
    __shared__ uint32_t s_buf[TW][128];
    OBuf ob{s_buf[threadIdx.x >> 6], 0};
    const unsigned long long hp = ctr->hpack;
    const uint32_t cnt = (uint32_t)(hp >> 32), total = (uint32_t)hp;
    if (cnt == 0) return;
    const int lane = lane_id();
    const uint64_t nwaves = (uint64_t)gridDim.x * TW;
    const uint64_t gw = blockIdx.x * (uint64_t)TW + (threadIdx.x >> 6);
    const uint64_t per = (total + nwaves - 1) / nwaves;
    uint32_t c = (uint32_t)min((uint64_t)total, gw * per);
    const uint32_t cend = (uint32_t)min((uint64_t)total, gw * per + per);
    if (c >= cend) return;
    uint32_t lo = 0, hi = cnt;  // largest h with hbase[h] <= c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (hbase[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    for (uint32_t h = lo; c < cend; ++h) {
        const uint32_t p = heavy[h], hb = hbase[h];
        const uint32_t hn = h + 1 < cnt ? hbase[h + 1] : total;
        const uint32_t u = F[p];
        uint32_t da;
        const uint32_t d = nv_degree(fwd, u, da);
        const uint32_t sa = fwd.off_a[u];
        const uint32_t sb = fwd.off_b ? fwd.off_b[u] : 0;
        const uint32_t c1 = min(cend, hn);
        const uint32_t e1 = min(d, (c1 - hb) * 64u);
        // Four 64-slot steps in flight: their targets, then their levels, then the offers.
        for (uint32_t eb = (c - hb) * 64u; eb < e1; eb += 256) {
            uint32_t v[4];
            int32_t lv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t e = eb + 64u * j + lane;
                v[j] = e < e1 ? (e < da ? fwd.tgt_a[sa + e] : fwd.tgt_b[sb + (e - da)]) : 0u;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) lv[j] = eb + 64u * j + lane < e1 ? level[v[j]] : 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t e = eb + 64u * j + lane;
                bool fresh = false;
                if (lv[j] < 0) fresh = atomicMin(&key[v[j]], ((unsigned long long)p << 32) | e) == ~0ull;
                o_push(fresh, v[j], cand, ctr, ob);
            }
        }
        c = c1;
    }
    o_flush(ob, cand, ctr);
}

__global__ void k_ogather_keys(const uint32_t* __restrict__ cand, uint32_t cnt,
                               const unsigned long long* __restrict__ key, uint64_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = key[cand[i]];
}

// Commit a level: the sorted new nodes get their level and parent and their
// key is reset; they are appended to `order` at `at`.
__global__ void k_ocommit(const uint32_t* __restrict__ sorted, const uint64_t* __restrict__ skeys, uint32_t cnt,
                          const uint32_t* __restrict__ F, int32_t* __restrict__ level, int32_t L,
                          unsigned long long* __restrict__ key, uint32_t* __restrict__ parent,
                          uint32_t* __restrict__ order, uint64_t at, uint32_t target, OCtr* ctr) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = sorted[i];
        level[v] = L;
        key[v] = ~0ull;
        if (parent) parent[v] = F[skeys[i] >> 32];
        order[at + i] = v;
        if (v == target) ctr->hit = 1;
    }
}

__global__ void k_oreset(const uint32_t* __restrict__ order, uint64_t cnt, int32_t* __restrict__ level) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x)
        level[order[i]] = -1;
}

__global__ void k_oseed(uint32_t seed, int32_t* level, uint32_t* order, uint32_t* parent) {
    level[seed] = 0;
    order[0] = seed;
    if (parent) parent[seed] = seed;
}

__global__ void k_fill_i32b(int32_t* p, uint64_t n, int32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Scratch for ordered traversals over one topology; level[] is kept all -1
// between traversals (each traversal resets exactly what it touched).
struct OWork {
    uint64_t n;
    DevBuf<int32_t> level;
    DevBuf<unsigned long long> key;
    DevBuf<uint32_t> order, cand, sorted, parent;
    DevBuf<uint64_t> ckeys, skeys;
    DevBuf<uint32_t> heavy, hbase;
    DevBuf<OCtr> ctr;
    OWork(uint64_t n_, bool with_parent, hipStream_t s)
        : n(n_), level(n_), key(n_), order(n_), cand(n_), sorted(n_), ckeys(n_), skeys(n_), heavy(n_), hbase(n_), ctr(1) {
        if (with_parent) parent.alloc(n_);
        k_fill_i32b<<<blocks_for(n, 256), 256, 0, s>>>(level.get(), n, -1);
        k_fill_u64<<<blocks_for(n, 256), 256, 0, s>>>(key.get(), n, ~0ull);
        check_launch();
    }
    OCtr read(hipStream_t s) {
        OCtr* h = static_cast<OCtr*>(pinned_scratch(sizeof(OCtr)));
        HIP_OK(hipMemcpyAsync(h, ctr.get(), sizeof(OCtr), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return *h;
    }
};

// Ordered BFS from `seed` for at most `max_levels` levels (0 = unbounded),
// stopping early once `target` is reached. Leaves the discovery order in
// w.order[0..*visited) and returns whether the target was reached.
bool ordered_bfs(uint32_t seed, uint32_t max_levels, uint32_t target, const NbrView& fwd, OWork& w, hipStream_t s,
                 uint64_t* visited, bool may_have_heavy) {
    k_oseed<<<1, 1, 0, s>>>(seed, w.level.get(), w.order.get(), w.parent.get());
    check_launch();
    uint64_t fstart = 0, flen = 1, total = 1;
    bool hit = false;
    for (int32_t L = 1; flen > 0 && (max_levels == 0 || (uint32_t)L <= max_levels); ++L) {
        HIP_OK(hipMemsetAsync(w.ctr.get(), 0, sizeof(OCtr), s));
        const uint32_t* F = w.order.get() + fstart;
        k_oexpand<<<blocks_for((flen + 63) / 64, TW, 1u << 16), TB, 0, s>>>(F, (uint32_t)flen, fwd, w.level.get(),
                                                                           w.key.get(), w.cand.get(), w.ctr.get(),
                                                                           w.heavy.get(), w.hbase.get());
        if (may_have_heavy)
            k_oexpand_heavy<<<2048, TB, 0, s>>>(F, fwd, w.level.get(), w.key.get(), w.cand.get(), w.ctr.get(),
                                                w.heavy.get(), w.hbase.get());
        check_launch();
        const uint32_t cnt = w.read(s).count;
        if (cnt == 0) break;
        k_ogather_keys<<<blocks_for(cnt, 256), 256, 0, s>>>(w.cand.get(), cnt, w.key.get(), w.ckeys.get());
        check_launch();
        prim::sort_pairs_u64_u32(w.ckeys.get(), w.skeys.get(), w.cand.get(), w.sorted.get(), cnt,
                                 32 + bits_for(flen), s);
        k_ocommit<<<blocks_for(cnt, 256), 256, 0, s>>>(w.sorted.get(), w.skeys.get(), cnt, F, w.level.get(), L,
                                                      w.key.get(), w.parent.get(), w.order.get(), total, target,
                                                      w.ctr.get());
        check_launch();
        fstart = total;
        flen = cnt;
        total += cnt;
        if (target != 0xFFFFFFFFu && w.read(s).hit) {
            hit = true;
            break;
        }
    }
    *visited = total;
    return hit;
}

void reset_touched(OWork& w, uint64_t visited, hipStream_t s) {
    k_oreset<<<blocks_for(visited, 256), 256, 0, s>>>(w.order.get(), visited, w.level.get());
    check_launch();
}

// Walks the parent links from tgt back to src (at most cap nodes), writing the
// path target-first.
__global__ void k_opath(const uint32_t* __restrict__ parent, uint32_t src, uint32_t tgt, uint32_t* __restrict__ path,
                        uint32_t cap, uint32_t* __restrict__ len) {
    uint32_t cur = tgt, l = 0;
    path[l++] = cur;
    while (cur != src && l < cap) {
        cur = parent[cur];
        path[l++] = cur;
    }
    *len = l;
}

// ---- cross-graph passes: one workgroup runs one query's whole traversal ----
//
// For graphs of at most TRAV_MULTI_MAX_N nodes a query (k-hop from each of its
// seeds in turn, or one shortest path) is a workgroup of WT threads, and one
// launch runs every query of a pass. The serial order is kept without a sort
// and without global atomics: a level's frontier edges are walked in (frontier
// position, slot) order, WT at a time; among the edges of one batch that reach
// a node not yet on any level, an LDS table keeps the smallest edge index per
// node, and that edge is the node's serial first discoverer (edges of earlier
// batches came first and have already put their nodes on the level). The
// winners are appended in edge order by a block scan. Only this workgroup
// touches the query's level, order and parent arrays, so its barriers order
// every access (the level array is -1 again after each traversal).
constexpr int WT = 1024;
constexpr uint32_t WT_HASH = 2048;

struct TravProb {
    NbrView fwd;
    const uint32_t* seeds;  // k-hop: its seeds; path: source, target
    int32_t* level;         // n, -1 on entry and exit
    uint32_t* order;        // n: discovery order of the current traversal
    uint32_t* parent;       // n (paths)
    uint32_t* out;          // k-hop: reached nodes of every seed in turn; path: source..target
    uint32_t* ends;         // k-hop: out's length after each seed; path: [0] = path length (0 = none)
    uint32_t n, nseeds, k;  // k-hop depth (paths: 0 = unbounded)
    uint32_t path;
};

// Exclusive block scan over WT threads; *total receives the sum (all threads).
__device__ __forceinline__ uint32_t wt_excl_scan(uint32_t x, uint32_t* s_wave, uint32_t* s_tot) {
    const uint32_t incl = wave_incl_scan(x);
    const int lane = lane_id(), wave = (int)(threadIdx.x >> 6);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        const uint32_t w = lane < WT / 64 ? s_wave[lane] : 0u;
        const uint32_t wi = wave_incl_scan(w);
        if (lane < WT / 64) s_wave[lane] = wi - w;
        if (lane == WT / 64 - 1) *s_tot = wi;
    }
    __syncthreads();
    return s_wave[wave] + incl - x;
}

__global__ __launch_bounds__(WT) void k_otrav_wg(const TravProb* __restrict__ probs) {
    __shared__ uint32_t s_pre[WT], s_u[WT], s_da[WT];
    __shared__ uint32_t s_hkey[WT_HASH], s_hmin[WT_HASH];
    __shared__ uint32_t s_wave[WT / 64];
    __shared__ uint32_t s_tot, s_hit;
    const TravProb Q = probs[blockIdx.x];
    const uint32_t tid = threadIdx.x, n = Q.n;
    const NbrView fwd = Q.fwd;
    const uint32_t runs = Q.path ? 1u : Q.nseeds;
    const uint32_t target = Q.path ? Q.seeds[1] : 0xFFFFFFFFu;
    uint32_t outpos = 0;
    for (uint32_t si = 0; si < runs; ++si) {
        const uint32_t seed = Q.seeds[si];
        if (seed >= n || (Q.path && target >= n) || (!Q.path && Q.k == 0)) {  // hop.rs:53-56, 91-93; bfs.rs:55-57
            if (tid == 0) Q.ends[si] = outpos;
            continue;
        }
        if (tid == 0) {
            Q.level[seed] = 0;
            Q.order[0] = seed;
            if (Q.path) Q.parent[seed] = seed;
            s_hit = 0u;
        }
        __syncthreads();
        uint32_t fs = 0, flen = 1, filled = 1;
        bool hit = false;
        for (int32_t L = 1; flen > 0 && (Q.k == 0 || (uint32_t)L <= Q.k) && !hit; ++L) {
            uint32_t added = 0;
            for (uint32_t c0 = 0; c0 < flen && !hit; c0 += WT) {
                uint32_t u = 0, d = 0, da = 0;
                if (c0 + tid < flen) {
                    u = Q.order[fs + c0 + tid];
                    d = nv_degree(fwd, u, da);
                }
                const uint32_t pre = wt_excl_scan(d, s_wave, &s_tot);
                const uint32_t E = s_tot;
                s_pre[tid] = pre;
                s_u[tid] = u;
                s_da[tid] = da;
                __syncthreads();
                for (uint32_t e0 = 0; e0 < E && !hit; e0 += WT) {
                    for (uint32_t h = tid; h < WT_HASH; h += WT) {
                        s_hkey[h] = 0xFFFFFFFFu;
                        s_hmin[h] = 0xFFFFFFFFu;
                    }
                    __syncthreads();
                    const uint32_t e = e0 + tid;
                    uint32_t v = 0, owner = 0;
                    int slot = -1;
                    if (e < E) {
                        uint32_t lo = 0, hi = WT;  // largest owner with s_pre[owner] <= e
                        while (hi - lo > 1) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (s_pre[mid] <= e)
                                lo = mid;
                            else
                                hi = mid;
                        }
                        owner = lo;
                        const uint32_t j = e - s_pre[owner], uu = s_u[owner], a = s_da[owner];
                        v = j < a ? fwd.tgt_a[fwd.off_a[uu] + j] : fwd.tgt_b[fwd.off_b[uu] + (j - a)];
                        if (Q.level[v] < 0) {
                            uint32_t h = (v * 0x9E3779B1u) >> 21;  // 11 bits: WT_HASH slots
                            for (uint32_t step = 0; step < WT_HASH; ++step, h = (h + 1) & (WT_HASH - 1)) {
                                const uint32_t k = atomicCAS(&s_hkey[h], 0xFFFFFFFFu, v);
                                if (k == 0xFFFFFFFFu || k == v) {
                                    atomicMin(&s_hmin[h], e);
                                    slot = (int)h;
                                    break;
                                }
                            }
                        }
                    }
                    __syncthreads();
                    const bool win = slot >= 0 && s_hmin[slot] == e;
                    const uint32_t pos = wt_excl_scan(win ? 1u : 0u, s_wave, &s_tot);
                    const uint32_t cnt = s_tot;
                    if (win) {
                        Q.order[filled + pos] = v;
                        Q.level[v] = L;
                        if (Q.path) {
                            Q.parent[v] = s_u[owner];
                            if (v == target) s_hit = 1u;
                        }
                    }
                    filled += cnt;
                    added += cnt;
                    __syncthreads();
                    hit = s_hit != 0u;
                }
            }
            fs += flen;
            flen = added;
        }
        if (Q.path) {
            if (tid == 0) {
                uint32_t len = 0;
                if (hit) {
                    uint32_t cur = target;
                    len = 1;
                    while (cur != seed && len < n) {
                        cur = Q.parent[cur];
                        ++len;
                    }
                    cur = target;
                    for (uint32_t i = len; i-- > 0;) {
                        Q.out[i] = cur;
                        cur = Q.parent[cur];
                    }
                }
                Q.ends[0] = len;
            }
        } else {
            for (uint32_t i = 1 + tid; i < filled; i += WT) Q.out[outpos + i - 1] = Q.order[i];
            outpos += filled - 1;
            if (tid == 0) Q.ends[si] = outpos;
        }
        __syncthreads();
        for (uint32_t i = tid; i < filled; i += WT) Q.level[Q.order[i]] = -1;
        __syncthreads();
    }
}

}  // namespace

uint64_t traverse_workspace_bytes(uint64_t n) {
    // OWork (level, key, order, cand, sorted, parent, ckeys, skeys, heavy,
    // hbase: 52 B per node) plus the radix sort's temporary storage for up to n
    // (u64, u32) pairs (a double buffer of 12 B per pair and its histograms).
    return n * 84 + (2u << 20);
}

void k_hop_seq(ursa_hip_topo* t, const uint32_t* seeds, uint64_t nseeds, uint32_t k, int dir,
               std::vector<uint32_t>& dst, std::vector<uint64_t>& ends) {
    const uint64_t n = t->n;
    ends.assign(nseeds, dst.size());
    if (k == 0 || n == 0 || nseeds == 0) return;  // hop.rs:53-56
    hipStream_t s = stream();
    NbrView fwd, rev;
    direction_views(t, dir, s, fwd, rev);
    OWork w(n, false, s);
    for (uint64_t i = 0; i < nseeds; ++i) {
        const uint32_t seed = seeds[i];
        if (seed < n) {  // hop.rs:91-93: unknown seed contributes nothing
            uint64_t visited = 0;
            ordered_bfs(seed, k, 0xFFFFFFFFu, fwd, w, s, &visited, t->m > O_HEAVY);
            const uint64_t reached = visited - 1;
            if (reached) {
                const size_t at = dst.size();
                dst.resize(at + reached);
                HIP_OK(hipMemcpyAsync(dst.data() + at, w.order.get() + 1, reached * sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, s));
            }
            reset_touched(w, visited, s);
            HIP_OK(hipStreamSynchronize(s));
        }
        ends[i] = dst.size();
    }
}

void k_hop(ursa_hip_topo* t, const uint32_t* seeds_user, uint64_t nseeds, uint32_t k, int dir, ursa_hip_pairs* res) {
    std::vector<uint32_t> seeds(nseeds);
    if (nseeds && is_device_ptr(seeds_user))
        HIP_OK(hipMemcpy(seeds.data(), seeds_user, nseeds * sizeof(uint32_t), hipMemcpyDeviceToHost));
    else if (nseeds)
        std::copy(seeds_user, seeds_user + nseeds, seeds.begin());
    std::vector<uint64_t> ends;
    k_hop_seq(t, seeds.data(), nseeds, k, dir, res->dst, ends);
    uint64_t at = 0;
    res->src.reserve(res->dst.size());
    for (uint64_t i = 0; i < nseeds; ++i) {
        res->src.insert(res->src.end(), ends[i] - at, seeds[i]);
        at = ends[i];
    }
}

void shortest_paths_seq(ursa_hip_topo* t, const uint32_t* src, const uint32_t* tgt, uint64_t count, int dir,
                        std::vector<std::vector<uint32_t>>& paths) {
    const uint64_t n = t->n;
    paths.assign(count, {});
    bool any = false;
    for (uint64_t i = 0; i < count; ++i) {
        if (src[i] >= n || tgt[i] >= n) continue;  // bfs.rs:55-57
        if (src[i] == tgt[i]) {                    // bfs.rs:58-60
            paths[i].push_back(src[i]);
            continue;
        }
        any = true;
    }
    if (!any) return;
    hipStream_t s = stream();
    NbrView fwd, rev;
    direction_views(t, dir, s, fwd, rev);
    OWork w(n, true, s);
    DevBuf<uint32_t> dpath(n), dlen(1);
    for (uint64_t i = 0; i < count; ++i) {
        if (src[i] >= n || tgt[i] >= n || src[i] == tgt[i]) continue;
        uint64_t visited = 0;
        const bool hit = ordered_bfs(src[i], 0, tgt[i], fwd, w, s, &visited, t->m > O_HEAVY);
        if (hit) {
            k_opath<<<1, 1, 0, s>>>(w.parent.get(), src[i], tgt[i], dpath.get(), (uint32_t)std::min<uint64_t>(n, visited),
                                    dlen.get());
            check_launch();
            uint32_t* h = static_cast<uint32_t*>(pinned_scratch(sizeof(uint32_t)));
            HIP_OK(hipMemcpyAsync(h, dlen.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            const uint32_t len = *h;
            paths[i].resize(len);
            HIP_OK(hipMemcpyAsync(paths[i].data(), dpath.get(), len * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        }
        reset_touched(w, visited, s);
        HIP_OK(hipStreamSynchronize(s));
        if (hit) std::reverse(paths[i].begin(), paths[i].end());
    }
}

namespace {

uint64_t al256(uint64_t b) { return (b + 255) & ~(uint64_t)255; }

// Device bytes of one task's arrays (level, order, parent, out, ends).
uint64_t trav_task_bytes(const TravTask& q) {
    const uint64_t n = q.t->n;
    const uint64_t outs = q.path ? n : (uint64_t)q.seeds.size() * n;
    return al256(n * 4) * (q.path ? 3 : 2) + al256(outs * 4) + al256(std::max<size_t>(q.seeds.size(), 1) * 4);
}

}  // namespace

bool traverse_multi_eligible(const ursa_hip_topo* t) { return t->n > 0 && t->n <= TRAV_MULTI_MAX_N; }

uint64_t traverse_multi_workspace_bytes(const std::vector<const TravTask*>& tasks) {
    uint64_t b = 0;
    for (const TravTask* q : tasks) b += trav_task_bytes(*q) + al256(q->seeds.size() * 4);
    return b + al256(tasks.size() * sizeof(TravProb)) + (64u << 10);
}

void traverse_multi(const std::vector<TravTask*>& tasks) {
    hipStream_t s = stream();
    std::vector<TravTask*> run;
    for (TravTask* q : tasks) {
        q->dst.clear();
        q->ends.assign(q->path ? 0 : q->seeds.size(), 0);
        q->found.clear();
        const uint64_t n = q->t->n;
        URSA_REQUIRE(traverse_multi_eligible(q->t), URSA_HIP_EINVAL, "graph not eligible for a cross-graph pass");
        if (q->path) {
            URSA_REQUIRE(q->seeds.size() == 2, URSA_HIP_EINVAL, "a path task takes a source and a target");
            if (q->seeds[0] >= n || q->seeds[1] >= n) continue;  // bfs.rs:55-57
            if (q->seeds[0] == q->seeds[1]) {                    // bfs.rs:58-60
                q->found.push_back(q->seeds[0]);
                continue;
            }
        } else if (q->k == 0 || q->seeds.empty()) {  // hop.rs:53-56
            continue;
        }
        run.push_back(q);
    }
    if (run.empty()) return;
    const uint32_t np = (uint32_t)run.size();
    // One state allocation: every task's level array first (one fill to -1),
    // then ends and outs of every task (one copy back), then order and parent.
    uint64_t lv_bytes = 0, eo_bytes = 0, rest = 0;
    for (TravTask* q : run) {
        const uint64_t n = q->t->n;
        lv_bytes += al256(n * 4);
        eo_bytes += al256(std::max<size_t>(q->seeds.size(), 1) * 4) + al256((q->path ? n : q->seeds.size() * n) * 4);
        rest += al256(n * 4) * (q->path ? 2 : 1);
    }
    DevBuf<char> state(lv_bytes + eo_bytes + rest);
    std::vector<TravProb> tab(np);
    std::vector<uint32_t> seeds;
    std::vector<uint64_t> ends_at(np), out_at(np), seed_at(np);
    uint64_t lv = 0, eo = lv_bytes, rs = lv_bytes + eo_bytes;
    for (uint32_t i = 0; i < np; ++i) {
        TravTask* q = run[i];
        const uint64_t n = q->t->n;
        NbrView fwd, rev;
        direction_views(q->t, q->dir, s, fwd, rev);
        TravProb& P = tab[i];
        P.fwd = fwd;
        P.n = (uint32_t)n;
        P.nseeds = (uint32_t)q->seeds.size();
        P.k = q->path ? 0u : q->k;
        P.path = q->path ? 1u : 0u;
        P.level = reinterpret_cast<int32_t*>(state.get() + lv);
        lv += al256(n * 4);
        ends_at[i] = eo;
        P.ends = reinterpret_cast<uint32_t*>(state.get() + eo);
        eo += al256(std::max<size_t>(q->seeds.size(), 1) * 4);
        out_at[i] = eo;
        P.out = reinterpret_cast<uint32_t*>(state.get() + eo);
        eo += al256((q->path ? n : q->seeds.size() * n) * 4);
        P.order = reinterpret_cast<uint32_t*>(state.get() + rs);
        rs += al256(n * 4);
        P.parent = nullptr;
        if (q->path) {
            P.parent = reinterpret_cast<uint32_t*>(state.get() + rs);
            rs += al256(n * 4);
        }
        seed_at[i] = seeds.size();
        seeds.insert(seeds.end(), q->seeds.begin(), q->seeds.end());
    }
    const uint64_t tab_bytes = al256(np * sizeof(TravProb));
    DevBuf<char> blob(tab_bytes + seeds.size() * 4);
    const uint32_t* d_seeds = reinterpret_cast<const uint32_t*>(blob.get() + tab_bytes);
    for (uint32_t i = 0; i < np; ++i) tab[i].seeds = d_seeds + seed_at[i];
    std::vector<char> img(tab_bytes + seeds.size() * 4);
    std::memcpy(img.data(), tab.data(), np * sizeof(TravProb));
    std::memcpy(img.data() + tab_bytes, seeds.data(), seeds.size() * 4);
    HIP_OK(hipMemcpyAsync(blob.get(), img.data(), img.size(), hipMemcpyHostToDevice, s));
    k_fill_i32b<<<blocks_for(lv_bytes / 4, 256), 256, 0, s>>>(reinterpret_cast<int32_t*>(state.get()), lv_bytes / 4, -1);
    check_launch();
    k_otrav_wg<<<np, WT, 0, s>>>(reinterpret_cast<const TravProb*>(blob.get()));
    check_launch();
    char* h = static_cast<char*>(pinned_scratch(eo_bytes));
    HIP_OK(hipMemcpyAsync(h, state.get() + lv_bytes, eo_bytes, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < np; ++i) {
        TravTask* q = run[i];
        const uint32_t* ends = reinterpret_cast<const uint32_t*>(h + (ends_at[i] - lv_bytes));
        const uint32_t* out = reinterpret_cast<const uint32_t*>(h + (out_at[i] - lv_bytes));
        if (q->path) {
            q->found.assign(out, out + ends[0]);
        } else {
            const uint32_t total = ends[q->seeds.size() - 1];
            q->dst.assign(out, out + total);
            for (size_t j = 0; j < q->seeds.size(); ++j) q->ends[j] = ends[j];
        }
    }
}

// Returns the path length (0 = None).
uint64_t shortest_path(ursa_hip_topo* t, uint32_t src, uint32_t tgt, int dir, std::vector<uint32_t>& path) {
    std::vector<std::vector<uint32_t>> paths;
    shortest_paths_seq(t, &src, &tgt, 1, dir, paths);
    path.swap(paths[0]);
    return path.size();
}

}  // namespace ursa

using namespace ursa;

extern "C" {

int ursa_hip_k_hop(ursa_hip_topo* t, const uint32_t* seeds, uint64_t nseeds, uint32_t k, int direction,
                   ursa_hip_pairs** out) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr && out != nullptr, URSA_HIP_EINVAL, "topology/out is NULL");
        URSA_REQUIRE(direction >= 0 && direction <= 2, URSA_HIP_EINVAL, "bad direction");
        URSA_REQUIRE(nseeds == 0 || seeds != nullptr, URSA_HIP_EINVAL, "seeds is NULL");
        *out = nullptr;
        auto* r = new ursa_hip_pairs;
        try {
            ursa::k_hop(t, seeds, nseeds, k, direction, r);
        } catch (...) {
            delete r;
            throw;
        }
        *out = r;
    });
}

int ursa_hip_pairs_size(const ursa_hip_pairs* p, uint64_t* count) {
    return guard([&] {
        URSA_REQUIRE(p != nullptr && count != nullptr, URSA_HIP_EINVAL, "NULL argument");
        *count = p->src.size();
    });
}

int ursa_hip_pairs_export(const ursa_hip_pairs* p, uint32_t* src, uint32_t* dst) {
    return guard([&] {
        URSA_REQUIRE(p != nullptr, URSA_HIP_EINVAL, "pairs handle is NULL");
        const uint64_t c = p->src.size();
        if (!c) return;
        if (src) std::copy(p->src.begin(), p->src.end(), src);
        if (dst) std::copy(p->dst.begin(), p->dst.end(), dst);
    });
}

int ursa_hip_pairs_destroy(ursa_hip_pairs* p) {
    delete p;
    return URSA_HIP_OK;
}

int ursa_hip_shortest_path(ursa_hip_topo* t, uint32_t source, uint32_t target, int direction, uint32_t* path,
                           uint64_t cap, uint64_t* len) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr && len != nullptr, URSA_HIP_EINVAL, "topology/len is NULL");
        URSA_REQUIRE(direction >= 0 && direction <= 2, URSA_HIP_EINVAL, "bad direction");
        std::vector<uint32_t> p;
        *len = ursa::shortest_path(t, source, target, direction, p);
        URSA_REQUIRE(p.size() <= cap, URSA_HIP_ERANGE, "path longer than the output capacity (see *len)");
        if (!p.empty()) {
            URSA_REQUIRE(path != nullptr, URSA_HIP_EINVAL, "path is NULL");
            std::copy(p.begin(), p.end(), path);
        }
    });
}

}  // extern "C"
