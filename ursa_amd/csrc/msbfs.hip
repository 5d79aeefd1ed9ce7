// msbfs.hip — multi-source BFS levels, 64 sources per u64 lane mask (the
// batched form of bfs_distances, ursa-core/src/algo/bfs.rs:18-42).
//
// Bit b of visit[v] / front[v] belongs to source b of the batch. Each level is
// one launch of either
//   pull  (k_ms_pull): every vertex with unvisited lanes ORs the front masks
//         of its reverse neighbours, stopping once all its unvisited lanes are
//         covered; the new lanes get level L. Writes next[] for every vertex,
//         so no clearing is needed.
//   push  (k_ms_push): for sparse frontiers, each frontier vertex pushes its
//         mask along its forward edges with an atomic OR into visit[]; the
//         lanes an OR newly sets are exactly the lanes claimed by this edge, so
//         every (source, vertex) level is written once. The first lane mask a
//         vertex receives appends it to the next queue.
// A lane's level is its hop distance from its source, so the result is the
// same i32 array the reference's serial loop gives for each source, for any mix
// of push and pull levels. All 64 traversals share each edge scan.
//
// The level loop runs without the host: a one-thread kernel closes the last
// level and opens the next (level number, done test, compaction, push or
// pull) in a device control block, every level kernel reads its arguments from
// that block and returns at once when it is not chosen, and a level's push or
// pull is one expand launch (plus one for its heavy vertices). The host queues
// levels four at a time (recorded once as a HIP graph and replayed for every
// group of every batch) and reads the done flag once per group.
//
// Algorithmic bytes counted per source as in single-source BFS (4 B per
// reached edge + 12 B per node); the batch shares them.
#include <functional>
#include <algorithm>
#include <vector>

#include "frontier.hpp"
#include "prim.hpp"

namespace ursa {

namespace {

constexpr int MB = 256;
constexpr int MW = MB / 64;

struct MsCtr {
    uint32_t qlen;
    uint32_t pad;
    // Push-frontier vertices deferred to the grid-wide pass (high 32 bits) and
    // the 64-slot chunks of their lists (low 32 bits), one atomic add each.
    unsigned long long hpack;
    unsigned long long newbits;  // (source, vertex) pairs settled this level
    unsigned long long edges;    // forward degree sum of the next frontier (push levels)
};

constexpr uint32_t MS_HEAVY = 1024;
constexpr int MS_LEVELS_PER_GROUP = 4;

// Device-side state of one batch's level loop.
struct MsCtl {
    int32_t L;                  // the level being built
    uint32_t done;
    uint32_t push;              // this level pushes (else it pulls)
    uint32_t compact;           // this level first compacts the frontier into a queue
    uint32_t qsel;              // queue buffer (0 / 1) holding the frontier queue
    uint32_t have_queue;        // the frontier is known as a queue
    uint32_t known_edges;       // fedges is exact
    uint32_t fresh;             // no level has run yet (nothing to close)
    unsigned long long qlen;    // frontier queue length, or after a pull the settled-pair count
    unsigned long long fedges;  // forward degree sum of the frontier queue
    unsigned long long valid;   // lanes whose source is in range
    int32_t* levels;            // the batch's level rows
};

__device__ __forceinline__ void write_levels(uint64_t bits, int32_t* levels, uint64_t n, uint64_t v, int32_t L) {
    while (bits) {
        const int b = __ffsll((unsigned long long)bits) - 1;
        levels[(uint64_t)b * n + v] = L;
        bits &= bits - 1;
    }
}


// Push one frontier edge u -> v carrying lane mask fm; returns whether this
// edge gave v its first mask of the level.
__device__ __forceinline__ bool ms_push_edge(uint32_t v, unsigned long long fm, unsigned long long* visit,
                                             unsigned long long* next, int32_t* levels, uint64_t n, int32_t L,
                                             unsigned long long& cnt) {
    const unsigned long long bits = fm & ~visit[v];
    if (!bits) return false;
    const unsigned long long old = atomicOr(&visit[v], bits);
    const unsigned long long mine = bits & ~old;
    if (!mine) return false;
    write_levels(mine, levels, n, v, L);
    cnt += __popcll(mine);
    return atomicOr(&next[v], mine) == 0;
}

__global__ void k_ms_fill_levels(int32_t* levels, uint64_t total) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x)
        levels[i] = -1;
}

// Seed: visit/front bits of each source, level 0, unique source queue (in
// queue buffer 0), and the control block of the batch's level loop.
__global__ void k_ms_seed(const uint32_t* src, uint32_t nb, uint64_t n, NbrView fwd, unsigned long long* visit,
                          unsigned long long* front, int32_t* levels, uint32_t* queue, MsCtl* ctl) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t q = 0;
    unsigned long long edges = 0, valid = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t s = src[b];
        if (s >= n) continue;
        valid |= 1ull << b;
        if (front[s] == 0) {
            queue[q++] = s;
            edges += nv_degree(fwd, s);
        }
        front[s] |= 1ull << b;
        visit[s] |= 1ull << b;
        levels[(uint64_t)b * n + s] = 0;
    }
    MsCtl c{};
    c.L = 0;
    c.qsel = 0;
    c.have_queue = 1;
    c.known_edges = 1;
    c.qlen = q;
    c.fedges = edges;
    c.valid = valid;
    c.levels = levels;
    c.fresh = 1;
    *ctl = c;
}

__global__ void k_ms_zero_u64(unsigned long long* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Closes the previous level (the next frontier's size from its counters) and
// opens the next: its number, the done test (empty frontier), and whether it
// compacts first and whether it pushes or pulls. A level pushes while its
// frontier's forward edges number fewer than fwd_edges / push_div and the
// frontier is a known queue, and pulls otherwise; after a pull only the
// settled-pair count is known, which bounds the new frontier's vertex count,
// so pulls continue until that count shows a small frontier, and a compaction
// then measures it.
__global__ void k_ms_level_step(MsCtl* ctl, MsCtr* ctr, uint64_t n, unsigned long long fwd_edges,
                                uint32_t push_div) {
    MsCtl c = *ctl;
    if (c.done) return;
    if (!c.fresh) {
        if (c.push) {
            c.qlen = ctr->qlen;
            c.fedges = ctr->edges;
            c.known_edges = 1;
            c.have_queue = 1;
            c.qsel ^= 1u;
        } else {
            c.qlen = ctr->newbits;
            c.have_queue = 0;
            c.known_edges = 0;
        }
    }
    c.fresh = 0;
    *ctr = MsCtr{};
    c.L += 1;
    if (c.qlen == 0) {
        c.done = 1;
    } else {
        c.compact = !c.have_queue && c.qlen * 64 < n;
        if (!c.compact)
            c.push = c.have_queue && c.known_edges ? c.fedges * push_div < fwd_edges : (c.have_queue && c.L == 1);
    }
    *ctl = c;
}

// After a compaction: the measured queue decides push or pull.
__global__ void k_ms_after_compact(MsCtl* ctl, MsCtr* ctr, unsigned long long fwd_edges, uint32_t push_div) {
    MsCtl c = *ctl;
    if (c.done || !c.compact) return;
    c.qlen = ctr->qlen;
    c.fedges = ctr->edges;
    c.have_queue = 1;
    c.known_edges = 1;
    if (c.qlen == 0)
        c.done = 1;
    else
        c.push = c.fedges * push_div < fwd_edges;
    *ctr = MsCtr{};
    *ctl = c;
}

// next[] = 0 before a push level.
__global__ void k_ms_clear_next(const MsCtl* ctl, unsigned long long* next, uint64_t n) {
    if (ctl->done || !ctl->push) return;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        next[i] = 0;
}

// Pull-side vertices with more reverse edges than this are ORed chunk by chunk
// by whole waves (k_ms_pull_heavy); OR is order-free, so the result is exact.
constexpr uint32_t PULL_HEAVY = 256;
constexpr uint32_t PULL_CHUNK = 4096;

// OR of the frontier masks over list slots [k, ke) until every wanted lane is
// covered; four slots (and their mask reads) in flight per step.
__device__ __forceinline__ unsigned long long ms_pull_list(const uint32_t* __restrict__ tgt, uint32_t k, uint32_t ke,
                                                           const unsigned long long* __restrict__ front,
                                                           unsigned long long want) {
    unsigned long long acc = 0;
    for (; k + 4 <= ke; k += 4) {
        const uint32_t w0 = tgt[k], w1 = tgt[k + 1], w2 = tgt[k + 2], w3 = tgt[k + 3];
        acc |= front[w0] | front[w1] | front[w2] | front[w3];
        if ((acc & want) == want) return acc;
    }
    for (; k < ke; ++k) {
        acc |= front[tgt[k]];
        if ((acc & want) == want) break;
    }
    return acc;
}

__device__ __forceinline__ void ms_pull(uint64_t n, NbrView rev, const unsigned long long* __restrict__ front,
                                                unsigned long long* __restrict__ next,
                                                unsigned long long* __restrict__ visit, const MsCtl* ctl,
                                                MsCtr* ctr) {
    const unsigned long long valid = ctl->valid;
    int32_t* levels = ctl->levels;
    const int32_t L = ctl->L;
    unsigned long long cnt = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        if (nv_degree(rev, (uint32_t)v) > PULL_HEAVY) continue;  // k_ms_pull_heavy + k_ms_pull_finish
        const unsigned long long vis = visit[v];
        const unsigned long long want = valid & ~vis;
        unsigned long long acc = 0;
        if (want) {
            acc = ms_pull_list(rev.tgt_a, rev.off_a[v], rev.off_a[v + 1], front, want);
            if (rev.off_b && (acc & want) != want)
                acc |= ms_pull_list(rev.tgt_b, rev.off_b[v], rev.off_b[v + 1], front, want & ~acc);
            acc &= want;
        }
        next[v] = acc;
        if (acc) {
            visit[v] = vis | acc;
            write_levels(acc, levels, n, v, L);
            cnt += __popcll(acc);
        }
    }
    block_add<MB>(cnt, &ctr->newbits);
}

// One wave per chunk of a heavy vertex's reverse list (out list then in list
// for Both): OR of the frontier masks, merged into hacc[h] with atomic OR.
__device__ __forceinline__ void ms_pull_heavy(const uint32_t* __restrict__ heavy,
                                                      const uint32_t* __restrict__ cbase, uint32_t nh, uint32_t chunks,
                                                      NbrView rev, const unsigned long long* __restrict__ front,
                                                      const unsigned long long* __restrict__ visit, const MsCtl* ctl,
                                                      unsigned long long* __restrict__ hacc) {
    const unsigned long long valid = ctl->valid;
    const uint64_t gw = blockIdx.x * (uint64_t)MW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * MW;
    for (uint64_t c = gw; c < chunks; c += nwaves) {
        uint32_t lo = 0, hi = nh;  // owner: largest h with cbase[h] <= c
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cbase[mid] <= c)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t h = lo, v = heavy[h];
        const unsigned long long want = valid & ~visit[v];
        if (!want) continue;
        uint32_t da;
        const uint32_t d = nv_degree(rev, v, da);
        const uint32_t e0 = ((uint32_t)c - cbase[h]) * PULL_CHUNK;
        const uint32_t e1 = min(d, e0 + PULL_CHUNK);
        unsigned long long acc = 0;
        const uint32_t sa = rev.off_a[v], sb = rev.off_b ? rev.off_b[v] : 0u;
        // Four 64-slot steps in flight: their slots, then their masks.
        for (uint32_t eb = e0; eb < e1; eb += 256) {
            uint32_t w[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t e = eb + 64u * r + lane_id();
                w[r] = e < e1 ? (e < da ? rev.tgt_a[sa + e] : rev.tgt_b[sb + (e - da)]) : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (w[r] != 0xFFFFFFFFu) acc |= front[w[r]];
        }
        for (int off = 32; off > 0; off >>= 1) acc |= __shfl_down(acc, off, 64);
        acc &= want;
        if (lane_id() == 0 && acc) atomicOr(&hacc[h], acc);
    }
}

__global__ void k_ms_pull_finish(const uint32_t* __restrict__ heavy, uint32_t nh,
                                 unsigned long long* __restrict__ hacc, unsigned long long* __restrict__ next,
                                 unsigned long long* __restrict__ visit, uint64_t n, const MsCtl* ctl, MsCtr* ctr) {
    if (ctl->done || ctl->push) return;
    int32_t* levels = ctl->levels;
    const int32_t L = ctl->L;
    const uint64_t h = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    unsigned long long cnt = 0;
    if (h < nh) {
        const uint32_t v = heavy[h];
        const unsigned long long acc = hacc[h] & ~visit[v];
        hacc[h] = 0;
        next[v] = acc;
        if (acc) {
            visit[v] |= acc;
            write_levels(acc, levels, n, v, L);
            cnt = __popcll(acc);
        }
    }
    cnt = wave_sum_u64(cnt);
    if (lane_id() == 0 && cnt) atomicAdd(&ctr->newbits, cnt);
}

__global__ void k_ms_heavy_flags(NbrView rev, uint64_t n, uint8_t* flag, uint32_t* iota, uint32_t* chunks) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = nv_degree(rev, (uint32_t)v);
        flag[v] = d > PULL_HEAVY;
        iota[v] = (uint32_t)v;
        chunks[v] = (d + PULL_CHUNK - 1) / PULL_CHUNK;
    }
}

__global__ void k_ms_gather(const uint32_t* __restrict__ idx, uint32_t cnt, const uint32_t* __restrict__ src,
                            uint32_t* __restrict__ out) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < cnt) out[i] = src[idx[i]];
}

__device__ __forceinline__ void ms_push(uint32_t* __restrict__ q0, uint32_t* __restrict__ q1, NbrView fwd,
                                                const unsigned long long* __restrict__ front,
                                                unsigned long long* __restrict__ next,
                                                unsigned long long* __restrict__ visit, uint64_t n, const MsCtl* ctl,
                                                MsCtr* ctr, uint32_t* __restrict__ heavy,
                                                uint32_t* __restrict__ hbase) {
    const uint32_t* q = ctl->qsel ? q1 : q0;
    uint32_t* nq = ctl->qsel ? q0 : q1;
    const uint64_t qlen = ctl->qlen;
    int32_t* levels = ctl->levels;
    const int32_t L = ctl->L;
    const int lane = lane_id();
    const uint64_t gw = blockIdx.x * (uint64_t)MW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * MW;
    unsigned long long cnt = 0, edges = 0;
    for (uint64_t base = gw * 64; base < qlen; base += nwaves * 64) {
        const uint64_t idx = base + lane;
        uint32_t u = 0, d = 0, da = 0, sa = 0, sb = 0;
        unsigned long long fm = 0;
        if (idx < qlen) {
            u = q[idx];
            d = nv_degree(fwd, u, da);
            sa = fwd.off_a[u];
            if (fwd.off_b) sb = fwd.off_b[u];
            fm = front[u];
            if (d > MS_HEAVY) {
                const unsigned long long old = atomicAdd(&ctr->hpack, (1ull << 32) | (unsigned long long)((d + 63) / 64));
                heavy[old >> 32] = u;
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
            const unsigned long long ofm = __shfl(fm, j, 64);
            bool first = false;
            uint32_t v = 0;
            if (t < total) {
                v = e < oa ? fwd.tgt_a[osa + e] : fwd.tgt_b[osb + (e - oa)];
                first = ms_push_edge(v, ofm, visit, next, levels, n, L, cnt);
            }
            const uint32_t slot = wave_append_slot(first, &ctr->qlen);
            if (first) {
                nq[slot] = v;
                edges += nv_degree(fwd, v);
            }
        }
    }
    cnt = wave_sum_u64(cnt);
    edges = wave_sum_u64(edges);
    if (lane == 0) {
        if (cnt) atomicAdd(&ctr->newbits, cnt);
        if (edges) atomicAdd(&ctr->edges, edges);
    }
}

// Heavy push-frontier vertices. Their 64-slot chunks are numbered in one
// sequence (entry h owns chunks hbase[h] .. hbase[h+1]-1, the last up to the
// total) and every wave takes a contiguous run of it, so hub lists spread over
// the grid and no wave walks the deferred list.
__device__ __forceinline__ void ms_push_heavy(const uint32_t* __restrict__ heavy, const uint32_t* __restrict__ hbase,
                                              NbrView fwd, const unsigned long long* __restrict__ front,
                                              unsigned long long* __restrict__ next,
                                              unsigned long long* __restrict__ visit, uint64_t n,
                                              uint32_t* __restrict__ q0, uint32_t* __restrict__ q1,
                                              const MsCtl* ctl, MsCtr* ctr) {
    uint32_t* nq = ctl->qsel ? q0 : q1;
    int32_t* levels = ctl->levels;
    const int32_t L = ctl->L;
    const unsigned long long hp = ctr->hpack;
    const uint32_t cnt_h = (uint32_t)(hp >> 32), total = (uint32_t)hp;
    if (cnt_h == 0) return;
    const int lane = lane_id();
    const uint64_t nwaves = (uint64_t)gridDim.x * MW;
    const uint64_t gw = blockIdx.x * (uint64_t)MW + (threadIdx.x >> 6);
    const uint64_t per = (total + nwaves - 1) / nwaves;
    uint32_t c = (uint32_t)min((uint64_t)total, gw * per);
    const uint32_t cend = (uint32_t)min((uint64_t)total, gw * per + per);
    unsigned long long cnt = 0, edges = 0;
    if (c < cend) {
        uint32_t lo = 0, hi = cnt_h;  // largest h with hbase[h] <= c
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (hbase[mid] <= c)
                lo = mid;
            else
                hi = mid;
        }
        for (uint32_t h = lo; c < cend; ++h) {
            const uint32_t u = heavy[h], hb = hbase[h];
            const uint32_t hn = h + 1 < cnt_h ? hbase[h + 1] : total;
            uint32_t da;
            const uint32_t d = nv_degree(fwd, u, da);
            const uint32_t sa = fwd.off_a[u];
            const uint32_t sb = fwd.off_b ? fwd.off_b[u] : 0;
            const unsigned long long fm = front[u];
            const uint32_t c1 = min(cend, hn);
            const uint32_t e1 = min(d, (c1 - hb) * 64u);
            for (uint32_t e0 = (c - hb) * 64u; e0 < e1; e0 += 64) {
                const uint32_t e = e0 + lane;
                bool first = false;
                uint32_t v = 0;
                if (e < e1) {
                    v = e < da ? fwd.tgt_a[sa + e] : fwd.tgt_b[sb + (e - da)];
                    first = ms_push_edge(v, fm, visit, next, levels, n, L, cnt);
                }
                const uint32_t slot = wave_append_slot(first, &ctr->qlen);
                if (first) {
                    nq[slot] = v;
                    edges += nv_degree(fwd, v);
                }
            }
            c = c1;
        }
    }
    cnt = wave_sum_u64(cnt);
    edges = wave_sum_u64(edges);
    if (lane == 0) {
        if (cnt) atomicAdd(&ctr->newbits, cnt);
        if (edges) atomicAdd(&ctr->edges, edges);
    }
}

// One level's expansion: the push over the frontier queue or the pull over all
// vertices, as the control block chose (uniform over the grid).
__global__ __launch_bounds__(MB) void k_ms_expand(uint32_t* __restrict__ q0, uint32_t* __restrict__ q1, NbrView fwd,
                                                  NbrView rev, const unsigned long long* __restrict__ front,
                                                  unsigned long long* __restrict__ next,
                                                  unsigned long long* __restrict__ visit, uint64_t n, const MsCtl* ctl,
                                                  MsCtr* ctr, uint32_t* __restrict__ heavy,
                                                  uint32_t* __restrict__ hbase) {
    if (ctl->done) return;
    if (ctl->push)
        ms_push(q0, q1, fwd, front, next, visit, n, ctl, ctr, heavy, hbase);
    else
        ms_pull(n, rev, front, next, visit, ctl, ctr);
}

// The heavy vertices of the level: the push frontier's long forward lists
// (queued by the push) or the pull side's long reverse lists, chunk by chunk.
__global__ __launch_bounds__(MB) void k_ms_expand_heavy(const uint32_t* __restrict__ heavy,
                                                        const uint32_t* __restrict__ hbase, NbrView fwd, NbrView rev,
                                                        const unsigned long long* __restrict__ front,
                                                        unsigned long long* __restrict__ next,
                                                        unsigned long long* __restrict__ visit, uint64_t n,
                                                        uint32_t* __restrict__ q0, uint32_t* __restrict__ q1,
                                                        const uint32_t* __restrict__ pheavy,
                                                        const uint32_t* __restrict__ pcbase, uint32_t pnh,
                                                        uint32_t pchunks, unsigned long long* __restrict__ hacc,
                                                        const MsCtl* ctl, MsCtr* ctr) {
    if (ctl->done) return;
    if (ctl->push)
        ms_push_heavy(heavy, hbase, fwd, front, next, visit, n, q0, q1, ctl, ctr);
    else if (pnh)
        ms_pull_heavy(pheavy, pcbase, pnh, pchunks, rev, front, visit, ctl, hacc);
}

// queue of vertices with a non-zero front mask (pull -> push hand-over).
__global__ __launch_bounds__(MB) void k_ms_compact(const unsigned long long* __restrict__ front, uint64_t n,
                                                   NbrView fwd, uint32_t* __restrict__ q0, uint32_t* __restrict__ q1,
                                                   const MsCtl* ctl, MsCtr* ctr) {
    if (ctl->done || !ctl->compact) return;
    uint32_t* q = ctl->qsel ? q1 : q0;
    __shared__ uint32_t s_cnt[MW + 1];
    unsigned long long edges = 0;
    for (uint64_t b0 = blockIdx.x * (uint64_t)MB; b0 < n; b0 += (uint64_t)gridDim.x * MB) {
        const uint64_t v = b0 + threadIdx.x;
        const bool take = v < n && front[v] != 0;
        const uint32_t slot = block_append_slot<MB>(take, &ctr->qlen, s_cnt);
        if (take) {
            q[slot] = (uint32_t)v;
            edges += nv_degree(fwd, (uint32_t)v);
        }
    }
    block_add<MB>(edges, &ctr->edges);
}

// TEPS numerator: sum over vertices of forward degree x reached lanes.
__global__ void k_ms_reached_edges(const unsigned long long* __restrict__ visit, NbrView fwd, uint64_t n,
                                   unsigned long long* total) {
    unsigned long long acc = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        acc += (unsigned long long)nv_degree(fwd, (uint32_t)v) * __popcll(visit[v]);
    block_add<MB>(acc, total);
}

}  // namespace

void direction_views(ursa_hip_topo* t, int dir, hipStream_t s, NbrView& fwd, NbrView& rev);

// Levels for up to 64 sources (row b of `levels` = source b) on one stream.
// The pull side's heavy-vertex list is cached on the topology per direction;
// the rest is per call (allocations are reused across calls by the device
// allocator), and the level groups' HIP graph is recorded once per call.
struct MsWork {
    uint64_t n;
    DevBuf<unsigned long long> visit, f0, f1, redges, hacc;
    DevBuf<uint32_t> q0, q1, dsrc, heavy, hbase;
    const uint32_t* pheavy = nullptr;
    const uint32_t* pcbase = nullptr;
    uint32_t pnh = 0, pchunks = 0;  // pull-side heavy vertices and their chunk count
    DevBuf<MsCtr> ctr;
    DevBuf<MsCtl> ctl;
    CapturedLaunches group;
    explicit MsWork(uint64_t n_)
        : n(n_), visit(n_), f0(n_), f1(n_), redges(1), q0(n_), q1(n_), dsrc(64), heavy(n_), hbase(n_), ctr(1),
          ctl(1) {}
    MsCtl read_ctl(hipStream_t s) {
        MsCtl* h = static_cast<MsCtl*>(pinned_scratch(sizeof(MsCtl)));
        HIP_OK(hipMemcpyAsync(h, ctl.get(), sizeof(MsCtl), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return *h;
    }
    // Heavy reverse-list vertices and chunk bases for the pull levels.
    void prepare_pull(ursa_hip_topo* t, int dir, const NbrView& rev, hipStream_t s) {
        {
            NoArena keep_out_of_arena;
            std::lock_guard<std::mutex> lk(t->cache_mu);
            if (!t->ms_pull_ready[dir]) {
                DevBuf<uint8_t> flag(n);
                DevBuf<uint32_t> iota(n), chunks(n);
                DevBuf<uint64_t> cnt(1);
                DevBuf<uint32_t> ph(n);
                k_ms_heavy_flags<<<blocks_for(n, 256, 1u << 16), 256, 0, s>>>(rev, n, flag.get(), iota.get(),
                                                                               chunks.get());
                check_launch();
                prim::select_flagged_u32(iota.get(), flag.get(), ph.get(), cnt.get(), n, s);
                uint64_t c = 0;
                HIP_OK(hipMemcpyAsync(&c, cnt.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
                HIP_OK(hipStreamSynchronize(s));
                const uint32_t nh = (uint32_t)c;
                DevBuf<uint32_t> cb(nh + 1);
                uint32_t nch = 0;
                if (nh) {
                    DevBuf<uint32_t> hch(nh);
                    k_ms_gather<<<blocks_for(nh, 256), 256, 0, s>>>(ph.get(), nh, chunks.get(), hch.get());
                    check_launch();
                    prim::exclusive_sum_u32(hch.get(), cb.get(), nh, s);
                    HIP_OK(hipMemcpyAsync(&nch, cb.get() + nh, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                }
                DevBuf<uint32_t> keep(nh ? nh : 1);  // the list at its own length
                if (nh) HIP_OK(hipMemcpyAsync(keep.get(), ph.get(), nh * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
                HIP_OK(hipStreamSynchronize(s));
                t->ms_pheavy[dir] = std::move(keep);
                t->ms_pcbase[dir] = std::move(cb);
                t->ms_pnh[dir] = nh;
                t->ms_pchunks[dir] = nch;
                t->ms_pull_ready[dir] = true;
            }
            pheavy = t->ms_pheavy[dir].get();
            pcbase = t->ms_pcbase[dir].get();
            pnh = t->ms_pnh[dir];
            pchunks = t->ms_pchunks[dir];
        }
        hacc.alloc(pnh ? pnh : 1);
        HIP_OK(hipMemsetAsync(hacc.get(), 0, (pnh ? pnh : 1) * sizeof(unsigned long long), s));
    }
};

// One batch of nb <= 64 sources: seed, then groups of MS_LEVELS_PER_GROUP
// levels until the control block reports an empty frontier. Every launch in a
// group has the same arguments for every group and batch of the call (the
// level, the direction and the batch's level rows come from the control
// block), so the group is one replayed graph.
void msbfs_batch(uint64_t n, uint64_t fwd_edges, const uint32_t* src_host, uint32_t nb, const NbrView& fwd,
                 const NbrView& rev, int32_t* levels, MsWork& w, hipStream_t s, uint64_t* reached_edges,
                 uint32_t push_div) {
    const unsigned gv = blocks_for(n, MB, GRID_CAP);
    k_ms_fill_levels<<<blocks_for(n * nb, 256, 1u << 16), 256, 0, s>>>(levels, n * nb);
    k_ms_zero_u64<<<gv, MB, 0, s>>>(w.visit.get(), n);
    k_ms_zero_u64<<<gv, MB, 0, s>>>(w.f0.get(), n);
    HIP_OK(hipMemcpyAsync(w.dsrc.get(), src_host, nb * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    k_ms_seed<<<1, 1, 0, s>>>(w.dsrc.get(), nb, n, fwd, w.visit.get(), w.f0.get(), levels, w.q0.get(), w.ctl.get());
    check_launch();
    MsCtl* ctl = w.ctl.get();
    MsCtr* ctr = w.ctr.get();
    auto group = [&] {
        for (int i = 0; i < MS_LEVELS_PER_GROUP; ++i) {
            // Level L = 2 g + i + 1 of any group g: its frontier is in f0 when L is odd.
            unsigned long long* front = (i & 1) ? w.f1.get() : w.f0.get();
            unsigned long long* next = (i & 1) ? w.f0.get() : w.f1.get();
            k_ms_level_step<<<1, 1, 0, s>>>(ctl, ctr, n, fwd_edges, push_div);
            k_ms_compact<<<gv, MB, 0, s>>>(front, n, fwd, w.q0.get(), w.q1.get(), ctl, ctr);
            k_ms_after_compact<<<1, 1, 0, s>>>(ctl, ctr, fwd_edges, push_div);
            k_ms_clear_next<<<gv, MB, 0, s>>>(ctl, next, n);
            timed("msbfs_expand", 0.0, s, [&] {
                k_ms_expand<<<GRID_CAP, MB, 0, s>>>(w.q0.get(), w.q1.get(), fwd, rev, front, next, w.visit.get(), n,
                                                    ctl, ctr, w.heavy.get(), w.hbase.get());
                k_ms_expand_heavy<<<GRID_CAP, MB, 0, s>>>(w.heavy.get(), w.hbase.get(), fwd, rev, front, next,
                                                          w.visit.get(), n,
                                                          w.q0.get(), w.q1.get(), w.pheavy, w.pcbase, w.pnh,
                                                          w.pchunks, w.hacc.get(), ctl, ctr);
                if (w.pnh)
                    k_ms_pull_finish<<<blocks_for(w.pnh, 256), 256, 0, s>>>(w.pheavy, w.pnh, w.hacc.get(), next,
                                                                           w.visit.get(), n, ctl, ctr);
            });
            check_launch();
        }
    };
    for (;;) {
        w.group.run(s, group);
        if (w.read_ctl(s).done) break;
    }
    if (reached_edges) {
        HIP_OK(hipMemsetAsync(w.redges.get(), 0, sizeof(unsigned long long), s));
        k_ms_reached_edges<<<gv, MB, 0, s>>>(w.visit.get(), fwd, n, w.redges.get());
        check_launch();
        unsigned long long hre = 0;
        HIP_OK(hipMemcpyAsync(&hre, w.redges.get(), sizeof(hre), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        *reached_edges += hre;
    }
}

void msbfs_levels_each(ursa_hip_topo* t, const std::vector<uint32_t>& src, int dir, int32_t* lv64, uint32_t push_div,
                       const std::function<void(uint64_t, uint32_t)>& each) {
    const uint64_t n = t->n;
    hipStream_t s = stream();
    NbrView fwd, rev;
    direction_views(t, dir, s, fwd, rev);
    MsWork w(n);
    w.prepare_pull(t, dir, rev, s);
    const uint64_t fwd_edges = dir == URSA_HIP_BOTH ? 2 * t->m : t->m;
    for (uint64_t b0 = 0; b0 < src.size(); b0 += 64) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(64, src.size() - b0);
        msbfs_batch(n, fwd_edges, src.data() + b0, nb, fwd, rev, lv64, w, s, nullptr, push_div);
        each(b0, nb);
    }
}

void msbfs_levels(ursa_hip_topo* t, const std::vector<uint32_t>& src, int dir, int32_t* out_dev,
                  uint64_t* reached_edges, uint32_t push_div) {
    const uint64_t n = t->n;
    hipStream_t s = stream();
    NbrView fwd, rev;
    direction_views(t, dir, s, fwd, rev);
    MsWork w(n);
    w.prepare_pull(t, dir, rev, s);
    const uint64_t fwd_edges = dir == URSA_HIP_BOTH ? 2 * t->m : t->m;
    for (uint64_t b0 = 0; b0 < src.size(); b0 += 64) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(64, src.size() - b0);
        msbfs_batch(n, fwd_edges, src.data() + b0, nb, fwd, rev, out_dev + b0 * n, w, s, reached_edges, push_div);
    }
}

}  // namespace ursa
