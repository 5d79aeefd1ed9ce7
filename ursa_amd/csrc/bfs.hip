// bfs.hip — direction-optimizing BFS hop levels (bfs_distances,
// ursa-core/src/algo/bfs.rs:18-42).
//
// Level-synchronous, switching per level between
//   top-down  (k_bfs_td): the frontier is a queue; each wave takes 64 queue
//             vertices, prefix-sums their degrees and walks the combined edge
//             list 64 edges at a time (owner found by a 6-step shuffle search),
//             claims unvisited targets with a CAS on level[] and appends them
//             to the next queue with one atomic per wave;
//   bottom-up (k_bfs_bu): the frontier is a bitmap; each unvisited vertex scans
//             its reverse neighbours and stops at the first one in the
//             frontier; each wave owns 64 consecutive vertices, so it writes its
//             two next-frontier bitmap words whole (ballot) with no atomics.
// The switch follows Beamer et al.: go bottom-up when the frontier's edge count
// exceeds the unexplored edges / 15, go back top-down once the frontier shrinks
// below n / 18. Hop levels are unique, so every schedule gives the same i32
// output as the reference's serial frontier loop.
//
// Directions (topology.rs:395-404): Out expands out-edges and checks
// in-neighbours bottom-up; In the reverse; Both concatenates out and in for both.
//
// Algorithmic bytes per source: 4 B per inspected edge (E_reach = sum of the
// forward degrees of reached vertices) + 12 B per node (offsets, level write,
// level/visited read).
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "frontier.hpp"
#include "rng.hpp"

namespace ursa {

namespace {

constexpr int BFS_BLOCK = 256;
constexpr int WAVES = BFS_BLOCK / 64;

struct BfsCtr {
    uint32_t qlen;  // next-queue length (top-down) / newly visited count (bottom-up)
    uint32_t pad;
    // Frontier vertices deferred to the grid-wide expansion (high 32 bits) and
    // the 64-slot chunks of their lists (low 32 bits), one atomic add each.
    unsigned long long hpack;
    unsigned long long edges;  // forward degree sum of the newly visited vertices
};

// Frontier vertices with more forward edges than this are expanded by the
// whole grid (k_bfs_td_heavy) instead of by one wave.
constexpr uint32_t TD_HEAVY = 1024;

// The level counters are double-buffered: the first kernel of each step
// clears the slot the next step will use, so no launch is spent on resets.
__device__ __forceinline__ void ctr_clear(BfsCtr* c) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->qlen = 0;
        c->hpack = 0;
        c->edges = 0;
    }
}

__global__ void k_bfs_init(int32_t* level, uint64_t n, uint32_t src, uint32_t* queue, BfsCtr* ctr, BfsCtr* next_ctr,
                           NbrView fwd) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        level[i] = (i == src) ? 0 : -1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        queue[0] = src;
        ctr->qlen = 1;
        ctr->hpack = 0;
        ctr->edges = nv_degree(fwd, src);
    }
    ctr_clear(next_ctr);
}

// Claim v for level L and append it to the next queue (wave-aggregated).
__device__ __forceinline__ void td_visit(bool active, uint32_t v, const NbrView& fwd, int32_t* level, int32_t L,
                                         uint32_t* nq, BfsCtr* ctr, uint64_t& edges) {
    bool claimed = false;
    if (active && level[v] < 0) claimed = atomicCAS(&level[v], -1, L) == -1;
    const uint32_t slot = wave_append_slot(claimed, &ctr->qlen);
    if (claimed) {
        nq[slot] = v;
        edges += nv_degree(fwd, v);
    }
}

__global__ __launch_bounds__(BFS_BLOCK) void k_bfs_td(const uint32_t* __restrict__ q, uint32_t qlen, NbrView fwd,
                                                      int32_t* __restrict__ level, int32_t L,
                                                      uint32_t* __restrict__ nq, BfsCtr* ctr,
                                                      uint32_t* __restrict__ heavy, uint32_t* __restrict__ hbase,
                                                      BfsCtr* next_ctr) {
    ctr_clear(next_ctr);
    const int lane = lane_id();
    const uint64_t gw = blockIdx.x * (uint64_t)WAVES + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    uint64_t edges = 0;
    for (uint64_t base = gw * 64; base < qlen; base += nwaves * 64) {
        const uint64_t idx = base + lane;
        uint32_t u = 0, d = 0, da = 0, sa = 0, sb = 0;
        if (idx < qlen) {
            u = q[idx];
            d = nv_degree(fwd, u, da);
            sa = fwd.off_a[u];
            if (fwd.off_b) sb = fwd.off_b[u];
            if (d > TD_HEAVY) {
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
            uint32_t v = 0;
            if (t < total) v = e < oa ? fwd.tgt_a[osa + e] : fwd.tgt_b[osb + (e - oa)];
            td_visit(t < total, v, fwd, level, L, nq, ctr, edges);
        }
    }
    edges = wave_sum_u64(edges);
    if (lane == 0 && edges) atomicAdd(&ctr->edges, edges);
}

// Heavy frontier vertices. Their 64-slot chunks are numbered in one sequence
// (entry h owns chunks hbase[h] .. hbase[h+1]-1, the last up to the total) and
// every wave takes a contiguous run of it, so hub lists spread over the grid
// and no wave walks the deferred list (which took milliseconds when a level
// held thousands of hubs).
__global__ __launch_bounds__(BFS_BLOCK) void k_bfs_td_heavy(const uint32_t* __restrict__ heavy,
                                                            const uint32_t* __restrict__ hbase, NbrView fwd,
                                                            int32_t* __restrict__ level, int32_t L,
                                                            uint32_t* __restrict__ nq, BfsCtr* ctr) {
    const unsigned long long hp = ctr->hpack;
    const uint32_t cnt = (uint32_t)(hp >> 32), total = (uint32_t)hp;
    if (cnt == 0) return;
    const int lane = lane_id();
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    const uint64_t gw = blockIdx.x * (uint64_t)WAVES + (threadIdx.x >> 6);
    const uint64_t per = (total + nwaves - 1) / nwaves;
    uint32_t c = (uint32_t)min((uint64_t)total, gw * per);
    const uint32_t cend = (uint32_t)min((uint64_t)total, gw * per + per);
    uint64_t edges = 0;
    if (c < cend) {
        uint32_t lo = 0, hi = cnt;  // largest h with hbase[h] <= c
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (hbase[mid] <= c)
                lo = mid;
            else
                hi = mid;
        }
        for (uint32_t h = lo; c < cend; ++h) {
            const uint32_t u = heavy[h], hb = hbase[h];
            const uint32_t hn = h + 1 < cnt ? hbase[h + 1] : total;
            uint32_t da;
            const uint32_t d = nv_degree(fwd, u, da);
            const uint32_t sa = fwd.off_a[u];
            const uint32_t sb = fwd.off_b ? fwd.off_b[u] : 0;
            const uint32_t c1 = min(cend, hn);
            const uint32_t e1 = min(d, (c1 - hb) * 64u);
            for (uint32_t e0 = (c - hb) * 64u; e0 < e1; e0 += 64) {
                const uint32_t e = e0 + lane;
                uint32_t v = 0;
                if (e < e1) v = e < da ? fwd.tgt_a[sa + e] : fwd.tgt_b[sb + (e - da)];
                td_visit(e < e1, v, fwd, level, L, nq, ctr, edges);
            }
            c = c1;
        }
    }
    edges = wave_sum_u64(edges);
    if (lane == 0 && edges) atomicAdd(&ctr->edges, edges);
}

// Bottom-up step: vertices of level L are those unvisited with a reverse
// neighbour in `front` (level L-1). Writes the whole next bitmap.
// A lane checks its own vertex's reverse list eight, then four slots at a time
// (the slot reads, then the bitmap reads, are independent). Vertices with more
// than BU_HEAVY reverse slots check their first four slots in their own lane
// like the others; if no parent is among them, they would hold their wave for
// the whole list, so the wave takes them together afterwards: 64 slots per
// step, coalesced, with the early exit decided by a ballot.
constexpr uint32_t BU_HEAVY = 64;
constexpr int BU_STEPS = 4;  // s24, 4 sources: 1 step 3.80 ms, 4 steps 3.53 ms, 8 steps 3.78 ms

__device__ __forceinline__ bool bu_scan_lane(const uint32_t* __restrict__ tgt, uint32_t k, uint32_t ke,
                                             const uint32_t* __restrict__ front) {
    for (; k + 8 <= ke; k += 8) {
        uint32_t wv[8], bv = 0;
#pragma unroll
        for (int r = 0; r < 8; ++r) wv[r] = tgt[k + r];
#pragma unroll
        for (int r = 0; r < 8; ++r) bv |= front[wv[r] >> 5] >> (wv[r] & 31);
        if (bv & 1u) return true;
    }
    for (; k + 4 <= ke; k += 4) {
        const uint32_t w0 = tgt[k], w1 = tgt[k + 1], w2 = tgt[k + 2], w3 = tgt[k + 3];
        const uint32_t b0 = front[w0 >> 5] >> (w0 & 31), b1 = front[w1 >> 5] >> (w1 & 31);
        const uint32_t b2 = front[w2 >> 5] >> (w2 & 31), b3 = front[w3 >> 5] >> (w3 & 31);
        if ((b0 | b1 | b2 | b3) & 1u) return true;
    }
    for (; k < ke; ++k) {
        const uint32_t w = tgt[k];
        if ((front[w >> 5] >> (w & 31)) & 1u) return true;
    }
    return false;
}

// A heavy vertex's reverse list checked by the whole wave, BU_STEPS 64-slot
// steps (their slot loads, then their bitmap loads) in flight per check.
__device__ __forceinline__ bool bu_scan_wave(const uint32_t* __restrict__ tgt, uint32_t k, uint32_t ke,
                                             const uint32_t* __restrict__ front) {
    const int lane = lane_id();
    for (; k < ke; k += 64 * BU_STEPS) {  // k, ke wave-uniform
        uint32_t w[BU_STEPS];
#pragma unroll
        for (int r = 0; r < BU_STEPS; ++r) {
            const uint32_t i = k + 64 * r + lane;
            w[r] = i < ke ? tgt[i] : 0xFFFFFFFFu;
        }
        uint32_t bv = 0;
#pragma unroll
        for (int r = 0; r < BU_STEPS; ++r)
            if (w[r] != 0xFFFFFFFFu) bv |= front[w[r] >> 5] >> (w[r] & 31);
        if (__ballot(bv & 1u)) return true;
    }
    return false;
}

__global__ __launch_bounds__(BFS_BLOCK) void k_bfs_bu(uint64_t n, NbrView rev, NbrView fwd,
                                                      const uint32_t* __restrict__ front,
                                                      uint32_t* __restrict__ next, int32_t* __restrict__ level,
                                                      int32_t L, BfsCtr* ctr, BfsCtr* next_ctr) {
    ctr_clear(next_ctr);
    const int lane = lane_id();
    const uint64_t gw = blockIdx.x * (uint64_t)WAVES + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    uint64_t edges = 0;
    uint32_t awake = 0;
    for (uint64_t base = gw * 64; base < n; base += nwaves * 64) {
        const uint64_t v = base + lane;
        bool found = false, heavy = false;
        if (v < n && level[v] < 0) {
            const uint32_t ka = rev.off_a[v], kea = rev.off_a[v + 1];
            uint32_t kb = 0, keb = 0;
            if (rev.off_b) {
                kb = rev.off_b[v];
                keb = rev.off_b[v + 1];
            }
            if ((kea - ka) + (keb - kb) > BU_HEAVY) {
                // the first four slots of the first list, in this lane
                found = bu_scan_lane(rev.tgt_a, ka, min(ka + 4, kea), front);
                heavy = !found;
            } else {
                found = bu_scan_lane(rev.tgt_a, ka, kea, front);
                if (!found && rev.off_b) found = bu_scan_lane(rev.tgt_b, kb, keb, front);
            }
        }
        uint64_t hm = __ballot(heavy);
        while (hm) {
            const int hl = __builtin_ctzll(hm);
            hm &= hm - 1;
            const uint32_t hv = (uint32_t)(base + hl);
            const uint32_t ha = rev.off_a[hv], hea = rev.off_a[hv + 1];
            bool hf = bu_scan_wave(rev.tgt_a, min(ha + 4, hea), hea, front);
            if (!hf && rev.off_b) hf = bu_scan_wave(rev.tgt_b, rev.off_b[hv], rev.off_b[hv + 1], front);
            if (lane == hl) found = hf;
        }
        if (found) {
            level[v] = L;
            edges += nv_degree(fwd, (uint32_t)v);
        }
        const uint64_t mask = __ballot(found);
        if (lane == 0) {
            next[base >> 5] = (uint32_t)mask;
            next[(base >> 5) + 1] = (uint32_t)(mask >> 32);
            awake += (uint32_t)__popcll(mask);
        }
    }
    block_add<BFS_BLOCK>((unsigned long long)edges, &ctr->edges);
    block_add<BFS_BLOCK>(awake, &ctr->qlen);
}

__global__ __launch_bounds__(BFS_BLOCK) void k_level_to_bits(const int32_t* __restrict__ level, uint64_t n, int32_t L,
                                                             uint32_t* __restrict__ bits) {
    const int lane = lane_id();
    const uint64_t gw = blockIdx.x * (uint64_t)WAVES + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    for (uint64_t base = gw * 64; base < n; base += nwaves * 64) {
        const uint64_t v = base + lane;
        const uint64_t mask = __ballot(v < n && level[v] == L);
        if (lane == 0) {
            bits[base >> 5] = (uint32_t)mask;
            bits[(base >> 5) + 1] = (uint32_t)(mask >> 32);
        }
    }
}

// Queue of the vertices whose bit is set in a frontier bitmap (the last
// bottom-up level), with one atomic per workgroup: each workgroup counts the
// set bits of its contiguous range of words, reserves that many slots, then
// writes the vertices in the same word order. (Appending per 256 vertices of a
// grid-stride loop had cost 35 K same-address atomics, 0.28 ms at scale 24.)
__global__ __launch_bounds__(BFS_BLOCK) void k_bits_to_queue(const uint32_t* __restrict__ bits, uint64_t nwords,
                                                             uint32_t* __restrict__ q, BfsCtr* ctr,
                                                             BfsCtr* next_ctr) {
    __shared__ uint32_t s_w[WAVES];
    __shared__ uint32_t s_base;
    ctr_clear(next_ctr);
    const uint64_t per = (nwords + gridDim.x - 1) / gridDim.x;
    const uint64_t w0 = blockIdx.x * per;
    const uint64_t w1 = w0 + per < nwords ? w0 + per : nwords;
    uint32_t cnt = 0;
    for (uint64_t w = w0 + threadIdx.x; w < w1; w += BFS_BLOCK) cnt += __popc(bits[w]);
    const uint32_t incl = wave_incl_scan(cnt);
    const int wid = (int)(threadIdx.x >> 6);
    if (lane_id() == 63) s_w[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < WAVES; ++i) {
            const uint32_t c = s_w[i];
            s_w[i] = tot;
            tot += c;
        }
        s_base = tot ? atomicAdd(&ctr->qlen, tot) : 0u;
    }
    __syncthreads();
    uint32_t slot = s_base + s_w[wid] + incl - cnt;
    for (uint64_t w = w0 + threadIdx.x; w < w1; w += BFS_BLOCK) {
        uint32_t b = bits[w];
        while (b) {
            const int k = __ffs(b) - 1;
            b &= b - 1;
            q[slot++] = (uint32_t)(w * 32 + (uint64_t)k);
        }
    }
}

__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

// Direction views of a topology (forward = expansion direction, reverse = the
// direction bottom-up checks).
void direction_views(ursa_hip_topo* t, int dir, hipStream_t s, NbrView& fwd, NbrView& rev) {
    const DevCsr& out = t->out;
    const DevCsr& in = t->incoming(s);
    const NbrView vo{out.off.get(), out.tgt.get(), nullptr, nullptr};
    const NbrView vi{in.off.get(), in.tgt.get(), nullptr, nullptr};
    const NbrView vb{out.off.get(), out.tgt.get(), in.off.get(), in.tgt.get()};
    if (dir == URSA_HIP_OUT) {
        fwd = vo;
        rev = vi;
    } else if (dir == URSA_HIP_IN) {
        fwd = vi;
        rev = vo;
    } else {
        fwd = vb;
        rev = vb;
    }
}

struct BfsWork {
    DevBuf<uint32_t> q0, q1, b0, b1, heavy, hbase;
    DevBuf<BfsCtr> ctr;  // two slots, used by alternate steps
    BfsCtr* h = nullptr;  // pinned mirror of the current slot
    explicit BfsWork(uint64_t n)
        : q0(n ? n : 1), q1(n ? n : 1), b0(((n + 63) / 64) * 2), b1(((n + 63) / 64) * 2), heavy(n ? n : 1),
          hbase(n ? n : 1), ctr(2) {
        HIP_OK(hipHostMalloc((void**)&h, sizeof(BfsCtr), hipHostMallocDefault));
    }
    ~BfsWork() {
        if (h) (void)hipHostFree(h);
    }
};

// One direction-optimizing BFS into level (device, n entries). Returns the
// number of levels run; *reached_edges = sum of forward degrees of reached
// vertices (the TEPS numerator).
static void bfs_one(ursa_hip_topo* t, uint32_t src, const NbrView& fwd, const NbrView& rev, uint64_t fwd_edges,
                    int32_t* level, BfsWork& w, hipStream_t s, uint64_t* reached_edges, double alpha, double beta) {
    const uint64_t n = t->n;
    const unsigned gb = blocks_for((n + 63) / 64, WAVES, GRID_CAP);
    const uint64_t nwords = ((n + 63) / 64) * 2;
    // Step k uses counter slot k & 1 and clears slot (k + 1) & 1.
    int step = 0;
    auto cur = [&]() { return w.ctr.get() + (step & 1); };
    auto nxt = [&]() { return w.ctr.get() + ((step + 1) & 1); };
    auto read_ctr = [&]() {
        HIP_OK(hipMemcpyAsync(w.h, cur(), sizeof(BfsCtr), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        ++step;
        return *w.h;
    };
    k_bfs_init<<<blocks_for(n, 256), 256, 0, s>>>(level, n, src, w.q0.get(), cur(), nxt(), fwd);
    check_launch();
    BfsCtr c = read_ctr();
    uint32_t* q = w.q0.get();
    uint32_t* nq = w.q1.get();
    uint32_t* front = w.b0.get();
    uint32_t* next = w.b1.get();
    uint64_t qlen = c.qlen;
    uint64_t scout = c.edges;
    uint64_t total_edges = scout;
    int64_t edges_to_check = (int64_t)fwd_edges;
    int32_t L = 1;
    const double bytes_node = 12.0 * (double)n;
    while (qlen > 0) {
        if ((double)scout > (double)edges_to_check / alpha) {
            // bottom-up phase
            k_level_to_bits<<<gb, BFS_BLOCK, 0, s>>>(level, n, L - 1, front);
            check_launch();
            uint64_t awake = qlen, old_awake;
            do {
                old_awake = awake;
                timed("bfs_bottom_up", bytes_node, s, [&] {
                    k_bfs_bu<<<gb, BFS_BLOCK, 0, s>>>(n, rev, fwd, front, next, level, L, cur(), nxt());
                });
                check_launch();
                c = read_ctr();
                awake = c.qlen;
                total_edges += c.edges;
                std::swap(front, next);
                ++L;
            } while (awake > 0 && (awake >= old_awake || (double)awake > (double)n / beta));
            if (awake == 0) break;
            k_bits_to_queue<<<gb, BFS_BLOCK, 0, s>>>(front, nwords, q, cur(), nxt());
            check_launch();
            c = read_ctr();
            qlen = c.qlen;
            scout = 1;
        } else {
            edges_to_check -= (int64_t)scout;
            const unsigned tb = blocks_for((qlen + 63) / 64, WAVES, 1u << 16);
            timed("bfs_top_down", 4.0 * (double)scout + 8.0 * (double)qlen, s, [&] {
                k_bfs_td<<<tb, BFS_BLOCK, 0, s>>>(q, (uint32_t)qlen, fwd, level, L, nq, cur(), w.heavy.get(),
                                                  w.hbase.get(), nxt());
                if (scout > TD_HEAVY) {
                    const unsigned hb = blocks_for(scout, BFS_BLOCK, 2048);
                    k_bfs_td_heavy<<<hb, BFS_BLOCK, 0, s>>>(w.heavy.get(), w.hbase.get(), fwd, level, L, nq, cur());
                }
            });
            check_launch();
            c = read_ctr();
            qlen = c.qlen;
            scout = c.edges;
            total_edges += c.edges;
            std::swap(q, nq);
            ++L;
        }
    }
    if (reached_edges) *reached_edges = total_edges;
}

void msbfs_levels(ursa_hip_topo* t, const std::vector<uint32_t>& src, int dir, int32_t* out_dev,
                  uint64_t* reached_edges, uint32_t push_div);

namespace {

std::mutex g_tune_mu;
std::map<uint64_t, TuneParams> g_tune_cache;  // degree-distribution key -> parameters

// Key of the forward degree distribution: log2 n, 2 log2 of the mean degree,
// log2 of max/mean degree, eighths of zero-degree vertices, and the direction.
uint64_t degree_key(ursa_hip_topo* t, int dir, const NbrView& fwd, hipStream_t s) {
    if (t->tune_key[dir]) return t->tune_key[dir];
    const uint64_t n = t->n;
    std::vector<uint32_t> a(n + 1), b;
    HIP_OK(hipMemcpyAsync(a.data(), fwd.off_a, (n + 1) * 4, hipMemcpyDeviceToHost, s));
    if (fwd.off_b) {
        b.resize(n + 1);
        HIP_OK(hipMemcpyAsync(b.data(), fwd.off_b, (n + 1) * 4, hipMemcpyDeviceToHost, s));
    }
    HIP_OK(hipStreamSynchronize(s));
    uint64_t zero = 0, maxd = 0, tot = 0;
    for (uint64_t u = 0; u < n; ++u) {
        uint64_t d = a[u + 1] - a[u];
        if (!b.empty()) d += b[u + 1] - b[u];
        zero += d == 0;
        maxd = std::max(maxd, d);
        tot += d;
    }
    const double mean = n ? (double)tot / (double)n : 0.0;
    auto lg = [](double x) { return x > 1.0 ? (uint64_t)std::lround(std::log2(x)) : 0ull; };
    const uint64_t k = (lg((double)n) & 0xFF) | ((lg(mean * mean) & 0xFF) << 8) |
                       ((lg(mean > 0 ? (double)maxd / mean : 0.0) & 0xFF) << 16) |
                       (((uint64_t)(n ? 8 * zero / n : 0) & 0xF) << 24) | ((uint64_t)dir << 28) | (1ull << 32);
    t->tune_key[dir] = k;
    return k;
}

}  // namespace

// The parameters a traversal in direction dir uses: measured on this graph,
// else cached for its degree-distribution key, else the defaults.
TuneParams effective_tune(ursa_hip_topo* t, int dir) {
    if (t->tune[dir].source != 0 || t->tune_key[dir] == 0) return t->tune[dir];
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(t->tune_key[dir]);
    if (it == g_tune_cache.end()) return t->tune[dir];
    TuneParams p = it->second;
    p.source = 2;
    return p;
}

// mode: 0 = automatic (64-lane multi-source BFS for 4 or more sources),
// 1 = one direction-optimizing traversal per source, 2 = multi-source.
void bfs_levels(ursa_hip_topo* t, const uint32_t* sources, uint64_t k, int dir, int32_t* out_user,
                uint64_t* reached_edges, int mode) {
    const uint64_t n = t->n;
    if (reached_edges) *reached_edges = 0;
    if (n == 0 || k == 0) return;
    hipStream_t s = stream();
    NbrView fwd, rev;
    direction_views(t, dir, s, fwd, rev);
    const uint64_t fwd_edges = dir == URSA_HIP_BOTH ? 2 * t->m : t->m;
    DevOut<int32_t> out(out_user, n * k);
    BfsWork w(n);
    std::vector<uint32_t> src(k);
    if (is_device_ptr(sources))
        HIP_OK(hipMemcpy(src.data(), sources, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
    else
        std::copy(sources, sources + k, src.begin());
    const bool multi = mode == 2 || (mode == 0 && k >= 4);
    const TuneParams tp = effective_tune(t, dir);
    if (multi) {
        uint64_t re = 0;
        msbfs_levels(t, src, dir, out.p, reached_edges ? &re : nullptr, tp.ms_push_div);
        out.finish(s);
        if (reached_edges) *reached_edges = re;
        return;
    }
    uint64_t sum = 0;
    for (uint64_t i = 0; i < k; ++i) {
        int32_t* lv = out.p + i * n;
        if (src[i] >= n) {  // bfs.rs:21-23: out-of-range source -> all -1
            k_fill_i32<<<blocks_for(n, 256), 256, 0, s>>>(lv, n, -1);
            check_launch();
            continue;
        }
        uint64_t re = 0;
        bfs_one(t, src[i], fwd, rev, fwd_edges, lv, w, s, &re, tp.bfs_alpha, tp.bfs_beta);
        sum += re;
    }
    out.finish(s);
    if (reached_edges) *reached_edges = sum;
}

// Empirical schedule search on this graph: the direction-switch pair (alpha,
// beta) over a 3 x 3 grid, timed on two sets of 4 single-source traversals,
// and the multi-source push rule over 3 divisors, timed on two 64-source
// batches. The
// winner is kept on the topology and cached for its degree-distribution key,
// so a graph with the same key reuses it without timing.
constexpr uint64_t kTuneSeed = 0x74756E65ull;  // sources the autotuner times on

TuneParams autotune(ursa_hip_topo* t, int dir, bool force) {
    const uint64_t n = t->n;
    hipStream_t s = stream();
    NbrView fwd, rev;
    direction_views(t, dir, s, fwd, rev);
    const uint64_t key = degree_key(t, dir, fwd, s);
    if (!force) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            t->tune[dir] = it->second;
            t->tune[dir].source = 2;
            return t->tune[dir];
        }
    }
    TuneParams best;
    if (n == 0 || t->m == 0) return best;
    const uint64_t fwd_edges = dir == URSA_HIP_BOTH ? 2 * t->m : t->m;
    // Sources: 128 vertices with forward edges in a seeded random order (8
    // single-source traversals, two 64-source batches). (An id
    // stride picks sources that time alike under every candidate: at R-MAT
    // scale 22 push divisors 5, 20 and 80 all measured 3.0-3.2 ms on them,
    // while on random sources 5 took 5.1 ms and the others 3.1 ms.)
    std::vector<uint32_t> a(n + 1), b;
    HIP_OK(hipMemcpyAsync(a.data(), fwd.off_a, (n + 1) * 4, hipMemcpyDeviceToHost, s));
    if (fwd.off_b) {
        b.resize(n + 1);
        HIP_OK(hipMemcpyAsync(b.data(), fwd.off_b, (n + 1) * 4, hipMemcpyDeviceToHost, s));
    }
    HIP_OK(hipStreamSynchronize(s));
    std::vector<uint32_t> live;
    for (uint64_t u = 0; u < n; ++u)
        if (a[u + 1] > a[u] || (!b.empty() && b[u + 1] > b[u])) live.push_back((uint32_t)u);
    std::vector<uint32_t> cand;
    const std::vector<uint32_t> order = shuffled_order(live.size(), kTuneSeed);
    for (uint64_t i = 0; i < 128 && i < live.size(); ++i) cand.push_back(live[order[i]]);
    if (cand.empty()) return best;
    DevBuf<int32_t> lv(n * 64);
    BfsWork w(n);
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    auto time_it = [&](auto&& body) {
        HIP_OK(hipEventRecord(e0, s));
        body();
        HIP_OK(hipEventRecord(e1, s));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        return ms;
    };
    // A candidate replaces the defaults (first in each list) only when it is
    // more than 5 % faster on each half of the sample separately, so that a
    // choice which helps one set of sources and hurts another stays out.
    const double alphas[3] = {15.0, 4.0, 60.0}, betas[3] = {18.0, 6.0, 54.0};
    const size_t ns = std::min<size_t>(8, cand.size()), nh = (ns + 1) / 2;
    float best_a = 1e30f, best_b = 1e30f;
    bfs_one(t, cand[0], fwd, rev, fwd_edges, lv.get(), w, s, nullptr, 15.0, 18.0);  // warm-up
    for (double al : alphas)
        for (double be : betas) {
            const float ta = time_it([&] {
                for (size_t i = 0; i < nh; ++i)
                    bfs_one(t, cand[i], fwd, rev, fwd_edges, lv.get() + i * n, w, s, nullptr, al, be);
            });
            const float tb = time_it([&] {
                for (size_t i = nh; i < ns; ++i)
                    bfs_one(t, cand[i], fwd, rev, fwd_edges, lv.get() + i * n, w, s, nullptr, al, be);
            });
            if (best_a == 1e30f || (ta < 0.95f * best_a && tb <= 0.95f * best_b)) {
                best_a = ta;
                best_b = tb;
                best.bfs_alpha = al;
                best.bfs_beta = be;
            }
        }
    if (cand.size() >= 4) {
        const uint32_t divs[3] = {20, 5, 80};
        const std::vector<uint32_t> b0(cand.begin(), cand.begin() + std::min<size_t>(64, cand.size()));
        const std::vector<uint32_t> b1(cand.begin() + b0.size(), cand.end());
        float best0 = 1e30f, best1 = 1e30f;
        msbfs_levels(t, b0, dir, lv.get(), nullptr, 20);  // warm-up
        for (uint32_t dv : divs) {
            const float t0 = time_it([&] { msbfs_levels(t, b0, dir, lv.get(), nullptr, dv); });
            const float t1 = b1.empty() ? 0.f : time_it([&] { msbfs_levels(t, b1, dir, lv.get(), nullptr, dv); });
            if (best0 == 1e30f || (t0 < 0.95f * best0 && t1 <= 0.95f * best1)) {
                best0 = t0;
                best1 = t1;
                best.ms_push_div = dv;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    best.source = 1;
    t->tune[dir] = best;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
    return best;
}

}  // namespace ursa

extern "C" {

int ursa_hip_bfs(ursa_hip_topo* t, const uint32_t* sources, uint32_t k, int direction, int32_t* levels) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(direction >= 0 && direction <= 2, URSA_HIP_EINVAL, "bad direction");
        if (t->n == 0 || k == 0) return;
        URSA_REQUIRE(sources != nullptr && levels != nullptr, URSA_HIP_EINVAL, "sources/levels is NULL");
        ursa::bfs_levels(t, sources, k, direction, levels, nullptr, 0);
    });
}

int ursa_hip_autotune(ursa_hip_topo* t, int direction, int force, double* alpha, double* beta, uint32_t* push_div,
                      int* source) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(direction >= 0 && direction <= 2, URSA_HIP_EINVAL, "bad direction");
        const ursa::TuneParams p = ursa::autotune(t, direction, force != 0);
        if (alpha) *alpha = p.bfs_alpha;
        if (beta) *beta = p.bfs_beta;
        if (push_div) *push_div = p.ms_push_div;
        if (source) *source = p.source;
    });
}

int ursa_hip_bfs_ex(ursa_hip_topo* t, const uint32_t* sources, uint32_t k, int direction, int32_t* levels,
                    uint64_t* reached_edges, int mode) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(direction >= 0 && direction <= 2, URSA_HIP_EINVAL, "bad direction");
        if (reached_edges) *reached_edges = 0;
        if (t->n == 0 || k == 0) return;
        URSA_REQUIRE(sources != nullptr && levels != nullptr, URSA_HIP_EINVAL, "sources/levels is NULL");
        URSA_REQUIRE(mode >= 0 && mode <= 2, URSA_HIP_EINVAL, "mode must be 0, 1 or 2");
        ursa::bfs_levels(t, sources, k, direction, levels, reached_edges, mode);
    });
}

}  // extern "C"
