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
// Algorithmic bytes counted per source as in single-source BFS (4 B per
// reached edge + 12 B per node); the batch shares them.
#include <algorithm>
#include <vector>

#include "frontier.hpp"

namespace ursa {

namespace {

constexpr int MB = 256;
constexpr int MW = MB / 64;

struct MsCtr {
    uint32_t qlen;
    uint32_t pad;
    unsigned long long newbits;  // (source, vertex) pairs settled this level
};

__device__ __forceinline__ void write_levels(uint64_t bits, int32_t* levels, uint64_t n, uint64_t v, int32_t L) {
    while (bits) {
        const int b = __ffsll((unsigned long long)bits) - 1;
        levels[(uint64_t)b * n + v] = L;
        bits &= bits - 1;
    }
}

__global__ void k_ms_fill_levels(int32_t* levels, uint64_t total) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x)
        levels[i] = -1;
}

// Seed: visit/front bits of each source, level 0, unique source queue.
__global__ void k_ms_seed(const uint32_t* src, uint32_t nb, uint64_t n, unsigned long long* visit,
                          unsigned long long* front, int32_t* levels, uint32_t* queue, MsCtr* ctr) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t q = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t s = src[b];
        if (s >= n) continue;
        if (front[s] == 0) queue[q++] = s;
        front[s] |= 1ull << b;
        visit[s] |= 1ull << b;
        levels[(uint64_t)b * n + s] = 0;
    }
    ctr->qlen = q;
}

__global__ __launch_bounds__(MB) void k_ms_pull(uint64_t n, NbrView rev, const unsigned long long* __restrict__ front,
                                                unsigned long long* __restrict__ next,
                                                unsigned long long* __restrict__ visit, unsigned long long valid,
                                                int32_t* __restrict__ levels, int32_t L, MsCtr* ctr) {
    unsigned long long cnt = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long vis = visit[v];
        const unsigned long long want = valid & ~vis;
        unsigned long long acc = 0;
        if (want) {
            for (uint32_t k = rev.off_a[v], ke = rev.off_a[v + 1]; k < ke; ++k) {
                acc |= front[rev.tgt_a[k]];
                if ((acc & want) == want) break;
            }
            if (rev.off_b && (acc & want) != want) {
                for (uint32_t k = rev.off_b[v], ke = rev.off_b[v + 1]; k < ke; ++k) {
                    acc |= front[rev.tgt_b[k]];
                    if ((acc & want) == want) break;
                }
            }
            acc &= want;
        }
        next[v] = acc;
        if (acc) {
            visit[v] = vis | acc;
            write_levels(acc, levels, n, v, L);
            cnt += __popcll(acc);
        }
    }
    cnt = wave_sum_u64(cnt);
    if (lane_id() == 0 && cnt) atomicAdd(&ctr->newbits, cnt);
}

__global__ __launch_bounds__(MB) void k_ms_push(const uint32_t* __restrict__ q, uint32_t qlen, NbrView fwd,
                                                const unsigned long long* __restrict__ front,
                                                unsigned long long* __restrict__ next,
                                                unsigned long long* __restrict__ visit, int32_t* __restrict__ levels,
                                                uint64_t n, int32_t L, uint32_t* __restrict__ nq, MsCtr* ctr) {
    const int lane = lane_id();
    const uint64_t gw = blockIdx.x * (uint64_t)MW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * MW;
    unsigned long long cnt = 0;
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
                const unsigned long long bits = ofm & ~visit[v];
                if (bits) {
                    const unsigned long long old = atomicOr(&visit[v], bits);
                    const unsigned long long mine = bits & ~old;
                    if (mine) {
                        write_levels(mine, levels, n, v, L);
                        cnt += __popcll(mine);
                        first = atomicOr(&next[v], mine) == 0;
                    }
                }
            }
            const uint32_t slot = wave_append_slot(first, &ctr->qlen);
            if (first) nq[slot] = v;
        }
    }
    cnt = wave_sum_u64(cnt);
    if (lane == 0 && cnt) atomicAdd(&ctr->newbits, cnt);
}

// queue of vertices with a non-zero front mask (pull -> push hand-over).
__global__ __launch_bounds__(MB) void k_ms_compact(const unsigned long long* __restrict__ front, uint64_t n,
                                                   uint32_t* __restrict__ q, MsCtr* ctr) {
    const uint64_t gw = blockIdx.x * (uint64_t)MW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * MW;
    for (uint64_t base = gw * 64; base < n; base += nwaves * 64) {
        const uint64_t v = base + lane_id();
        const bool take = v < n && front[v] != 0;
        const uint32_t slot = wave_append_slot(take, &ctr->qlen);
        if (take) q[slot] = (uint32_t)v;
    }
}

// TEPS numerator: sum over vertices of forward degree x reached lanes.
__global__ void k_ms_reached_edges(const unsigned long long* __restrict__ visit, NbrView fwd, uint64_t n,
                                   unsigned long long* total) {
    unsigned long long acc = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        acc += (unsigned long long)nv_degree(fwd, (uint32_t)v) * __popcll(visit[v]);
    acc = wave_sum_u64(acc);
    if (lane_id() == 0 && acc) atomicAdd(total, acc);
}

}  // namespace

void direction_views(ursa_hip_topo* t, int dir, hipStream_t s, NbrView& fwd, NbrView& rev);

// Levels for up to 64 sources (row b of `levels` = source b) on one stream.
struct MsWork {
    uint64_t n;
    DevBuf<unsigned long long> visit, f0, f1, redges;
    DevBuf<uint32_t> q0, q1, dsrc;
    DevBuf<MsCtr> ctr;
    MsCtr* h = nullptr;
    explicit MsWork(uint64_t n_) : n(n_), visit(n_), f0(n_), f1(n_), redges(1), q0(n_), q1(n_), dsrc(64), ctr(1) {
        HIP_OK(hipHostMalloc((void**)&h, sizeof(MsCtr), hipHostMallocDefault));
    }
    ~MsWork() {
        if (h) (void)hipHostFree(h);
    }
    MsCtr read(hipStream_t s) {
        HIP_OK(hipMemcpyAsync(h, ctr.get(), sizeof(MsCtr), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return *h;
    }
};

// One batch of nb <= 64 sources. Push while the frontier holds fewer than
// n / push_div vertices, pull otherwise.
void msbfs_batch(uint64_t n, const uint32_t* src_host, uint32_t nb, const NbrView& fwd, const NbrView& rev,
                 int32_t* levels, MsWork& w, hipStream_t s, uint64_t* reached_edges, uint32_t push_div) {
    const unsigned gv = blocks_for(n, MB, 1u << 16);
    k_ms_fill_levels<<<blocks_for(n * nb, 256, 1u << 16), 256, 0, s>>>(levels, n * nb);
    HIP_OK(hipMemsetAsync(w.visit.get(), 0, n * sizeof(unsigned long long), s));
    HIP_OK(hipMemsetAsync(w.f0.get(), 0, n * sizeof(unsigned long long), s));
    HIP_OK(hipMemcpyAsync(w.dsrc.get(), src_host, nb * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemsetAsync(w.ctr.get(), 0, sizeof(MsCtr), s));
    k_ms_seed<<<1, 1, 0, s>>>(w.dsrc.get(), nb, n, w.visit.get(), w.f0.get(), levels, w.q0.get(), w.ctr.get());
    check_launch();
    unsigned long long valid = 0;
    for (uint32_t b = 0; b < nb; ++b)
        if (src_host[b] < n) valid |= 1ull << b;
    unsigned long long* front = w.f0.get();
    unsigned long long* next = w.f1.get();
    uint32_t* q = w.q0.get();
    uint32_t* nq = w.q1.get();
    uint64_t qlen = w.read(s).qlen;
    bool have_queue = true;
    for (int32_t L = 1; qlen > 0; ++L) {
        HIP_OK(hipMemsetAsync(w.ctr.get(), 0, sizeof(MsCtr), s));
        const bool push = qlen * (uint64_t)push_div < n;
        if (push) {
            if (!have_queue) {
                k_ms_compact<<<gv / MW + 1, MB, 0, s>>>(front, n, q, w.ctr.get());
                check_launch();
                qlen = w.read(s).qlen;
                HIP_OK(hipMemsetAsync(w.ctr.get(), 0, sizeof(MsCtr), s));
                if (qlen == 0) break;
            }
            HIP_OK(hipMemsetAsync(next, 0, n * sizeof(unsigned long long), s));
            timed("msbfs_push", 0.0, s, [&] {
                k_ms_push<<<blocks_for((qlen + 63) / 64, MW, 1u << 16), MB, 0, s>>>(q, (uint32_t)qlen, fwd, front, next,
                                                                                  w.visit.get(), levels, n, L, nq,
                                                                                  w.ctr.get());
            });
            check_launch();
            const MsCtr c = w.read(s);
            qlen = c.qlen;
            std::swap(q, nq);
            have_queue = true;
        } else {
            timed("msbfs_pull", 0.0, s, [&] {
                k_ms_pull<<<gv, MB, 0, s>>>(n, rev, front, next, w.visit.get(), valid, levels, L, w.ctr.get());
            });
            check_launch();
            const MsCtr c = w.read(s);
            // The vertex count of the new frontier is not known after a pull; the
            // settled-pair count bounds it and picks the next direction.
            qlen = c.newbits;
            have_queue = false;
        }
        std::swap(front, next);
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

void msbfs_levels(ursa_hip_topo* t, const std::vector<uint32_t>& src, int dir, int32_t* out_dev,
                  uint64_t* reached_edges, uint32_t push_div) {
    const uint64_t n = t->n;
    hipStream_t s = stream();
    NbrView fwd, rev;
    direction_views(t, dir, s, fwd, rev);
    MsWork w(n);
    for (uint64_t b0 = 0; b0 < src.size(); b0 += 64) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(64, src.size() - b0);
        msbfs_batch(n, src.data() + b0, nb, fwd, rev, out_dev + b0 * n, w, s, reached_edges, push_div);
    }
}

}  // namespace ursa
