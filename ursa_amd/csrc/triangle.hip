// triangle.hip — per-node triangle counts (triangle_count,
// ursa-core/src/algo/triangle.rs:23-63) over the undirected view
// (Topology::undirected, topology.rs:353-389), accumulated in u64 (the
// reference's u32 accumulator wraps on large graphs; u64 never does for m < 2^32).
//
// Orientation: every undirected edge {u, v} is kept once, from the endpoint of
// lower rank, rank(x) = (degree(x), x). Each out-list N+(u) stays sorted by id
// and has at most sqrt(2m) entries. A triangle u < v < w (in rank) is then
// found exactly once: while processing u, as w in N+(u) ∩ N+(v) for v in N+(u).
// Counting: one wavefront per vertex u (four per workgroup); see k_tc_count.
// The wave flattens its (v, candidate) pairs over 64 lanes with a degree
// prefix-sum. Integer sums are order-independent, so counts are exact and
// identical to the node-iterator reference.
//
// Algorithmic bytes (the sorted-merge model): 4 * sum over oriented edges (u,v)
// of (d+(u) + d+(v)), plus 12 B per node.
#include <vector>

#include "frontier.hpp"
#include "prim.hpp"

namespace ursa {

namespace {

constexpr int TCB = 256;
constexpr int TCW = TCB / 64;
constexpr uint32_t STAGE_CAP = 512;  // N+(u) entries staged in LDS per wave
constexpr uint32_t HASH_CAP = 1024;  // hash slots per wave (>= 2 * STAGE_CAP)

__global__ void k_tc_flags(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt,
                           const uint32_t* __restrict__ owner, uint64_t slots, uint8_t* __restrict__ keep) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < slots;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = owner[k], v = tgt[k];
        const uint32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
        keep[k] = (dv > du) || (dv == du && v > u);
    }
}

__global__ void k_tc_workflags(const uint32_t* __restrict__ ooff, uint64_t n, uint8_t* flag, uint32_t* iota) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        flag[u] = (ooff[u + 1] - ooff[u]) >= 2;
        iota[u] = (uint32_t)u;
    }
}

// sum over oriented edges (u,v) of d+(u) + d+(v): the merge-intersection
// length, the traffic model of sorted-list triangle counting.
__global__ void k_tc_pairs(const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ otgt, uint64_t n,
                           unsigned long long* total) {
    unsigned long long acc = 0;
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = ooff[u], z = ooff[u + 1];
        for (uint32_t k = a; k < z; ++k) {
            const uint32_t v = otgt[k];
            acc += (z - a) + (ooff[v + 1] - ooff[v]);
        }
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0 && acc) atomicAdd(total, acc);
}

__device__ __forceinline__ bool sorted_contains(const uint32_t* a, uint32_t len, uint32_t x) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t y = a[mid];
        if (y < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < len && a[lo] == x;
}

// Index of x in sorted a[0..len), or -1.
__device__ __forceinline__ int sorted_find(const uint32_t* a, uint32_t len, uint32_t x) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < len && a[lo] == x) ? (int)lo : -1;
}

__device__ __forceinline__ uint32_t tc_hash(uint32_t x, uint32_t mask) { return (x * 2654435761u >> 7) & mask; }

// One wavefront per vertex u with |N+(u)| >= 2. When N+(u) fits (<= STAGE_CAP
// entries) it is staged in LDS as a sorted list, an open-addressing hash set
// (slot -> list index) and a per-entry hit counter. For each v in N+(u) the
// wave intersects N+(u) and N+(v) from the cheaper side:
//   A  (|N+(v)| <= 8 |N+(u)|): walk N+(v) from HBM, probe the LDS hash;
//   B  otherwise:              walk N+(u) from LDS, binary-search N+(v) in HBM.
// A hit w closes the triangle (u, v, w): u's count is wave-reduced, and v and w
// (both in N+(u)) are credited in the LDS counters, flushed to HBM once per u.
__global__ __launch_bounds__(TCB) void k_tc_count(const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ otgt,
                                                  const uint32_t* __restrict__ work, uint64_t nwork,
                                                  unsigned long long* __restrict__ tri) {
    __shared__ uint32_t s_list[TCW][STAGE_CAP];
    __shared__ uint32_t s_cnt[TCW][STAGE_CAP];
    __shared__ uint32_t s_key[TCW][HASH_CAP];
    __shared__ uint16_t s_idx[TCW][HASH_CAP];
    const int wid = threadIdx.x >> 6;
    const int lane = lane_id();
    uint32_t* list = s_list[wid];
    uint32_t* cnt = s_cnt[wid];
    uint32_t* key = s_key[wid];
    uint16_t* hidx = s_idx[wid];
    for (uint64_t g = blockIdx.x; g * TCW < nwork; g += gridDim.x) {
        const uint64_t wi = g * TCW + wid;
        const bool active = wi < nwork;
        const uint32_t u = active ? work[wi] : 0;
        const uint32_t a = active ? ooff[u] : 0;
        const uint32_t du = active ? ooff[u + 1] - a : 0;
        const bool staged = du <= STAGE_CAP;
        uint32_t hmask = 0;
        if (staged && du) {
            uint32_t h = 16;
            while (h < 2 * du) h <<= 1;
            hmask = h - 1;
            for (uint32_t i = lane; i < h; i += 64) key[i] = 0xFFFFFFFFu;
            for (uint32_t i = lane; i < du; i += 64) {
                list[i] = otgt[a + i];
                cnt[i] = 0;
            }
        }
        __syncthreads();
        if (staged && du) {
            for (uint32_t i = lane; i < du; i += 64) {
                const uint32_t x = list[i];
                uint32_t slot = tc_hash(x, hmask);
                while (atomicCAS(&key[slot], 0xFFFFFFFFu, x) != 0xFFFFFFFFu) slot = (slot + 1) & hmask;
                hidx[slot] = (uint16_t)i;
            }
        }
        __syncthreads();
        const uint32_t* ulist = staged ? list : otgt + a;
        unsigned long long cu = 0;
        for (uint32_t i0 = 0; i0 < du; i0 += 64) {
            const uint32_t i = i0 + lane;
            uint32_t v = 0, dv = 0, sv = 0, c = 0;
            bool mode_a = true;
            if (i < du) {
                v = ulist[i];
                sv = ooff[v];
                dv = ooff[v + 1] - sv;
                mode_a = dv <= 8 * du;
                c = mode_a ? dv : du;
            }
            const uint32_t incl = wave_incl_scan(c);
            const uint32_t excl = incl - c;
            const uint32_t total = __shfl(incl, 63, 64);
            for (uint32_t t0 = 0; t0 < total; t0 += 64) {
                const uint32_t t = t0 + lane;
                const int j = wave_owner(excl, t);
                const uint32_t e = t - __shfl(excl, j, 64);
                const uint32_t ov = __shfl(v, j, 64);
                const uint32_t osv = __shfl(sv, j, 64);
                const uint32_t odv = __shfl(dv, j, 64);
                const bool oa = __shfl((int)mode_a, j, 64) != 0;
                if (t >= total) continue;
                int hit = -1;  // index of w in N+(u)
                uint32_t w;
                if (oa) {
                    w = otgt[osv + e];
                    if (staged) {
                        uint32_t slot = tc_hash(w, hmask);
                        uint32_t k;
                        while ((k = key[slot]) != 0xFFFFFFFFu) {
                            if (k == w) {
                                hit = hidx[slot];
                                break;
                            }
                            slot = (slot + 1) & hmask;
                        }
                    } else {
                        hit = sorted_find(ulist, du, w);
                    }
                } else {
                    w = ulist[e];
                    if (sorted_contains(otgt + osv, odv, w)) hit = (int)e;
                }
                if (hit >= 0) {
                    ++cu;
                    if (staged) {
                        atomicAdd(&cnt[hit], 1u);
                        atomicAdd(&cnt[i0 + j], 1u);
                    } else {
                        atomicAdd(&tri[w], 1ull);
                        atomicAdd(&tri[ov], 1ull);
                    }
                }
            }
        }
        cu = wave_sum_u64(cu);
        if (lane == 0 && cu) atomicAdd(&tri[u], cu);
        __syncthreads();
        if (staged)
            for (uint32_t i = lane; i < du; i += 64)
                if (cnt[i]) atomicAdd(&tri[list[i]], (unsigned long long)cnt[i]);
        __syncthreads();
    }
}

}  // namespace

static void tc_orient(ursa_hip_topo* t, hipStream_t s) {
    NoArena keep_out_of_arena;
    std::lock_guard<std::mutex> lk(t->cache_mu);
    if (t->tc_built) return;
    const uint64_t n = t->n;
    const DevCsr& U = t->undirected(s);
    const uint64_t slots = U.slots;
    t->tc_off.alloc(n + 1);
    uint64_t kept = 0;
    if (slots) {
        DevBuf<uint32_t> owner(slots);
        slot_owners(n, U, owner.get(), s);
        DevBuf<uint8_t> keep(slots);
        k_tc_flags<<<blocks_for(slots, 256), 256, 0, s>>>(U.off.get(), U.tgt.get(), owner.get(), slots, keep.get());
        check_launch();
        DevBuf<uint32_t> oown(slots);
        t->tc_tgt.alloc(slots);
        DevBuf<uint64_t> cnt(1);
        prim::select_flagged_u32(U.tgt.get(), keep.get(), t->tc_tgt.get(), cnt.get(), slots, s);
        prim::select_flagged_u32(owner.get(), keep.get(), oown.get(), cnt.get(), slots, s);
        HIP_OK(hipMemcpyAsync(&kept, cnt.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        offsets_from_sorted(oown.get(), kept, n, t->tc_off.get(), s);
    } else {
        t->tc_tgt.alloc(1);
        offsets_from_sorted(nullptr, 0, n, t->tc_off.get(), s);
    }
    // Work list: vertices with at least two oriented neighbours.
    DevBuf<uint8_t> flag(n ? n : 1);
    DevBuf<uint32_t> iota(n ? n : 1);
    t->tc_work.alloc(n ? n : 1);
    DevBuf<uint64_t> wc(1);
    uint64_t nwork = 0;
    if (n) {
        k_tc_workflags<<<blocks_for(n, 256), 256, 0, s>>>(t->tc_off.get(), n, flag.get(), iota.get());
        check_launch();
        prim::select_flagged_u32(iota.get(), flag.get(), t->tc_work.get(), wc.get(), n, s);
        HIP_OK(hipMemcpyAsync(&nwork, wc.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    }
    DevBuf<unsigned long long> pairs(1);
    HIP_OK(hipMemsetAsync(pairs.get(), 0, sizeof(unsigned long long), s));
    if (kept) {
        k_tc_pairs<<<blocks_for(n, 256, 1u << 14), 256, 0, s>>>(t->tc_off.get(), t->tc_tgt.get(), n, pairs.get());
        check_launch();
    }
    unsigned long long hp = 0;
    HIP_OK(hipMemcpyAsync(&hp, pairs.get(), sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    t->tc_slots = kept;
    t->tc_nwork = nwork;
    t->tc_pair_work = (double)hp;
    t->tc_built = true;
}

void triangle_prepare(ursa_hip_topo* t) { tc_orient(t, stream()); }

void triangle_count(ursa_hip_topo* t, uint64_t* out_user) {
    const uint64_t n = t->n;
    if (n == 0) return;  // triangle.rs:59-61
    hipStream_t s = stream();
    tc_orient(t, s);
    DevOut<uint64_t> out(out_user, n);
    HIP_OK(hipMemsetAsync(out.p, 0, n * sizeof(uint64_t), s));
    if (t->tc_nwork) {
        const double bytes = 4.0 * t->tc_pair_work + 12.0 * (double)n;
        const unsigned grid = (unsigned)std::min<uint64_t>((t->tc_nwork + TCW - 1) / TCW, 1u << 16);
        timed("triangle_count", bytes, s, [&] {
            k_tc_count<<<grid, TCB, 0, s>>>(t->tc_off.get(), t->tc_tgt.get(), t->tc_work.get(), t->tc_nwork,
                                            reinterpret_cast<unsigned long long*>(out.p));
        });
        check_launch();
    }
    out.finish(s);
}

}  // namespace ursa

extern "C" int ursa_hip_triangle_count(ursa_hip_topo* t, uint64_t* out) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (t->n == 0) return;
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        ursa::triangle_count(t, out);
    });
}
