// triangle.hip — per-node triangle counts (triangle_count,
// ursa-core/src/algo/triangle.rs:23-63) over the undirected view
// (Topology::undirected, topology.rs:353-389), accumulated in u64 (the
// reference's u32 accumulator wraps on large graphs; u64 never does for m < 2^32).
//
// Orientation: every undirected edge {u, v} is kept once, from the endpoint of
// lower rank, rank(x) = (degree(x), x). Each out-list N+(u) stays sorted by id
// and has at most sqrt(2m) entries. A triangle u < v < w (in rank) is then
// found exactly once: while processing u, as w in N+(u) ∩ N+(v) for v in N+(u).
// Counting: one wavefront per work item, alone in its workgroup (lists of up to
// STAGE_CAP entries), or one workgroup of four per item (longer lists); see
// k_tc_count and k_tc_count_big.
// A work item is a run of consecutive slots of one N+(u): the oriented slots
// are cut at every list start and wherever the running intersection cost
// crosses a multiple of TC_CHUNK, so a hub's list is shared by many waves
// instead of serialising the tail of the launch on one. The wave flattens its
// (v, candidate) pairs over 64 lanes with a degree prefix-sum. Integer sums are order-independent, so counts are exact and
// identical to the node-iterator reference.
//
// Algorithmic bytes (the sorted-merge model): 4 * sum over oriented edges (u,v)
// of (d+(u) + d+(v)), plus 12 B per node.
#include <vector>

#include "frontier.hpp"
#include "prim.hpp"

namespace ursa {

namespace {

constexpr int TCB = 64;  // one wave per workgroup of per-wave items, so waves never wait for each other
constexpr int TCW = TCB / 64;
constexpr uint64_t TC_SMALL_GRID = 1u << 22;  // workgroups of k_tc_count (one item each up to this)
constexpr int TCB_BIG = 256;  // threads of a workgroup sharing one long list
constexpr uint32_t STAGE_CAP = 256;  // N+(u) entries staged in LDS per wave
constexpr uint32_t HASH_CAP = 512;   // hash slots per wave (>= 2 * STAGE_CAP)
constexpr uint32_t TC_BUCKET = 8;    // hash slots per bucket
constexpr uint32_t MID_CAP = 1024;   // N+(u) entries staged per workgroup in the 12 KB table
constexpr uint32_t BIG_CAP = 2048;   // N+(u) entries staged in LDS per workgroup (24 KB table)
constexpr uint64_t TC_CHUNK = 32768;  // intersection steps per work item (about)
constexpr int TC_ILP = 2;            // 64-step groups whose loads are in flight together
constexpr uint32_t TC_LONG = 32;     // walks at least this long are taken by the whole wave
constexpr int TC_LILP = 8;           // 64-step chunks of long walks in flight together

// Intersection cost of oriented slot k = (u, v), as k_tc_count spends it:
// d+(v) hash probes (mode A) or d+(u) binary searches (mode B).
__device__ __forceinline__ uint32_t tc_slot_cost(uint32_t du, uint32_t dv) { return dv <= 8 * du ? dv : du; }

__global__ void k_tc_cost(const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ otgt,
                          const uint32_t* __restrict__ oown, uint64_t kept, uint32_t* __restrict__ cost) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < kept; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = oown[k], v = otgt[k];
        const uint32_t du = ooff[u + 1] - ooff[u], dv = ooff[v + 1] - ooff[v];
        cost[k] = du >= 2 ? tc_slot_cost(du, dv) : 0u;
    }
}

// Item starts: the first slot of every list with at least two entries, and
// every slot whose exclusive cost prefix enters a new TC_CHUNK window.
__global__ void k_tc_item_flags(const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ oown,
                                const uint64_t* __restrict__ pre, uint64_t kept, uint8_t* __restrict__ flag,
                                uint32_t* __restrict__ iota) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < kept; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = oown[k];
        const uint32_t a = ooff[u], du = ooff[u + 1] - a;
        bool f = false;
        // Lists staged per workgroup or read from HBM are cut 512x coarser (about
        // 16 M steps): their staging and count flush are paid once per item.
        const uint64_t win = du > STAGE_CAP ? 512 * TC_CHUNK : TC_CHUNK;
        if (du >= 2) f = (k == a) || (pre[k] / win != pre[k - 1] / win);
        flag[k] = f;
        iota[k] = (uint32_t)k;
    }
}

// Per item: owner, end slot (the next item's start, cut at the list end) and
// class flags: staged per wave (|N+(u)| <= STAGE_CAP) or read from HBM
// (> BIG_CAP), staged per workgroup in a table for up to MID_CAP entries, or
// in one for up to BIG_CAP. The smaller table leaves room for more
// workgroups per CU.
__global__ void k_tc_item_meta(const uint32_t* __restrict__ istart, const uint32_t* __restrict__ oown,
                               const uint32_t* __restrict__ ooff, uint64_t nitems, uint64_t kept,
                               uint32_t* __restrict__ iu, uint32_t* __restrict__ iend, uint8_t* __restrict__ mid,
                               uint8_t* __restrict__ big, uint8_t* __restrict__ small) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nitems; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = oown[istart[i]];
        const uint32_t next = i + 1 < nitems ? istart[i + 1] : (uint32_t)kept;
        const uint32_t du = ooff[u + 1] - ooff[u];
        iu[i] = u;
        iend[i] = min(next, ooff[u + 1]);
        mid[i] = du > STAGE_CAP && du <= MID_CAP;
        big[i] = du > MID_CAP && du <= BIG_CAP;
        small[i] = du <= STAGE_CAP || du > BIG_CAP;
    }
}

__global__ void k_tc_flags(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt,
                           const uint32_t* __restrict__ owner, uint64_t slots, uint8_t* __restrict__ keep) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < slots;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = owner[k], v = tgt[k];
        const uint32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
        keep[k] = (dv > du) || (dv == du && v > u);
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

// Multiplicative hash, bits 7.. of x * 0x9E3779B1. Tables have fewer than 2^17
// buckets, so only product bits below 24 are kept, and those depend only on
// the low 24 bits of x and of the multiplier: a full-rate 24-bit multiply
// gives the same value.
__device__ __forceinline__ uint32_t tc_hash(uint32_t x, uint32_t mask) { return (__umul24(x, 0x3779B1u) >> 7) & mask; }
static_assert(2 * BIG_CAP / TC_BUCKET <= (1u << 17), "tc_hash keeps product bits below 24");

// The LDS hash set: buckets of TC_BUCKET = 8 slots (32 bytes), key x homed
// in bucket tc_hash(x) and placed in its first free slot, spilling to the next
// bucket when full (load factor <= 1/2, so a spill is rare). Slots of a bucket
// fill in order, so a bucket is full exactly when its last slot is taken, and
// a key can lie past its home bucket only if that bucket is full. A lookup
// reads its home bucket with two 16-byte loads and tests all eight slots at
// once in vector arithmetic: d_i = key_i ^ x is zero only at a match (keys of
// a table are distinct, so at most one matches), and the match's slot number
// is read off the minima of the d_i over the slots with each index bit set.
// `bmask` is the bucket count minus one. Returns x's slot, or -1.
__device__ __forceinline__ int tc_probe_bucket(const uint32_t* key, uint32_t b, uint32_t x, bool& full) {
    const uint4 k0 = reinterpret_cast<const uint4*>(key)[2 * b];
    const uint4 k1 = reinterpret_cast<const uint4*>(key)[2 * b + 1];
    full = k1.w != 0xFFFFFFFFu;
    const uint32_t d0 = k0.x ^ x, d1 = k0.y ^ x, d2 = k0.z ^ x, d3 = k0.w ^ x;
    const uint32_t d4 = k1.x ^ x, d5 = k1.y ^ x, d6 = k1.z ^ x, d7 = k1.w ^ x;
    const uint32_t mhi = min(min(d4, d5), min(d6, d7));
    if (min(min(min(d0, d1), min(d2, d3)), mhi) != 0) return -1;
    const uint32_t f0 = 1u - min(min(min(d1, d3), min(d5, d7)), 1u);  // match in an odd slot
    const uint32_t f1 = 1u - min(min(min(d2, d3), min(d6, d7)), 1u);  // slot 2, 3, 6 or 7
    const uint32_t f2 = 1u - min(mhi, 1u);                            // slot 4..7
    return (int)(b * TC_BUCKET + f0 + 2 * f1 + 4 * f2);
}

__device__ __forceinline__ int tc_probe(const uint32_t* key, uint32_t bmask, uint32_t x) {
    uint32_t b = tc_hash(x, bmask);
    bool full;
    int r = tc_probe_bucket(key, b, x, full);
    if (r < 0 && full) {  // rare: follow the spill into the next buckets
        for (;;) {
            b = (b + 1) & bmask;
            r = tc_probe_bucket(key, b, x, full);
            if (r >= 0 || !full) break;
        }
    }
    return r;
}

// Records a triangle (u, v, w) found with w at hash slot wslot (staged) or
// w itself (not staged): w's credit; v's credit and u's are added by the caller.
// Slot counters are 16-bit, two per LDS word (a slot is credited at most
// once per v of the item plus its own hits, both below 2^16).
__device__ __forceinline__ void tc_cnt_add(uint32_t* cnt, int slot, uint32_t k) {
    atomicAdd(&cnt[slot >> 1], k << (16 * (slot & 1)));
}

__device__ __forceinline__ void tc_credit_w(bool staged, int wslot, uint32_t w, uint32_t* cnt,
                                            unsigned long long* __restrict__ tri) {
    if (staged)
        tc_cnt_add(cnt, wslot, 1u);
    else
        atomicAdd(&tri[w], 1ull);
}

// Is w in N+(u)? Sets wslot when staged. (A bloom bitmap in front of the
// table cannot pay on a wave: the table is read whenever any of its 64 lanes
// passes, which with about one hit in ten steps is nearly always.)
__device__ __forceinline__ bool tc_member(bool staged, const uint32_t* key, uint32_t hmask,
                                          const uint32_t* ulist, uint32_t du, uint32_t w, int& wslot) {
    if (staged) {
        wslot = tc_probe(key, hmask, w);
        return wslot >= 0;
    }
    return sorted_contains(ulist, du, w);
}

// The intersections of one work item, slots [s0, s1) of N+(u), by one wave
// that takes every `step`-th group of 64 slots starting at s0 + first. With
// `staged`, N+(u) sits in an LDS open-addressing hash set whose slots also
// carry hit counters. For each v in the range the wave intersects N+(u) and
// N+(v) from the cheaper side:
//   A  (|N+(v)| <= 8 |N+(u)|): walk N+(v) from HBM, probe N+(u);
//   B  otherwise:              walk N+(u), binary-search N+(v) in HBM.
// A v whose walk has at least TC_LONG steps is walked by the whole wave in
// 64-step chunks; the shorter walks of the group are flattened over the 64
// lanes with a prefix sum and an owner search.
// A hit w closes the triangle (u, v, w): the returned count (wave-uniform) is
// u's share, and v and w (both in N+(u)) are credited in the slot counters (or
// in HBM when N+(u) is not staged). Integer sums are order-independent.
__device__ __forceinline__ unsigned long long tc_item(const uint32_t* __restrict__ ooff,
                                                      const uint32_t* __restrict__ otgt, uint32_t a, uint32_t du,
                                                      uint32_t s0, uint32_t s1, uint32_t first, uint32_t step,
                                                      bool staged, const uint32_t* key, uint32_t* cnt,
                                                      uint32_t hmask,
                                                      unsigned long long* __restrict__ tri) {
    const int lane = lane_id();
    const uint32_t* ulist = otgt + a;
    unsigned long long cu = 0;
    for (uint32_t i0 = s0 + first; i0 < s1; i0 += step) {
        const uint32_t i = i0 + lane;
        uint32_t v = 0, dv = 0, sv = 0, c = 0;
        int vslot = -1;
        bool mode_a = true;
        if (i < s1) {
            v = ulist[i];
            sv = ooff[v];
            dv = ooff[v + 1] - sv;
            mode_a = dv <= 8 * du;
            c = mode_a ? dv : du;
            if (staged) vslot = tc_probe(key, hmask, v);
        }
        // Long walks: cut into 64-step chunks walked by the whole wave, lane l
        // taking step l of the chunk. A group of up to TC_LILP chunks of one
        // list has its loads in flight together; groups never cross lists, so
        // the list's data is read once per group. The owner of chunk ch is the
        // last lane whose chunk prefix is <= ch (prefixes are non-decreasing).
        // Loads are clamped into the list instead of masked, and lanes past
        // its end test a repeated element and discard the result.
        const uint32_t nch = c >= TC_LONG ? (c + 63) / 64 : 0;
        const uint32_t chi = wave_incl_scan(nch);
        const uint32_t che = chi - nch;
        const uint32_t nchunks = __shfl(chi, 63, 64);
        for (uint32_t ch0 = 0; ch0 < nchunks;) {
            const int j = (int)__popcll(__ballot(che <= ch0)) - 1;
            const uint32_t ng = min((uint32_t)TC_LILP, __builtin_amdgcn_readlane(chi, j) - ch0);
            const uint32_t oc = __builtin_amdgcn_readlane(c, j);
            const uint32_t osv = __builtin_amdgcn_readlane(sv, j);
            const bool oa = __builtin_amdgcn_readlane((int)mode_a, j) != 0;
            const uint32_t e0 = (ch0 - __builtin_amdgcn_readlane(che, j)) * 64 + lane;
            const uint32_t* walk = oa ? otgt + osv : ulist;  // mode A walks N+(v), mode B N+(u)
            uint32_t w[TC_LILP];
#pragma unroll
            for (int r = 0; r < TC_LILP; ++r)
                if ((uint32_t)r < ng) w[r] = walk[min(e0 + 64 * r, oc - 1)];
            uint32_t h = 0;
#pragma unroll
            for (int r = 0; r < TC_LILP; ++r) {
                if ((uint32_t)r >= ng) break;
                const bool ok = e0 + 64 * r < oc;
                int wslot = -1;
                bool hit;
                if (oa && staged) {
                    wslot = tc_probe(key, hmask, w[r]);
                    hit = ok && wslot >= 0;
                } else if (oa) {
                    hit = ok && sorted_contains(ulist, du, w[r]);
                } else {
                    hit = ok && sorted_contains(otgt + osv, __builtin_amdgcn_readlane(dv, j), w[r]);
                    if (hit && staged) wslot = tc_probe(key, hmask, w[r]);
                }
                if (hit) tc_credit_w(staged, wslot, w[r], cnt, tri);
                h += (uint32_t)__popcll(__ballot(hit));
            }
            cu += h;
            if (lane == 0 && h) {
                if (staged)
                    tc_cnt_add(cnt, __builtin_amdgcn_readlane(vslot, j), h);
                else
                    atomicAdd(&tri[__builtin_amdgcn_readlane(v, j)], (unsigned long long)h);
            }
            ch0 += ng;
        }
        // Short walks, flattened.
        if (c >= TC_LONG) c = 0;
        const uint32_t incl = wave_incl_scan(c);
        const uint32_t excl = incl - c;
        const uint32_t total = __shfl(incl, 63, 64);
        for (uint32_t t0 = 0; t0 < total; t0 += 64 * TC_ILP) {
            uint32_t w[TC_ILP], ovv[TC_ILP], osvv[TC_ILP], odvv[TC_ILP];
            int ovs[TC_ILP];
            bool oa[TC_ILP], ok[TC_ILP];
#pragma unroll
            for (int r = 0; r < TC_ILP; ++r) {
                const uint32_t t = t0 + 64 * r + lane;
                const int j = wave_owner(excl, t);
                const uint32_t e = t - __shfl(excl, j, 64);
                ovv[r] = __shfl(v, j, 64);
                osvv[r] = __shfl(sv, j, 64);
                odvv[r] = __shfl(dv, j, 64);
                ovs[r] = __shfl(vslot, j, 64);
                oa[r] = __shfl((int)mode_a, j, 64) != 0;
                ok[r] = t < total;
                w[r] = 0;
                if (ok[r]) w[r] = oa[r] ? otgt[osvv[r] + e] : ulist[e];
            }
#pragma unroll
            for (int r = 0; r < TC_ILP; ++r) {
                bool hit = false;
                int wslot = -1;
                if (!ok[r]) {
                    // nothing to test
                } else if (oa[r]) {
                    hit = tc_member(staged, key, hmask, ulist, du, w[r], wslot);
                } else {
                    hit = sorted_contains(otgt + osvv[r], odvv[r], w[r]);
                    if (hit && staged) wslot = tc_probe(key, hmask, w[r]);
                }
                if (hit) {
                    tc_credit_w(staged, wslot, w[r], cnt, tri);
                    if (staged)
                        tc_cnt_add(cnt, ovs[r], 1u);
                    else
                        atomicAdd(&tri[ovv[r]], 1ull);
                }
                cu += (uint64_t)__popcll(__ballot(ok[r] && hit));
            }
        }
    }
    return cu;
}

// Makes a table's LDS writes visible to its readers: the whole workgroup, or,
// for a table owned by one wavefront (WAVE), that wavefront alone, whose LDS
// accesses complete in program order, so only the compiler must be kept from
// moving accesses across.
template <bool WAVE>
__device__ __forceinline__ void tc_sync() {
    if (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    } else {
        __syncthreads();
    }
}

// Stages N+(u) into an LDS hash set of 2^k >= du / 4 buckets (at least two
// slots per entry) with zeroed counters, using `nthreads` threads. Ends with a
// barrier; returns the bucket mask.
template <bool WAVE>
__device__ __forceinline__ uint32_t tc_stage(const uint32_t* __restrict__ otgt, uint32_t a, uint32_t du,
                                             uint32_t tid, uint32_t nthreads, uint32_t* key, uint32_t* cnt) {
    uint32_t nb = 2;
    while (nb * TC_BUCKET < 2 * du) nb <<= 1;
    const uint32_t h = nb * TC_BUCKET;
    for (uint32_t i = tid; i < h; i += nthreads) key[i] = 0xFFFFFFFFu;
    for (uint32_t i = tid; i < h / 2; i += nthreads) cnt[i] = 0;
    tc_sync<WAVE>();
    for (uint32_t i = tid; i < du; i += nthreads) {
        const uint32_t x = otgt[a + i];
        uint32_t b = tc_hash(x, nb - 1);
        for (;;) {
            bool done = false;
            for (uint32_t k = 0; k < TC_BUCKET && !done; ++k)
                done = atomicCAS(&key[b * TC_BUCKET + k], 0xFFFFFFFFu, x) == 0xFFFFFFFFu;
            if (done) break;
            b = (b + 1) & (nb - 1);
        }
    }
    tc_sync<WAVE>();
    return nb - 1;
}

// Adds the slot counters of a staged list into tri. Ends with a barrier.
template <bool WAVE>
__device__ __forceinline__ void tc_flush(const uint32_t* key, const uint32_t* cnt, uint32_t h, uint32_t tid,
                                         uint32_t nthreads, unsigned long long* __restrict__ tri) {
    tc_sync<WAVE>();
    for (uint32_t i = tid; i < h; i += nthreads) {
        const uint32_t c = (cnt[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
        if (c) atomicAdd(&tri[key[i]], (unsigned long long)c);
    }
    tc_sync<WAVE>();
}

// One wavefront per item whose list has at most STAGE_CAP entries (staged in
// the wave's LDS) or more than BIG_CAP (read from HBM). Every wave owns its
// table and walks its own items, never waiting for the other waves.
__global__ __launch_bounds__(TCB, 8) void k_tc_count(const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ otgt,
                                                  const uint32_t* __restrict__ istart, const uint32_t* __restrict__ iend,
                                                  const uint32_t* __restrict__ iu, uint64_t nwork,
                                                  unsigned long long* __restrict__ tri) {
    __shared__ __align__(16) uint32_t s_key[TCW][HASH_CAP];
    __shared__ uint32_t s_cnt[TCW][HASH_CAP / 2];
    const int wid = threadIdx.x >> 6;
    const int lane = lane_id();
    uint32_t* key = s_key[wid];
    uint32_t* cnt = s_cnt[wid];
    const uint64_t nwaves = (uint64_t)gridDim.x * TCW;
    for (uint64_t wi = blockIdx.x * (uint64_t)TCW + wid; wi < nwork; wi += nwaves) {
        const uint32_t u = iu[wi];
        const uint32_t a = ooff[u];
        const uint32_t du = ooff[u + 1] - a;
        const uint32_t s0 = istart[wi] - a, s1 = iend[wi] - a;
        const bool staged = du <= STAGE_CAP;
        const uint32_t hmask = tc_stage<true>(otgt, a, staged ? du : 0, lane, 64, key, cnt);
        const unsigned long long cu =
            tc_item(ooff, otgt, a, du, s0, s1, 0, 64, staged, key, cnt, hmask, tri);
        if (lane == 0 && cu) atomicAdd(&tri[u], cu);
        tc_flush<true>(key, cnt, (hmask + 1) * TC_BUCKET, lane, 64, tri);
    }
}

// One workgroup per item whose list has more than STAGE_CAP and at most CAP
// entries (MID_CAP or BIG_CAP, the item classes above): the list is staged once
// in the group's LDS and the group's TCB_BIG / 64 waves share the item's slots.
template <uint32_t CAP>
__global__ __launch_bounds__(TCB_BIG) void k_tc_count_big(const uint32_t* __restrict__ ooff,
                                                             const uint32_t* __restrict__ otgt,
                                                             const uint32_t* __restrict__ istart,
                                                             const uint32_t* __restrict__ iend,
                                                             const uint32_t* __restrict__ iu, uint64_t nwork,
                                                             unsigned long long* __restrict__ tri) {
    __shared__ __align__(16) uint32_t b_key[2 * CAP];
    __shared__ uint32_t b_cnt[CAP];
    const int wid = threadIdx.x >> 6;
    const int lane = lane_id();
    for (uint64_t wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        const uint32_t u = iu[wi];
        const uint32_t a = ooff[u];
        const uint32_t du = ooff[u + 1] - a;
        const uint32_t s0 = istart[wi] - a, s1 = iend[wi] - a;
        const uint32_t hmask = tc_stage<false>(otgt, a, du, threadIdx.x, TCB_BIG, b_key, b_cnt);
        const unsigned long long cu =
            tc_item(ooff, otgt, a, du, s0, s1, 64u * wid, TCB_BIG, true, b_key, b_cnt, hmask, tri);
        if (lane == 0 && cu) atomicAdd(&tri[u], cu);
        tc_flush<false>(b_key, b_cnt, (hmask + 1) * TC_BUCKET, threadIdx.x, TCB_BIG, tri);
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
    DevBuf<uint32_t> oown;  // owner of each oriented slot
    if (slots) {
        DevBuf<uint32_t> owner(slots);
        slot_owners(n, U, owner.get(), s);
        DevBuf<uint8_t> keep(slots);
        k_tc_flags<<<blocks_for(slots, 256), 256, 0, s>>>(U.off.get(), U.tgt.get(), owner.get(), slots, keep.get());
        check_launch();
        oown.alloc(slots);
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
    // Work items (see the file comment).
    uint64_t nwork = 0;
    if (kept) {
        DevBuf<uint32_t> cost(kept), iota(kept);
        DevBuf<uint64_t> pre(kept + 1);
        DevBuf<uint8_t> flag(kept);
        k_tc_cost<<<blocks_for(kept, 256), 256, 0, s>>>(t->tc_off.get(), t->tc_tgt.get(), oown.get(), kept,
                                                       cost.get());
        check_launch();
        prim::exclusive_sum_u32_to_u64(cost.get(), pre.get(), kept, s);
        k_tc_item_flags<<<blocks_for(kept, 256), 256, 0, s>>>(t->tc_off.get(), oown.get(), pre.get(), kept,
                                                             flag.get(), iota.get());
        check_launch();
        DevBuf<uint32_t> istart(kept), iu(kept), iend(kept);
        DevBuf<uint64_t> wc(1);
        prim::select_flagged_u32(iota.get(), flag.get(), istart.get(), wc.get(), kept, s);
        uint64_t nitems = 0;
        HIP_OK(hipMemcpyAsync(&nitems, wc.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        DevBuf<uint8_t> mid(nitems ? nitems : 1), big(nitems ? nitems : 1), small(nitems ? nitems : 1);
        if (nitems) {
            k_tc_item_meta<<<blocks_for(nitems, 256), 256, 0, s>>>(istart.get(), oown.get(), t->tc_off.get(), nitems,
                                                                  kept, iu.get(), iend.get(), mid.get(), big.get(),
                                                                  small.get());
            check_launch();
        }
        // Split the items into the per-wave list and the per-workgroup list.
        auto part = [&](const uint8_t* f, DevBuf<uint32_t>& a, DevBuf<uint32_t>& b, DevBuf<uint32_t>& c) -> uint64_t {
            uint64_t cnt = 0;
            a.alloc(nitems ? nitems : 1);
            b.alloc(nitems ? nitems : 1);
            c.alloc(nitems ? nitems : 1);
            if (!nitems) return 0;
            prim::select_flagged_u32(istart.get(), f, a.get(), wc.get(), nitems, s);
            prim::select_flagged_u32(iend.get(), f, b.get(), wc.get(), nitems, s);
            prim::select_flagged_u32(iu.get(), f, c.get(), wc.get(), nitems, s);
            HIP_OK(hipMemcpyAsync(&cnt, wc.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            return cnt;
        };
        nwork = part(small.get(), t->tc_istart, t->tc_iend, t->tc_iu);
        t->tc_nbig = part(mid.get(), t->tc_bstart, t->tc_bend, t->tc_bu);
        t->tc_nhuge = part(big.get(), t->tc_hstart, t->tc_hend, t->tc_hu);
    } else {
        for (auto* b : {&t->tc_istart, &t->tc_iend, &t->tc_iu, &t->tc_bstart, &t->tc_bend, &t->tc_bu, &t->tc_hstart,
                        &t->tc_hend, &t->tc_hu})
            b->alloc(1);
        t->tc_nbig = 0;
        t->tc_nhuge = 0;
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
    if (t->tc_nwork || t->tc_nbig || t->tc_nhuge) {
        const double bytes = 4.0 * t->tc_pair_work + 12.0 * (double)n;
        auto* tri = reinterpret_cast<unsigned long long*>(out.p);
        timed("triangle_count", bytes, s, [&] {
            if (t->tc_nhuge)
                k_tc_count_big<BIG_CAP><<<(unsigned)std::min<uint64_t>(t->tc_nhuge, 1u << 16), TCB_BIG, 0, s>>>(
                    t->tc_off.get(), t->tc_tgt.get(), t->tc_hstart.get(), t->tc_hend.get(), t->tc_hu.get(),
                    t->tc_nhuge, tri);
            if (t->tc_nbig)
                k_tc_count_big<MID_CAP><<<(unsigned)std::min<uint64_t>(t->tc_nbig, 1u << 16), TCB_BIG, 0, s>>>(
                    t->tc_off.get(), t->tc_tgt.get(), t->tc_bstart.get(), t->tc_bend.get(), t->tc_bu.get(),
                    t->tc_nbig, tri);
            if (t->tc_nwork)
                k_tc_count<<<(unsigned)std::min<uint64_t>((t->tc_nwork + TCW - 1) / TCW, TC_SMALL_GRID), TCB, 0, s>>>(
                    t->tc_off.get(), t->tc_tgt.get(), t->tc_istart.get(), t->tc_iend.get(), t->tc_iu.get(),
                    t->tc_nwork, tri);
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
