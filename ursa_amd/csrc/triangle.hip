// triangle.hip — per-node triangle counts (triangle_count,
// ursa-core/src/algo/triangle.rs:23-63) over the undirected view
// (Topology::undirected, topology.rs:353-389), accumulated in u64 (the
// reference's u32 accumulator wraps on large graphs; u64 never does for m < 2^32).
//
// Orientation: every undirected edge {u, v} is kept once, from the endpoint of
// lower rank, rank(x) = (degree(x), x), and has at most sqrt(2m) entries per
// out-list. The oriented graph is stored with its vertices numbered by
// decreasing rank (new id 0 = highest rank), so every edge points to a smaller
// new id and N+(u) lies in [0, u); each out-list is sorted by new id. Counts
// are made per new id and mapped back to the caller's ids at the end.
// A triangle u < v < w (in rank) is found exactly once, as w in
// N+(u) ∩ N+(v) for the pair u -> v, intersected from one side: N+(v) walked
// against a table of N+(u), or, when N+(v) is a hub's list, the entries of
// N+(u) before v (the only ones that can close a triangle: w lies in [0, v))
// walked against a table of N+(v) (tc_walk_out / tc_walk_in); the second kind
// of work runs over the in-lists In+(v) of the oriented graph. At s22 the cut
// halves that kind's walks (6.87 G to 3.44 G entries under the d+(u) < d+(v)
// rule) and lets hub pairs take the shorter side.
// Tables: an LDS hash set (lists of up to BIG_CAP entries), or, for an owner
// u below BM_BITS whose list is longer than STAGE_CAP, a bitmap of N+(u) over
// [0, u) with per-word prefix counts (a lookup is one LDS word; the member's
// index, for its hit counter, is the prefix plus a popcount). On R-MAT graphs
// the owners of long lists are the few highest-ranked vertices (at scale 22
// all 35 K lists longer than 256 belong to new ids below 35.5 K; 89 % of the
// probes go to them), so the bitmap serves nearly all the work.
// Counting: one wavefront per work item, alone in its workgroup (hash tables
// of up to STAGE_CAP entries), or one workgroup of four per item (longer
// lists); see k_tc_count, k_tc_count_big and k_tc_count_bm.
// A work item is a run of consecutive slots of one list (an out-list, or an
// in-list for the second kind): the slots are cut at every list start and
// wherever the running intersection cost crosses a multiple of TC_CHUNK, so a
// hub's list is shared by many waves instead of serialising the tail of the
// launch on one. The wave flattens its (partner, candidate) pairs over 64
// lanes with a degree prefix-sum. Integer sums are order-independent, so
// counts are exact and identical to the node-iterator reference.
//
// Algorithmic bytes: 4 B per walked entry, 12 B per oriented pair (the
// partner id and its list offsets) and 12 B per node (the one-sided walk
// model); the sorted-merge model, 4 * sum over oriented edges (u,v) of
// (d+(u) + d+(v)) plus 12 B per node, is reported beside it.
#include <atomic>
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
constexpr uint32_t BM_BITS = 65536;  // bitmap tables cover new ids [0, BM_BITS) (8 KB)
constexpr uint32_t BM_WORDS = BM_BITS / 32;
constexpr uint64_t TC_CHUNK = 131072;  // intersection steps per work item (about; 65536 measured 3-4 % slower once hub-side walks stop below the owner)
constexpr uint32_t TC_MAX_SLOTS = 32768;  // list slots per work item at most (keeps hit counters below 2^16)
constexpr int TC_ILP = 4;            // 64-step groups whose loads are in flight together (s22: 2 -> 4 cut 25.9 to 21.4 ms, 8 gave 22.1)
constexpr uint32_t TC_LONG = 16;     // walks at least this long are taken by the whole wave (8 and 16 alone measured alike, 64 20 % slower)
constexpr int TC_LILP = 8;           // 64-step chunks of long walks in flight together

// Each oriented edge x -> y (a pair; y in N+(x), so y < x) is intersected once.
// A triangle z < y < x closes at z in N+(x) ∩ N+(y), and every such z lies
// below y: in N+(x) only the entries before y can close one. So a pair whose y
// is the first entry of N+(x) (position 0) has no triangle, and otherwise the
// pair is walked from y's side (the first pos(y) entries of N+(x) walked
// against y's table, by an item over y's in-list In+(y), kind 1) when N+(y) is
// longer than TC_HUB entries and than that prefix, else from x's side (all of
// N+(y) walked against x's table, by an item over x's out-list, kind 0). So the
// shorter walk is taken wherever y's list is a hub's; pairs among short lists
// stay on x's side, where each table is staged once and no partner needs an
// atomic credit (on graphs of even, small degrees splitting them as well
// doubled the staging and the launches). With whole out-lists walked on y's
// side and the rule d+(x) < d+(y), a hub threshold of 128 measured 21.8 ms at
// s22, 256 20.4 ms, 512 21.2 ms and 1024 33.4 ms; with the prefix walks and
// the rule above, 64 measured 21.5 ms, 128 21.1 ms, 256 18.6 ms and 512 20.1 ms.
constexpr uint32_t TC_HUB = 256;
// Kind 1 when y's list is a hub's and longer than the prefix of N+(x) below y.
__device__ __forceinline__ bool tc_from_owner(uint32_t d_y, uint32_t pos) { return d_y > TC_HUB && pos < d_y; }
// Walk length of the pair at position pos of N+(x) on x's side (kind 0).
__device__ __forceinline__ uint32_t tc_walk_out(uint32_t pos, uint32_t d_y) {
    return pos >= 1 && !tc_from_owner(d_y, pos) ? d_y : 0u;
}
// Walk length on y's side (kind 1).
__device__ __forceinline__ uint32_t tc_walk_in(uint32_t pos, uint32_t d_y) {
    return pos >= 1 && tc_from_owner(d_y, pos) ? pos : 0u;
}

// Walk length of every out-list slot (kind 0).
__global__ void k_tc_cost_out(const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ otgt,
                              const uint32_t* __restrict__ oown, uint64_t kept, uint32_t* __restrict__ cost) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < kept; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = oown[k], y = otgt[k];
        cost[k] = tc_walk_out((uint32_t)(k - ooff[x]), ooff[y + 1] - ooff[y]);
    }
}

__global__ void k_tc_iota(uint64_t kept, uint32_t* __restrict__ out) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < kept; k += (uint64_t)gridDim.x * blockDim.x)
        out[k] = (uint32_t)k;
}

// In-list slot i holds the out-slot islot[i] of its pair x -> y: the partner
// x, and the walk length on y's side (kind 1), the entries of N+(x) before y.
__global__ void k_tc_in_lists(const uint32_t* __restrict__ islot, const uint32_t* __restrict__ oown,
                              const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ otgt, uint64_t kept,
                              uint32_t* __restrict__ isrc, uint32_t* __restrict__ iwalk) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < kept; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = islot[i];
        const uint32_t x = oown[k], y = otgt[k];
        isrc[i] = x;
        iwalk[i] = tc_walk_in(k - ooff[x], ooff[y + 1] - ooff[y]);
    }
}

// Item starts: the first slot of every list with any work, and every slot
// whose exclusive cost prefix enters a new TC_CHUNK window. `loff` holds the
// list offsets (out- or in-lists), `ooff` the out-list offsets that size the
// owner's table.
__global__ void k_tc_item_flags(const uint32_t* __restrict__ loff, const uint32_t* __restrict__ ooff,
                                const uint32_t* __restrict__ owner, const uint32_t* __restrict__ cost,
                                const uint64_t* __restrict__ pre, uint64_t kept, uint8_t* __restrict__ flag,
                                uint32_t* __restrict__ iota) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < kept; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = owner[k];
        const uint32_t a = loff[o], e = loff[o + 1], du = ooff[o + 1] - ooff[o];
        // Lists staged in the mid-size workgroup table are cut 8x coarser (about
        // 1 M steps), so their larger staging and count flush are paid less
        // often, while the longest lists still spread over several workgroups
        // instead of finishing the launch alone (at s22: 30.8 ms against 32.5
        // uncut, 31.8 at 32x and 37.8 at 2x). The few lists of the large table
        // and those read from HBM (no staging) keep the plain window, so that
        // their launches are not left to a handful of long items.
        const uint64_t win = du > STAGE_CAP && du <= MID_CAP ? 8 * TC_CHUNK : TC_CHUNK;
        const bool work = pre[e - 1] + cost[e - 1] > pre[a];
        flag[k] = work && ((k == a) || (pre[k] / win != pre[k - 1] / win) || ((k - a) % TC_MAX_SLOTS == 0));
        iota[k] = (uint32_t)k;
    }
}

// Per item: owner, end slot (the next item's start, cut at the list end) and
// class flags: staged per wave (|N+(owner)| <= STAGE_CAP) or read from HBM
// (> BIG_CAP); per workgroup in a bitmap when the owner's new id is below
// BM_BITS; else per workgroup in a hash table for up to MID_CAP entries, or in
// one for up to BIG_CAP. The smaller table leaves room for more workgroups per
// CU.
__global__ void k_tc_item_meta(const uint32_t* __restrict__ istart, const uint32_t* __restrict__ owner,
                               const uint32_t* __restrict__ loff, const uint32_t* __restrict__ ooff,
                               uint64_t nitems, uint64_t kept, uint32_t* __restrict__ iu,
                               uint32_t* __restrict__ iend, uint8_t* __restrict__ mid, uint8_t* __restrict__ big,
                               uint8_t* __restrict__ small, uint8_t* __restrict__ bm, uint32_t bm_limit) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nitems; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = owner[istart[i]];
        const uint32_t next = i + 1 < nitems ? istart[i + 1] : (uint32_t)kept;
        const uint32_t du = ooff[o + 1] - ooff[o];
        const bool group = du > STAGE_CAP && du <= BIG_CAP;
        const bool bitmap = group && o < bm_limit;
        iu[i] = o;
        iend[i] = min(next, loff[o + 1]);
        bm[i] = bitmap;
        mid[i] = !bitmap && du > STAGE_CAP && du <= MID_CAP;
        big[i] = !bitmap && du > MID_CAP && du <= BIG_CAP;
        small[i] = !group;
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

// Bitmap table lookup: w's index in the sorted N+(u), or -1. `lim` is the
// bitmap's bit count (32 per staged word); bits at or above u are zero, since
// N+(u) lies in [0, u).
__device__ __forceinline__ int tc_bm_find(const uint32_t* bm, const uint16_t* pre, uint32_t lim, uint32_t w) {
    if (w >= lim) return -1;
    const uint32_t word = bm[w >> 5];
    const uint32_t b = w & 31u;
    if (!((word >> b) & 1u)) return -1;
    return (int)pre[w >> 5] + __popc(word & ((1u << b) - 1u));
}

// A staged table's lookup, bitmap (BM) or hash: `key` holds the bitmap words
// or the hash keys, `mask` the bit limit or the bucket mask. The returned index
// (a hash slot, or the member's position in the list) selects a 16-bit hit
// counter.
template <bool BM>
__device__ __forceinline__ int tc_find(const uint32_t* key, const uint16_t* pre, uint32_t mask, uint32_t w) {
    if constexpr (BM)
        return tc_bm_find(key, pre, mask, w);
    else
        return tc_probe(key, mask, w);
}

// Hit counters are 16-bit, two per LDS word (an entry is credited at most once
// per v of the item plus its own hits, both at most BIG_CAP, below 2^16).
__device__ __forceinline__ void tc_cnt_add(uint32_t* cnt, int slot, uint32_t k) {
    atomicAdd(&cnt[slot >> 1], k << (16 * (slot & 1)));
}

// Records a triangle (u, v, w) found with w at table index wslot (staged) or
// w itself (not staged): w's credit; v's credit and u's are added by the caller.
template <bool BM>
__device__ __forceinline__ void tc_credit_w(bool staged, int wslot, uint32_t w, uint32_t* cnt,
                                            unsigned long long* __restrict__ tri) {
    if (staged)
        tc_cnt_add(cnt, wslot, 1u);
    else
        atomicAdd(&tri[w], 1ull);
}

// Is w in N+(u)? Sets wslot when staged. (A bloom bitmap in front of the
// hash table cannot pay on a wave: the table is read whenever any of its 64
// lanes passes, which with about one hit in ten steps is nearly always.)
template <bool BM>
__device__ __forceinline__ bool tc_member(bool staged, const uint32_t* key, const uint16_t* pre, uint32_t hmask,
                                          const uint32_t* ulist, uint32_t du, uint32_t w, int& wslot) {
    if (staged) {
        wslot = tc_find<BM>(key, pre, hmask, w);
        return wslot >= 0;
    }
    return sorted_contains(ulist, du, w);
}

// The intersections of one work item, slots [s0, s1) of a list of the
// owner u (its out-list N+(u), or with IN its in-list In+(u); `plist` is the
// concatenated lists of that kind), by one wave that takes every `step`-th
// group of 64 slots starting at s0 + first. With `staged`, N+(u) sits in an
// LDS open-addressing hash set whose slots also carry hit counters; otherwise
// membership is a binary search of N+(u) in HBM. For each partner v of the
// range whose pair falls to this side, the wave walks N+(v) from HBM — all of
// it (out-lists, tc_walk_out), or the entries before u (in-lists: `iwalk`,
// tc_walk_in) — and tests each entry against N+(u). A v whose walk has at least TC_LONG steps
// is walked by the whole wave in 64-step chunks; the shorter walks of the group
// are flattened over the 64 lanes with a prefix sum and an owner search.
// A hit w closes a triangle of u, v and w: the returned count (wave-uniform) is
// u's share, w (in N+(u)) is credited in the slot counters (or in HBM when N+(u)
// is not staged), and so is v when v is in N+(u) (out-lists); an in-list
// partner is credited in HBM. Integer sums are order-independent.
template <bool IN, bool BM = false>
__device__ __forceinline__ unsigned long long tc_item(const uint32_t* __restrict__ ooff,
                                                      const uint32_t* __restrict__ otgt,
                                                      const uint32_t* __restrict__ plist,
                                                      const uint32_t* __restrict__ iwalk, uint32_t a, uint32_t du,
                                                      uint32_t s0, uint32_t s1, uint32_t first, uint32_t step,
                                                      bool staged, const uint32_t* key, uint32_t* cnt,
                                                      uint32_t hmask,
                                                      unsigned long long* __restrict__ tri,
                                                      const uint16_t* pre = nullptr) {
    const int lane = lane_id();
    const uint32_t* ulist = otgt + a;
    unsigned long long cu = 0;
    // Partner credits: the slot counter of v in N+(u), or v's count in HBM.
    const bool vtable = staged && !IN;
    for (uint32_t i0 = s0 + first; i0 < s1; i0 += step) {
        const uint32_t i = i0 + lane;
        uint32_t v = 0, sv = 0, c = 0;
        int vslot = -1;
        if (i < s1) {
            v = plist[i];
            sv = ooff[v];
            c = IN ? iwalk[i] : tc_walk_out(i - a, ooff[v + 1] - sv);
            if (vtable && c) vslot = tc_find<BM>(key, pre, hmask, v);
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
            const uint32_t* walk = otgt + __builtin_amdgcn_readlane(sv, j);
            const uint32_t e0 = (ch0 - __builtin_amdgcn_readlane(che, j)) * 64 + lane;
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
                if (staged) {
                    wslot = tc_find<BM>(key, pre, hmask, w[r]);
                    hit = ok && wslot >= 0;
                } else {
                    hit = ok && sorted_contains(ulist, du, w[r]);
                }
                if (hit) tc_credit_w<BM>(staged, wslot, w[r], cnt, tri);
                h += (uint32_t)__popcll(__ballot(hit));
            }
            cu += h;
            if (lane == 0 && h) {
                if (vtable)
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
            uint32_t w[TC_ILP], ovv[TC_ILP];
            int ovs[TC_ILP];
            bool ok[TC_ILP];
#pragma unroll
            for (int r = 0; r < TC_ILP; ++r) {
                const uint32_t t = t0 + 64 * r + lane;
                const int j = wave_owner(excl, t);
                const uint32_t e = t - __shfl(excl, j, 64);
                ovv[r] = __shfl(v, j, 64);
                ovs[r] = __shfl(vslot, j, 64);
                const uint32_t osv = __shfl(sv, j, 64);
                ok[r] = t < total;
                w[r] = 0;
                if (ok[r]) w[r] = otgt[osv + e];
            }
#pragma unroll
            for (int r = 0; r < TC_ILP; ++r) {
                bool hit = false;
                int wslot = -1;
                if (ok[r]) hit = tc_member<BM>(staged, key, pre, hmask, ulist, du, w[r], wslot);
                if (hit) {
                    tc_credit_w<BM>(staged, wslot, w[r], cnt, tri);
                    if (vtable)
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

// One wavefront per item whose owner's out-list has at most STAGE_CAP entries
// (staged in the wave's LDS) or more than BIG_CAP (read from HBM). Every wave
// owns its table and walks its own items, never waiting for the other waves.
// `plist`: the out-lists (otgt) or, with IN, the in-lists.
template <bool IN>
__global__ __launch_bounds__(TCB, 8) void k_tc_count(const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ otgt,
                                                  const uint32_t* __restrict__ plist,
                                                  const uint32_t* __restrict__ iwalk,
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
        const bool staged = du <= STAGE_CAP;
        const uint32_t hmask = tc_stage<true>(otgt, a, staged ? du : 0, lane, 64, key, cnt);
        const unsigned long long cu =
            tc_item<IN>(ooff, otgt, plist, iwalk, a, du, istart[wi], iend[wi], 0, 64, staged, key, cnt, hmask, tri);
        if (lane == 0 && cu) atomicAdd(&tri[u], cu);
        tc_flush<true>(key, cnt, (hmask + 1) * TC_BUCKET, lane, 64, tri);
    }
}

// One workgroup per item whose owner's out-list has more than STAGE_CAP and at
// most CAP entries (MID_CAP or BIG_CAP, the item classes above): the list is
// staged once in the group's LDS and the group's TCB_BIG / 64 waves share the
// item's slots.
template <uint32_t CAP, bool IN>
__global__ __launch_bounds__(TCB_BIG) void k_tc_count_big(const uint32_t* __restrict__ ooff,
                                                             const uint32_t* __restrict__ otgt,
                                                             const uint32_t* __restrict__ plist,
                                                             const uint32_t* __restrict__ iwalk,
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
        const uint32_t hmask = tc_stage<false>(otgt, a, du, threadIdx.x, TCB_BIG, b_key, b_cnt);
        const unsigned long long cu = tc_item<IN>(ooff, otgt, plist, iwalk, a, du, istart[wi], iend[wi], 64u * wid,
                                                  TCB_BIG, true, b_key, b_cnt, hmask, tri);
        if (lane == 0 && cu) atomicAdd(&tri[u], cu);
        tc_flush<false>(b_key, b_cnt, (hmask + 1) * TC_BUCKET, threadIdx.x, TCB_BIG, tri);
    }
}

// One workgroup per item whose owner u (new id below BM_BITS) has more than
// STAGE_CAP and at most BIG_CAP out-entries: N+(u), which lies in [0, u), is
// staged as a bitmap of ceil(u / 32) words with the number of members before
// each word, and one u32 hit counter per member in list order (the list is
// sorted, so a member's index is its prefix count plus the set bits below it
// in its word), two 16-bit counters per LDS word.
template <bool IN>
__global__ __launch_bounds__(TCB_BIG) void k_tc_count_bm(const uint32_t* __restrict__ ooff,
                                                        const uint32_t* __restrict__ otgt,
                                                        const uint32_t* __restrict__ plist,
                                                        const uint32_t* __restrict__ iwalk,
                                                        const uint32_t* __restrict__ istart,
                                                        const uint32_t* __restrict__ iend,
                                                        const uint32_t* __restrict__ iu, uint64_t nwork,
                                                        unsigned long long* __restrict__ tri) {
    constexpr int WPT = BM_WORDS / TCB_BIG;  // bitmap words per thread in the prefix pass
    __shared__ uint32_t s_bm[BM_WORDS];
    __shared__ uint16_t s_pre[BM_WORDS];
    __shared__ uint32_t s_cnt[BIG_CAP / 2];
    __shared__ uint32_t s_wave[TCB_BIG / 64];
    const int wid = threadIdx.x >> 6;
    const int lane = lane_id();
    const uint32_t tid = threadIdx.x;
    for (uint64_t wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        const uint32_t u = iu[wi];
        const uint32_t a = ooff[u];
        const uint32_t du = ooff[u + 1] - a;
        const uint32_t nw = (u + 31) / 32;  // u < BM_BITS, so nw <= BM_WORDS
        for (uint32_t i = tid; i < nw; i += TCB_BIG) s_bm[i] = 0;
        for (uint32_t i = tid; i < (du + 1) / 2; i += TCB_BIG) s_cnt[i] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < du; i += TCB_BIG) {
            const uint32_t x = otgt[a + i];  // < u by the numbering; the test only guards LDS
            if (x < nw * 32u) atomicOr(&s_bm[x >> 5], 1u << (x & 31u));
        }
        __syncthreads();
        // Prefix counts: thread t sums words [WPT t, WPT t + WPT), a block
        // exclusive scan gives its start.
        uint32_t loc = 0;
#pragma unroll
        for (int r = 0; r < WPT; ++r) {
            const uint32_t wd = WPT * tid + r;
            if (wd < nw) loc += __popc(s_bm[wd]);
        }
        const uint32_t incl = wave_incl_scan(loc);
        if (lane == 63) s_wave[wid] = incl;
        __syncthreads();
        uint32_t run = incl - loc;
        for (int w = 0; w < wid; ++w) run += s_wave[w];
#pragma unroll
        for (int r = 0; r < WPT; ++r) {
            const uint32_t wd = WPT * tid + r;
            if (wd < nw) {
                s_pre[wd] = (uint16_t)run;
                run += __popc(s_bm[wd]);
            }
        }
        __syncthreads();
        const unsigned long long cu = tc_item<IN, true>(ooff, otgt, plist, iwalk, a, du, istart[wi], iend[wi], 64u * wid,
                                                        TCB_BIG, true, s_bm, s_cnt, nw * 32u, tri, s_pre);
        if (lane == 0 && cu) atomicAdd(&tri[u], cu);
        __syncthreads();
        for (uint32_t i = tid; i < du; i += TCB_BIG) {
            const uint32_t c = (s_cnt[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
            if (c) atomicAdd(&tri[otgt[a + i]], (unsigned long long)c);
        }
        __syncthreads();
    }
}

// Counts per new id -> counts per caller id.
__global__ void k_tc_unmap(const unsigned long long* __restrict__ tnew, const uint32_t* __restrict__ newid,
                           uint64_t n, unsigned long long* __restrict__ out) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = tnew[newid[v]];
}

// Rank keys: ascending order of (~degree, ~id) is decreasing (degree, id).
__global__ void k_tc_rank_keys(const uint32_t* __restrict__ uoff, uint64_t n, uint64_t* __restrict__ keys,
                               uint32_t* __restrict__ iota) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = uoff[v + 1] - uoff[v];
        keys[v] = ((uint64_t)(~d) << 32) | (uint64_t)(~(uint32_t)v);
        iota[v] = (uint32_t)v;
    }
}

__global__ void k_tc_newid(const uint32_t* __restrict__ old_of_new, uint64_t n, uint32_t* __restrict__ newid) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        newid[old_of_new[i]] = (uint32_t)i;
}

// Oriented (owner, target) pairs in new ids, packed owner-major for the sort.
__global__ void k_tc_pack(const uint32_t* __restrict__ own, const uint32_t* __restrict__ tgt, uint64_t m,
                          const uint32_t* __restrict__ newid, uint64_t* __restrict__ keys) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x)
        keys[k] = ((uint64_t)newid[own[k]] << 32) | newid[tgt[k]];
}

__global__ void k_tc_unpack(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ own,
                            uint32_t* __restrict__ tgt) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
        own[k] = (uint32_t)(keys[k] >> 32);
        tgt[k] = (uint32_t)keys[k];
    }
}

}  // namespace

std::atomic<bool> g_tc_bitmaps{true};

static void tc_orient(ursa_hip_topo* t, hipStream_t s) {
    NoArena keep_out_of_arena;
    std::lock_guard<std::mutex> lk(t->cache_mu);
    if (t->tc_built) return;
    const uint64_t n = t->n;
    const DevCsr& U = t->undirected(s);
    const uint64_t slots = U.slots;
    t->tc_off.alloc(n + 1);
    // New ids: vertices in decreasing (degree, id) order, so that every
    // oriented edge points to a smaller new id.
    t->tc_new.alloc(n ? n : 1);
    if (n) {
        DevBuf<uint64_t> rk(n), rks(n);
        DevBuf<uint32_t> iota(n), old_of_new(n);
        k_tc_rank_keys<<<blocks_for(n, 256), 256, 0, s>>>(U.off.get(), n, rk.get(), iota.get());
        check_launch();
        prim::sort_pairs_u64_u32(rk.get(), rks.get(), iota.get(), old_of_new.get(), n, 64, s);
        k_tc_newid<<<blocks_for(n, 256), 256, 0, s>>>(old_of_new.get(), n, t->tc_new.get());
        check_launch();
    }
    uint64_t kept = 0;
    DevBuf<uint32_t> oown;  // owner of each oriented slot
    if (slots) {
        DevBuf<uint32_t> owner(slots);
        slot_owners(n, U, owner.get(), s);
        DevBuf<uint8_t> keep(slots);
        k_tc_flags<<<blocks_for(slots, 256), 256, 0, s>>>(U.off.get(), U.tgt.get(), owner.get(), slots, keep.get());
        check_launch();
        DevBuf<uint32_t> ktgt(slots), kown(slots);
        DevBuf<uint64_t> cnt(1);
        prim::select_flagged_u32(U.tgt.get(), keep.get(), ktgt.get(), cnt.get(), slots, s);
        prim::select_flagged_u32(owner.get(), keep.get(), kown.get(), cnt.get(), slots, s);
        HIP_OK(hipMemcpyAsync(&kept, cnt.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        // Relabel both ends and sort owner-major: the out-lists in new ids,
        // each sorted by new id.
        oown.alloc(kept ? kept : 1);
        t->tc_tgt.alloc(kept ? kept : 1);
        if (kept) {
            DevBuf<uint64_t> pk(kept), pks(kept);
            k_tc_pack<<<blocks_for(kept, 256), 256, 0, s>>>(kown.get(), ktgt.get(), kept, t->tc_new.get(), pk.get());
            check_launch();
            prim::sort_keys_u64(pk.get(), pks.get(), kept, 32 + bits_for(n), s);
            k_tc_unpack<<<blocks_for(kept, 256), 256, 0, s>>>(pks.get(), kept, oown.get(), t->tc_tgt.get());
            check_launch();
        }
        offsets_from_sorted(oown.get(), kept, n, t->tc_off.get(), s);
    } else {
        t->tc_tgt.alloc(1);
        offsets_from_sorted(nullptr, 0, n, t->tc_off.get(), s);
    }
    // In-lists of the DAG: the kept out-slots regrouped by target with a
    // stable sort, so each In+(v) lists its sources in increasing order; each
    // in-slot keeps its source and its walk length on the target's side.
    DevBuf<uint32_t> iown;  // target of each in-list slot
    t->tc_ioff.alloc(n + 1);
    if (kept) {
        iown.alloc(kept);
        t->tc_isrc.alloc(kept);
        t->tc_iwalk.alloc(kept);
        int bits = 1;
        while (bits < 32 && (1ull << bits) < n) ++bits;
        {
            DevBuf<uint32_t> slot(kept), islot(kept);
            k_tc_iota<<<blocks_for(kept, 256), 256, 0, s>>>(kept, slot.get());
            check_launch();
            prim::sort_pairs_u32(t->tc_tgt.get(), iown.get(), slot.get(), islot.get(), kept, bits, s);
            k_tc_in_lists<<<blocks_for(kept, 256), 256, 0, s>>>(islot.get(), oown.get(), t->tc_off.get(),
                                                               t->tc_tgt.get(), kept, t->tc_isrc.get(),
                                                               t->tc_iwalk.get());
            check_launch();
        }
        offsets_from_sorted(iown.get(), kept, n, t->tc_ioff.get(), s);
    } else {
        t->tc_isrc.alloc(1);
        t->tc_iwalk.alloc(1);
        offsets_from_sorted(nullptr, 0, n, t->tc_ioff.get(), s);
    }
    // Work items (see the file comment), for out-lists ([0]) and in-lists ([1]).
    uint64_t walk_steps = 0;  // list entries walked against a table, both kinds
    for (int kind = 0; kind < 2; ++kind) {
        auto& items = t->tc_items[kind];
        const uint32_t* loff = kind ? t->tc_ioff.get() : t->tc_off.get();
        const uint32_t* owner = kind ? iown.get() : oown.get();
        for (auto& it : items) {
            it.start.alloc(1);
            it.end.alloc(1);
            it.owner.alloc(1);
            it.n = 0;
        }
        if (!kept) continue;
        DevBuf<uint32_t> cost(kind ? 1 : kept), iota(kept);
        DevBuf<uint64_t> pre(kept + 1);
        DevBuf<uint8_t> flag(kept);
        if (kind == 0) {
            k_tc_cost_out<<<blocks_for(kept, 256), 256, 0, s>>>(t->tc_off.get(), t->tc_tgt.get(), owner, kept,
                                                               cost.get());
            check_launch();
        }
        const uint32_t* costp = kind ? t->tc_iwalk.get() : cost.get();
        prim::exclusive_sum_u32_to_u64(costp, pre.get(), kept, s);
        {
            uint64_t last_pre = 0;
            uint32_t last_cost = 0;
            HIP_OK(hipMemcpyAsync(&last_pre, pre.get() + (kept - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
            HIP_OK(hipMemcpyAsync(&last_cost, costp + (kept - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            walk_steps += last_pre + last_cost;
        }
        k_tc_item_flags<<<blocks_for(kept, 256), 256, 0, s>>>(loff, t->tc_off.get(), owner, costp, pre.get(),
                                                             kept, flag.get(), iota.get());
        check_launch();
        DevBuf<uint32_t> istart(kept), iu(kept), iend(kept);
        DevBuf<uint64_t> wc(1);
        prim::select_flagged_u32(iota.get(), flag.get(), istart.get(), wc.get(), kept, s);
        uint64_t nitems = 0;
        HIP_OK(hipMemcpyAsync(&nitems, wc.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        if (!nitems) continue;
        DevBuf<uint8_t> cls[4] = {DevBuf<uint8_t>(nitems), DevBuf<uint8_t>(nitems), DevBuf<uint8_t>(nitems),
                                  DevBuf<uint8_t>(nitems)};
        k_tc_item_meta<<<blocks_for(nitems, 256), 256, 0, s>>>(istart.get(), owner, loff, t->tc_off.get(), nitems,
                                                              kept, iu.get(), iend.get(), cls[1].get(), cls[2].get(),
                                                              cls[0].get(), cls[3].get(),
                                                              g_tc_bitmaps.load() ? BM_BITS : 0u);
        check_launch();
        // Split the items into the per-wave class and the three per-workgroup classes.
        for (int c = 0; c < 4; ++c) {
            auto& it = items[c];
            it.start.alloc(nitems);
            it.end.alloc(nitems);
            it.owner.alloc(nitems);
            prim::select_flagged_u32(istart.get(), cls[c].get(), it.start.get(), wc.get(), nitems, s);
            prim::select_flagged_u32(iend.get(), cls[c].get(), it.end.get(), wc.get(), nitems, s);
            prim::select_flagged_u32(iu.get(), cls[c].get(), it.owner.get(), wc.get(), nitems, s);
            HIP_OK(hipMemcpyAsync(&it.n, wc.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
        }
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
    t->tc_pair_work = (double)hp;
    t->tc_walk_steps = walk_steps;
    t->tc_built = true;
}

void triangle_prepare(ursa_hip_topo* t) { tc_orient(t, stream()); }

// Device temporaries of one triangle_count call on a prepared topology with a
// device output: the counts per new id (the arena aligns each block to 256 B).
uint64_t triangle_workspace_bytes(uint64_t n) { return n * sizeof(uint64_t) + 4096; }

void triangle_count(ursa_hip_topo* t, uint64_t* out_user) {
    const uint64_t n = t->n;
    if (n == 0) return;  // triangle.rs:59-61
    hipStream_t s = stream();
    tc_orient(t, s);
    DevOut<uint64_t> out(out_user, n);
    DevBuf<unsigned long long> tnew(n);  // counts per new id
    HIP_OK(hipMemsetAsync(tnew.get(), 0, n * sizeof(uint64_t), s));
    uint64_t nitems = 0;
    for (const auto& kind : t->tc_items)
        for (const auto& it : kind) nitems += it.n;
    if (nitems) {
        // Bytes the one-sided walk must read: each walked entry (4), each
        // oriented pair's partner id and offsets (12), per node 12.
        const double bytes = 4.0 * (double)t->tc_walk_steps + 12.0 * (double)t->tc_slots + 12.0 * (double)n;
        auto* tri = tnew.get();
        const uint32_t* ooff = t->tc_off.get();
        const uint32_t* otgt = t->tc_tgt.get();
        timed("triangle_count", bytes, s, [&] {
            // In-list items run on the side stream beside the out-list items, so
            // that each launch's last items overlap the other kind's work.
            const auto& in_items = t->tc_items[1];
            const bool side = in_items[0].n + in_items[1].n + in_items[2].n + in_items[3].n > 0;
            hipStream_t s1 = s;
            if (side) fork_side(s, &s1);
            for (int kind = 0; kind < 2; ++kind) {
                const auto& it = t->tc_items[kind];
                const hipStream_t ks = kind ? s1 : s;
                const uint32_t* plist = kind ? t->tc_isrc.get() : otgt;
                const uint32_t* iwalk = kind ? t->tc_iwalk.get() : nullptr;
                auto grid_big = [](uint64_t k) { return (unsigned)std::min<uint64_t>(k, 1u << 16); };
                const unsigned grid_small = (unsigned)std::min<uint64_t>((it[0].n + TCW - 1) / TCW, TC_SMALL_GRID);
                if (it[3].n) {
                    if (kind == 0)
                        k_tc_count_bm<false><<<grid_big(it[3].n), TCB_BIG, 0, ks>>>(
                            ooff, otgt, plist, iwalk, it[3].start.get(), it[3].end.get(), it[3].owner.get(), it[3].n, tri);
                    else
                        k_tc_count_bm<true><<<grid_big(it[3].n), TCB_BIG, 0, ks>>>(
                            ooff, otgt, plist, iwalk, it[3].start.get(), it[3].end.get(), it[3].owner.get(), it[3].n, tri);
                }
                if (kind == 0) {
                    if (it[2].n)
                        k_tc_count_big<BIG_CAP, false><<<grid_big(it[2].n), TCB_BIG, 0, ks>>>(
                            ooff, otgt, plist, iwalk, it[2].start.get(), it[2].end.get(), it[2].owner.get(), it[2].n, tri);
                    if (it[1].n)
                        k_tc_count_big<MID_CAP, false><<<grid_big(it[1].n), TCB_BIG, 0, ks>>>(
                            ooff, otgt, plist, iwalk, it[1].start.get(), it[1].end.get(), it[1].owner.get(), it[1].n, tri);
                    if (it[0].n)
                        k_tc_count<false><<<grid_small, TCB, 0, ks>>>(ooff, otgt, plist, iwalk, it[0].start.get(),
                                                                      it[0].end.get(), it[0].owner.get(), it[0].n, tri);
                } else {
                    if (it[2].n)
                        k_tc_count_big<BIG_CAP, true><<<grid_big(it[2].n), TCB_BIG, 0, ks>>>(
                            ooff, otgt, plist, iwalk, it[2].start.get(), it[2].end.get(), it[2].owner.get(), it[2].n, tri);
                    if (it[1].n)
                        k_tc_count_big<MID_CAP, true><<<grid_big(it[1].n), TCB_BIG, 0, ks>>>(
                            ooff, otgt, plist, iwalk, it[1].start.get(), it[1].end.get(), it[1].owner.get(), it[1].n, tri);
                    if (it[0].n)
                        k_tc_count<true><<<grid_small, TCB, 0, ks>>>(ooff, otgt, plist, iwalk, it[0].start.get(),
                                                                     it[0].end.get(), it[0].owner.get(), it[0].n, tri);
                }
            }
            if (side) join_side(s);
        });
        check_launch();
    }
    k_tc_unmap<<<blocks_for(n, 256), 256, 0, s>>>(tnew.get(), t->tc_new.get(), n,
                                                  reinterpret_cast<unsigned long long*>(out.p));
    check_launch();
    out.finish(s);
}

}  // namespace ursa

extern "C" int ursa_hip_triangle_bitmaps(int on) {
    ursa::g_tc_bitmaps.store(on != 0);
    return URSA_HIP_OK;
}

extern "C" int ursa_hip_triangle_work(ursa_hip_topo* t, uint64_t* oriented_edges, uint64_t* merge_steps,
                                      uint64_t* walk_steps) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        ursa::triangle_prepare(t);
        if (oriented_edges) *oriented_edges = t->tc_slots;
        if (merge_steps) *merge_steps = (uint64_t)t->tc_pair_work;
        if (walk_steps) *walk_steps = t->tc_walk_steps;
    });
}

extern "C" int ursa_hip_triangle_count(ursa_hip_topo* t, uint64_t* out) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (t->n == 0) return;
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        ursa::triangle_count(t, out);
    });
}
