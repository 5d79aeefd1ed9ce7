// frontier.hpp — device helpers shared by the frontier kernels (BFS, k-hop,
// shortest path, Brandes): the directional neighbour view and wave-level
// scan / ballot utilities for 64-lane wavefronts.
#pragma once

#include "topology.hpp"

namespace ursa {

// A vertex's neighbour list in one traversal direction: one CSR, or two CSRs
// concatenated (out then in) for Direction::Both — the walk of
// Topology::for_each_neighbor (topology.rs:395-404).
struct NbrView {
    const uint32_t* off_a;
    const uint32_t* tgt_a;
    const uint32_t* off_b;  // null unless two lists are concatenated
    const uint32_t* tgt_b;
};

__device__ __forceinline__ uint32_t nv_degree(const NbrView& g, uint32_t u, uint32_t& deg_a) {
    deg_a = g.off_a[u + 1] - g.off_a[u];
    uint32_t d = deg_a;
    if (g.off_b) d += g.off_b[u + 1] - g.off_b[u];
    return d;
}

__device__ __forceinline__ uint32_t nv_degree(const NbrView& g, uint32_t u) {
    uint32_t da;
    return nv_degree(g, u, da);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = lane_id();
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;  // valid in lane 0
}

// Largest lane j with excl[j] <= t (excl non-decreasing across lanes, excl[0] = 0).
__device__ __forceinline__ int wave_owner(uint32_t excl, uint32_t t) {
    int j = 0;
    for (int step = 32; step > 0; step >>= 1) {
        const int cand = j + step;
        const uint32_t v = __shfl(excl, cand & 63, 64);
        if (cand < 64 && v <= t) j = cand;
    }
    return j;
}

// Wave-aggregated append: every lane with `take` gets a distinct slot in a
// queue whose length lives at *len; one atomic per wave.
__device__ __forceinline__ uint32_t wave_append_slot(bool take, uint32_t* len) {
    const uint64_t mask = __ballot(take);
    uint32_t base = 0;
    if (mask) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        if (lane_id() == leader) base = atomicAdd(len, (uint32_t)__popcll(mask));
        base = __shfl(base, leader, 64);
    }
    return base + (uint32_t)__popcll(mask & lanemask_lt());
}

}  // namespace ursa
