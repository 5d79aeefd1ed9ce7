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

// Counters shared by a whole grid: same-address atomics from every wave
// serialise in one L2 channel (about 10 ns each on MI355X), which costs more
// than the work of a frontier kernel over millions of vertices. These helpers
// make it one atomic per workgroup; kernels that use them also cap their grid
// (GRID_CAP workgroups, grid-stride loops), so a launch issues at most
// GRID_CAP atomics per counter.
constexpr unsigned GRID_CAP = 2048;

// Adds every thread's v into *ctr with one atomic per workgroup. Called once,
// by every thread of the block, at the end of the kernel.
template <int BLOCK, class T>
__device__ __forceinline__ void block_add(T v, T* ctr) {
    __shared__ T s_part[BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane_id() == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = 0;
        for (int w = 0; w < BLOCK / 64; ++w) t += s_part[w];
        if (t) atomicAdd(ctr, t);
    }
    __syncthreads();  // s_part may be reused by another call
}

// Queue slot for each taking thread, one atomic per workgroup and call. Every
// thread of the block must call it (block-uniform control flow); s_cnt is an
// LDS array of BLOCK/64 + 1 words owned by the caller.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_append_slot(bool take, uint32_t* len, uint32_t* s_cnt) {
    const int wave = (int)(threadIdx.x >> 6);
    const uint64_t mask = __ballot(take);
    if (lane_id() == 0) s_cnt[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            const uint32_t c = s_cnt[w];
            s_cnt[w] = tot;
            tot += c;
        }
        s_cnt[BLOCK / 64] = tot ? atomicAdd(len, tot) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s_cnt[BLOCK / 64] + s_cnt[wave] + (uint32_t)__popcll(mask & lanemask_lt());
    __syncthreads();  // s_cnt is rewritten by the next call
    return slot;
}

}  // namespace ursa
