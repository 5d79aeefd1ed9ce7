// topology.hpp — the device-resident CSR topology shared by every kernel.
//
// HBM layout (all u32; m <= UINT32_MAX is enforced, topology.rs:61-64):
//   out.off[n+1], out.tgt[m], out.eid[m]   — built eagerly (Topology::build)
//   in.off[n+1],  in.tgt[m],  in.eid[m]    — lazy transpose (Topology::incoming)
//   und.off[n+1], und.tgt[slots]           — lazy undirected view (Topology::undirected)
//   out_deg[n]                             — cached with the out-CSR
// Offsets are u32 on the device (ursa-core keeps u64 on the host,
// topology.rs:40-44); every offset value is <= m, so u32 is lossless.
#pragma once

#include <mutex>

#include "common.hpp"

namespace ursa {

struct DevCsr {
    DevBuf<uint32_t> off;
    DevBuf<uint32_t> tgt;
    DevBuf<uint32_t> eid;
    uint64_t slots = 0;
    bool built = false;
};

// Stable grouping of m rows by key (< n): the device counterpart of
// Adjacency::build (topology.rs:54-187). Within each key segment rows keep
// their original order, so the result is byte-identical to the CPU build.
void build_grouped(uint64_t n, uint64_t m, const uint32_t* keys, const uint32_t* other, DevCsr& out,
                   hipStream_t s);

// owner[k] = node whose segment holds slot k of `csr` (needs csr.slots > 0).
void slot_owners(uint64_t n, const DevCsr& csr, uint32_t* owner, hipStream_t s);
// The same for offsets off[n+1] over m > 0 slots.
void slot_owners(uint64_t n, uint64_t m, const uint32_t* off, uint32_t* owner, hipStream_t s);

// off[u] = first index whose key >= u (u in [0, n]) for keys sorted ascending.
void offsets_from_sorted(const uint32_t* keys, uint64_t m, uint64_t n, uint32_t* off, hipStream_t s);

}  // namespace ursa

namespace ursa {
// Traversal schedule parameters picked by the autotuner (bfs.hip). They change
// only which kernel runs when, never a result.
struct TuneParams {
    double bfs_alpha = 15.0;   // top-down -> bottom-up when frontier edges > unexplored / alpha
    double bfs_beta = 18.0;    // bottom-up -> top-down when the frontier < n / beta
    uint32_t ms_push_div = 20; // multi-source push when frontier edges * div < all edges
    int source = 0;            // 0 defaults, 1 measured on this graph, 2 taken from the key cache
};
}  // namespace ursa

struct ursa_hip_topo {
    uint64_t n = 0, m = 0;
    ursa::DevCsr out, in, und;
    ursa::DevBuf<uint32_t> out_deg;
    std::mutex build_mu;  // guards the lazy in / und builds
    std::mutex cache_mu;  // guards algorithm-side caches below
    // PageRank's private copy of the in-CSR with renumbered vertices, its
    // row-start bits and per-chunk row / flags (see pagerank.hip).
    ursa::DevBuf<uint32_t> pr_old_of_new, pr_new_of_old, pr_off, pr_tgt, pr_eid, pr_outdeg;
    ursa::DevBuf<uint32_t> pr_head, pr_crow;
    ursa::DevBuf<uint8_t> pr_cflag;
    uint64_t pr_nz = 0, pr_chunks = 0;
    bool pr_ready = false;
    uint64_t pr_max_in = ~0ull;  // largest in-degree (computed on first use)
    // PageRank's binned sweep (see pagerank.hip): the renumbered in-slots in
    // tiles (block of 8192 sources, bin of up to 4096 consecutive rows; rows
    // ascending), each padded to whole chunks of 8 slots; source-block-major:
    // each slot's source index in its block (u16) and edge id, and each
    // chunk's place in the bin-major order; bin-major: each slot's row index
    // in its bin (u16); the first row of each bin; the work items of both phases.
    ursa::DevBuf<uint16_t> bn_src, bn_dst;
    ursa::DevBuf<uint32_t> bn_eid, bn_place, bn_bstart, bn_item1, bn_item2;
    uint64_t bn_slots = 0, bn_n1 = 0, bn_n2 = 0, bn_bins = 0;
    bool bn_ready = false;
    // Multi-source BFS pull side, per direction (see msbfs.hip): the vertices
    // whose reverse lists are cut into chunks and their chunk bases; depends
    // on the graph only.
    ursa::DevBuf<uint32_t> ms_pheavy[3], ms_pcbase[3];
    uint32_t ms_pnh[3] = {0, 0, 0}, ms_pchunks[3] = {0, 0, 0};
    bool ms_pull_ready[3] = {false, false, false};
    // Degree-oriented undirected DAG for triangle counting (see triangle.hip),
    // its vertices numbered by decreasing (degree, id) (tc_new: the new id of
    // each vertex): out-lists N+(u), in-lists In+(v), and work items over
    // either list kind ([0] out, [1] in) in four classes ([0] per wave, [1] per
    // workgroup with the mid-size hash table, [2] with the large one, [3] with
    // a bitmap table).
    ursa::DevBuf<uint32_t> tc_off, tc_tgt, tc_ioff, tc_isrc, tc_new;
    ursa::DevBuf<uint32_t> tc_iwalk;  // per in-list slot x -> v: entries of N+(x) walked on v's side (before v), or 0
    struct TcItems {
        ursa::DevBuf<uint32_t> start, end, owner;
        uint64_t n = 0;
    };
    TcItems tc_items[2][4];
    uint64_t tc_slots = 0;
    double tc_pair_work = 0.0;     // sum over oriented pairs of d+(u) + d+(v) (sorted-merge steps)
    uint64_t tc_walk_steps = 0;     // entries the one-sided walks read, both item kinds
    bool tc_built = false;
    // Personalized PageRank's split of high in-degree rows into fixed
    // segments (see ppr.hip); depends on the graph only.
    ursa::DevBuf<uint32_t> ppr_hrow, ppr_hsoff, ppr_sbeg, ppr_send;
    uint64_t ppr_H = 0, ppr_S = 0, ppr_dangling = 0;
    bool ppr_ready = false;
    bool ppr_bn_auto = false;  // the automatic schedule sweeps this graph binned

    ursa::TuneParams tune[3];  // per traversal direction
    uint64_t tune_key[3] = {0, 0, 0};

    const ursa::DevCsr& incoming(hipStream_t s);
    const ursa::DevCsr& undirected(hipStream_t s);
};
