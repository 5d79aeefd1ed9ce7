// traverse.hpp — order-exact k-hop and shortest-path traversals run one after
// another in one workspace (used by the C ABI and the request server).
#pragma once

#include <vector>

#include "common.hpp"

namespace ursa {

// Device bytes one sequence of traversals on an n-node graph allocates.
uint64_t traverse_workspace_bytes(uint64_t n);
// k_hop for each seed in turn (hop.rs:52-130): appends the reached nodes of
// seed i to dst and sets ends[i] to dst's size after it.
void k_hop_seq(ursa_hip_topo* t, const uint32_t* seeds, uint64_t nseeds, uint32_t k, int dir,
               std::vector<uint32_t>& dst, std::vector<uint64_t>& ends);
// shortest_path (bfs.rs:48-102) for each (src[i], tgt[i]); an empty path is None.
void shortest_paths_seq(ursa_hip_topo* t, const uint32_t* src, const uint32_t* tgt, uint64_t count, int dir,
                        std::vector<std::vector<uint32_t>>& paths);

// Cross-graph passes: each task is one k-hop query (seeds in turn, depth k)
// or one shortest path (seeds = {source, target}) on a graph of at most
// TRAV_MULTI_MAX_N nodes; one launch runs a whole pass, one workgroup per task,
// and every result equals k_hop_seq / shortest_paths_seq on the same query.
constexpr uint64_t TRAV_MULTI_MAX_N = 1u << 16;
struct TravTask {
    ursa_hip_topo* t = nullptr;
    bool path = false;
    std::vector<uint32_t> seeds;
    uint32_t k = 0;
    int dir = 0;
    std::vector<uint32_t> dst;   // k-hop: reached nodes, seed after seed
    std::vector<uint64_t> ends;  // k-hop: dst's size after each seed
    std::vector<uint32_t> found; // path: source..target (empty = None)
};
bool traverse_multi_eligible(const ursa_hip_topo* t);
uint64_t traverse_multi_workspace_bytes(const std::vector<const TravTask*>& tasks);
void traverse_multi(const std::vector<TravTask*>& tasks);

}  // namespace ursa
