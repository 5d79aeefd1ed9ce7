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

}  // namespace ursa
