// ppr.hpp — personalized PageRank solver shared by the C ABI and the request
// server (definition and determinism rules in ppr.hip).
#pragma once

#include "common.hpp"

namespace ursa {

constexpr int PPR_MAX_LANES = 16;

struct PprQuery {
    const uint32_t* seeds;  // host memory
    uint32_t nseeds;
    double damping;
    uint32_t max_iter;
    double tol;
};

// PageRank's working copy of the in-CSR (cached on the topology; pagerank.hip).
void pagerank_prepare(ursa_hip_topo* t, hipStream_t s);
// Builds the graph-only heavy-row segmentation over that copy (cached on the
// topology).
void ppr_prepare(ursa_hip_topo* t, hipStream_t s);
// Device bytes one solve of k lanes allocates (outputs staged when host_out).
uint64_t ppr_workspace_bytes(const ursa_hip_topo* t, int k, bool host_out);
// Solves k (1..16) queries on t together; outs[q] is f64[n] in host or device
// memory; iters_out[q] receives each query's iteration count.
void ppr_solve(ursa_hip_topo* t, const PprQuery* qs, int k, double* const* outs, uint32_t* iters_out);

}  // namespace ursa
