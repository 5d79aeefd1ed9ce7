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

// Cross-graph passes: one-lane solves on several graphs swept together, each
// query giving the bits ppr_solve gives it alone. A graph qualifies with at
// most PPR_MULTI_MAX_N vertices on the pull schedule (builds the caches
// ppr_prepare builds).
constexpr uint64_t PPR_MULTI_MAX_N = 1u << 19;
// Limits of one cross-graph pass: problems, and vertices summed over them.
constexpr size_t PPR_MULTI_MAX_PASS = 256;
constexpr uint64_t PPR_MULTI_MAX_NODES = 8u << 20;
bool ppr_multi_eligible(ursa_hip_topo* t);
// Device bytes one pass over these graphs allocates.
uint64_t ppr_multi_workspace_bytes(const ursa_hip_topo* const* ts, int count);
// Solves query i on ts[i] (every graph eligible); outs[i] is f64[n_i] in host
// or device memory.
void ppr_solve_multi(ursa_hip_topo* const* ts, const PprQuery* qs, int count, double* const* outs,
                     uint32_t* iters_out);
// Whether t (already prepared) is likely eligible, without building anything.
bool ppr_multi_candidate(const ursa_hip_topo* t);
// Query i on ts[i]: eligible graphs in cross-graph passes (within the limits
// above), the others one ppr_solve each.
void ppr_solve_many(ursa_hip_topo* const* ts, const PprQuery* qs, int count, double* const* outs,
                    uint32_t* iters_out);

}  // namespace ursa
