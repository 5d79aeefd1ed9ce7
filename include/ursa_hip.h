/*
 * ursa_hip.h — C ABI of the MI355X (gfx950) backend for ursa-core's CSR
 * algorithm kernels.
 *
 * Each entry point replaces one ursa-core function (cited file:line relative
 * to the cldixon/ursa 0.2.0 tree) and keeps its argument meaning, output
 * layout and edge-case behaviour. Outputs go to caller-allocated buffers of
 * the stated length; a buffer may be host memory or device memory (device
 * pointers are written in place, with no copy back). Inputs likewise.
 *
 * Error convention: every function returns URSA_HIP_OK (0) or a negative
 * status; ursa_hip_last_error() returns a thread-local message for the last
 * failure on the calling thread. No function aborts the process.
 *
 * Threading: any number of threads may call concurrently on one topology.
 * The lazy incoming/undirected builds run once per topology (topology.rs:
 * 257-260). Each calling thread gets its own HIP stream (triangle counting
 * also uses a second one, joined before the call returns). Temporary device
 * buffers are kept for reuse on the thread's own stream rather than freed;
 * memory from ursa_hip_device_alloc is never kept.
 */
#ifndef URSA_HIP_H
#define URSA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define URSA_HIP_OK 0
#define URSA_HIP_EINVAL (-22)  /* bad argument (length mismatch, id out of range, ...) */
#define URSA_HIP_ENOMEM (-12)  /* device or host allocation failed */
#define URSA_HIP_EDEVICE (-5)  /* HIP runtime error */
#define URSA_HIP_ERANGE (-34)  /* size beyond a documented cap (edge cap, output cap) */
#define URSA_HIP_EINTERNAL (-1)

/* Direction — ursa-core/src/topology.rs:27-32 */
typedef enum { URSA_HIP_OUT = 0, URSA_HIP_IN = 1, URSA_HIP_BOTH = 2 } ursa_hip_direction;

typedef struct ursa_hip_topo ursa_hip_topo;

/* ---- lifecycle ---- */
/* Select and initialise the device (optional: the first call on a thread
 * initialises device 0). */
int ursa_hip_init(int device);
int ursa_hip_shutdown(void);
const char* ursa_hip_last_error(void);
/* Synchronise the calling thread's stream. */
int ursa_hip_synchronize(void);
/* Serving hook: run this thread's work on the caller's HIP stream (a
 * hipStream_t, passed as void*), or on the thread's own non-blocking stream
 * again with NULL. Calls still return only when their results are written. */
int ursa_hip_set_stream(void* stream);

/* ---- BFS-based statistics — ursa-plan/src/stats.rs ---- */
/* bfs_source_scan (stats.rs:100-140): over every source (out-of-range ones
 * reach nothing) and every vertex at finite distance d >= 1: *sum = sum of d
 * (f64), *count = number of such pairs, *max_dist = largest d (0 if none). */
int ursa_hip_bfs_source_scan(ursa_hip_topo* t, const uint32_t* sources, uint64_t k, int direction,
                             double* sum, uint64_t* count, int32_t* max_dist);
/* avg_path_length (stats.rs:142-152): sources = every node, or the strided
 * sample of round(sample*n) nodes (stats.rs:80-98); out-edges; 0.0 when no pair
 * is reachable. */
int ursa_hip_avg_path_length(ursa_hip_topo* t, int has_sample, double sample, double* out);
/* diameter (stats.rs:154-168): max distance from every node (approximate = 0)
 * or from the strided sample of min(128/n, 1) (approximate = 1); out-edges. */
int ursa_hip_diameter(ursa_hip_topo* t, int approximate, int64_t* out);
/* closeness (ursa-core/src/algo/closeness.rs:69-137, unweighted, out-edges):
 * reachable / sum of finite distances >= 1 per source, 0.0 if it reaches
 * nothing. sources = NULL: every node (out f64[n]); else out[i] for sources[i]
 * (f64[k]; an out-of-range source gives 0.0). */
int ursa_hip_closeness(ursa_hip_topo* t, const uint32_t* sources, uint64_t k, double* out);

/* closeness_weighted (closeness.rs:143-222): for sources[i] (all nodes in
 * order when sources is NULL; out-of-range sources give 0.0), reachable /
 * sum of the shortest-path costs over out-edges with weights[edge row]
 * (f64[m], non-negative: URSA_HIP_EINVAL otherwise), the sum taken in node
 * order as the reference takes it. Costs are exactly Dijkstra's (min over
 * paths of the left-to-right rounded path cost), so values are bit-identical. */
int ursa_hip_closeness_weighted(ursa_hip_topo* t, const double* weights, const uint32_t* sources, uint64_t k,
                                double* out);
/* clustering_coefficient (ursa-core/src/algo/clustering.rs:21-40):
 * 2 T(u) / (k (k - 1)) over the undirected view, 0.0 for degree < 2; f64[n]. */
int ursa_hip_clustering_coefficient(ursa_hip_topo* t, double* out);

/* ---- ingest — ursa-core/src/id_map.rs ---- */
/* IdMap::from_edge_batches / intern_i64 (id_map.rs:220-319) on the device:
 * dense ids in first-seen order over the rows, src before dst within a row.
 * src/dst: int64[m] (host or device); src_dense/dst_dense: u32[m];
 * user_ids (may be NULL): int64 with room for 2m entries, receives the user id
 * of each dense id (IdMap::user_id_array, id_map.rs:359-365); *n_out = number
 * of distinct ids. 2m must fit in u32 (URSA_HIP_ERANGE otherwise). */
int ursa_hip_intern_i64(uint64_t m, const int64_t* src, const int64_t* dst, uint32_t* src_dense,
                        uint32_t* dst_dense, int64_t* user_ids, uint64_t* n_out);
/* IdMap::from_edge_batches over a stream of (src, dst) batches (id_map.rs:207-292):
 * batch b has lens[b] rows in srcs[b] / dsts[b] (int64, host or device; each
 * batch may be anywhere). Rows are numbered across the batches in order, so the
 * ids are those ursa_hip_intern_i64 gives the rows laid end to end;
 * src_dense/dst_dense: u32[sum lens] in that order; user_ids and n_out as
 * above. The batches are not concatenated. */
int ursa_hip_intern_i64_batches(uint64_t nbatches, const uint64_t* lens, const int64_t* const* srcs,
                                const int64_t* const* dsts, uint32_t* src_dense, uint32_t* dst_dense,
                                int64_t* user_ids, uint64_t* n_out);

/* ---- topology — ursa-core/src/topology.rs ---- */
/* Topology::build (topology.rs:276-298): stable grouping of the m edge rows by
 * source, within-segment order = original row order (topology.rs:109-114).
 * src_dense/dst_dense: u32[m], each < n. m must be <= UINT32_MAX
 * (topology.rs:61-64; URSA_HIP_ERANGE otherwise). */
int ursa_hip_topology_create(uint64_t n, uint64_t m, const uint32_t* src_dense,
                             const uint32_t* dst_dense, ursa_hip_topo** out);
/* A topology from an existing out-CSR (the Adjacency Topology::build produces,
 * topology.rs:35-47): offsets u64[n+1] (0 first, m last, non-decreasing),
 * targets u32[m] < n, edge_ids u32[m] a permutation of 0..m-1 (the original
 * row of each slot). Host or device buffers; validated (URSA_HIP_EINVAL). The
 * transpose and undirected view are derived from it as after create. */
int ursa_hip_topology_create_from_csr(uint64_t n, uint64_t m, const uint64_t* offsets,
                                      const uint32_t* targets, const uint32_t* edge_ids,
                                      ursa_hip_topo** out);
int ursa_hip_topology_destroy(ursa_hip_topo* t);
int ursa_hip_topology_size(const ursa_hip_topo* t, uint64_t* n, uint64_t* m);
/* Topology::incoming (topology.rs:343-348): build the transpose now. */
int ursa_hip_topology_incoming(ursa_hip_topo* t);
/* Topology::undirected (topology.rs:353-389): build now; *slots = target count. */
int ursa_hip_topology_undirected(ursa_hip_topo* t, uint64_t* slots);
/* Export a CSR to caller buffers (host or device): which = 0 out, 1 incoming,
 * 2 undirected (edge_ids must be NULL). offsets are u64[n+1] as in
 * Adjacency::offsets (topology.rs:35-47); targets/edge_ids u32[m] (or
 * u32[slots] for undirected). Any output pointer may be NULL. */
int ursa_hip_topology_export(ursa_hip_topo* t, int which, uint64_t* offsets, uint32_t* targets,
                             uint32_t* edge_ids);

/* ---- kernels — ursa-core/src/algo/ ---- */
/* degree (degree.rs:12-28) -> u32[n] */
int ursa_hip_degree(ursa_hip_topo* t, int direction, uint32_t* out);

/* pagerank / pagerank_weighted (pagerank.rs:98-169) with PageRankParams
 * (pagerank.rs:21-36) -> f64[n]. weights: f64[m] by original edge row, or NULL.
 * iters_out (may be NULL): iterations executed; the loop stops after the first
 * iteration whose L1 delta < tol (pagerank.rs:87-90). */
int ursa_hip_pagerank(ursa_hip_topo* t, const double* weights, double damping,
                      uint32_t max_iter, double tol, double* out, uint32_t* iters_out);

/* bfs_distances (bfs.rs:18-42) for k sources -> i32[k * n] row-major (row i =
 * levels from sources[i]; -1 = unreachable). An out-of-range source yields a
 * row of -1 (bfs.rs:21-23). */
int ursa_hip_bfs(ursa_hip_topo* t, const uint32_t* sources, uint32_t k, int direction,
                 int32_t* levels);

/* k_hop (hop.rs:52-75) -> (seeds_out, reached) pairs in the reference's order:
 * seeds in input order, reached nodes in discovery order. k == 0 or an
 * unknown seed contributes nothing; repeated seeds repeat their pairs. */
typedef struct ursa_hip_pairs ursa_hip_pairs;
int ursa_hip_k_hop(ursa_hip_topo* t, const uint32_t* seeds, uint64_t nseeds, uint32_t k,
                   int direction, ursa_hip_pairs** out);
int ursa_hip_pairs_size(const ursa_hip_pairs* p, uint64_t* count);
int ursa_hip_pairs_export(const ursa_hip_pairs* p, uint32_t* src, uint32_t* dst);
int ursa_hip_pairs_destroy(ursa_hip_pairs* p);

/* shortest_path (bfs.rs:48-102): *len = 0 for None, else the node count of
 * the path written to path[0..*len) (source first). The path is the one the
 * reference returns (first discoverer of every node). URSA_HIP_ERANGE if
 * *len > cap. */
int ursa_hip_shortest_path(ursa_hip_topo* t, uint32_t source, uint32_t target, int direction,
                           uint32_t* path, uint64_t cap, uint64_t* len);

/* betweenness (betweenness.rs:153-177): has_sample = 0 for None (exact),
 * otherwise ceil(sample*n) seed-shuffled sources and an n/k scale
 * (betweenness.rs:349-362); has_seed = 0 uses the reference's default seed.
 * -> f64[n]. */
int ursa_hip_betweenness(ursa_hip_topo* t, int has_sample, double sample, int has_seed,
                         uint64_t seed, double* out);
/* Brandes over an explicit source list (accumulate_sources,
 * betweenness.rs:45-70); out[v] *= scale afterwards when apply_scale. Sources
 * >= n are skipped. */
int ursa_hip_betweenness_sources(ursa_hip_topo* t, const uint32_t* sources, uint64_t k,
                                 int apply_scale, double scale, double* out);

/* triangle_count (triangle.rs:58-63) -> u64[n] (the reference accumulates in
 * u32, which wraps above 2^32 triangles per node). */
int ursa_hip_triangle_count(ursa_hip_topo* t, uint64_t* out);

/* connected_components_weak (components.rs:58-68) -> u32[n]. Each label is
 * the minimum dense id of the node's component (a valid representative under
 * the reference's contract, and its canonical form). */
int ursa_hip_cc_weak(ursa_hip_topo* t, uint32_t* labels);

/* No counterpart in the reference. ursa_hip_cc_weak for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): same labels, but no hipMalloc / hipFree or other
 * allocation, no host copy and no wait. `labels` must be device memory of n
 * u32; the work is queued on the calling thread's stream (ursa_hip_set_stream,
 * else the library's stream of that thread) and the labels are ready when that
 * stream reaches it. Scratch is a fixed device array (16 MiB per device, kept
 * until the process ends), so n is limited to 4194304 (URSA_HIP_ERANGE above);
 * calls that overlap on one device share it and race. */
int ursa_hip_cc_weak_noalloc(ursa_hip_topo* t, uint32_t* labels);

#ifdef __cplusplus
}
#endif
#endif /* URSA_HIP_H */
