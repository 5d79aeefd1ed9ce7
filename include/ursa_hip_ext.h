/*
 * ursa_hip_ext.h — entry points of the gfx950 backend that have no
 * counterpart in ursa-core: kernel timing and seeded workload synthesis.
 * Callers: benchmarks and tests (they need device-side timing and large
 * synthetic graphs without a host round trip).
 */
#ifndef URSA_HIP_EXT_H
#define URSA_HIP_EXT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ursa_hip_topo ursa_hip_topo;

/* ---- per-kernel timing ----
 * When enabled, every launch of a named kernel is bracketed by HIP events on
 * the stream it is launched on. query() synchronises those events and
 * returns the summed duration, the launch count and the algorithmic bytes
 * booked against the kernel (its per-launch traffic model). */
int ursa_hip_profile_enable(int on);
int ursa_hip_profile_reset(void);
int ursa_hip_profile_query(const char* kernel, double* total_ms, uint64_t* launches,
                           double* algo_bytes);
/* Names of all kernels seen since the last reset, '\n'-separated. Returns the
 * length needed (excluding the terminator). */
uint64_t ursa_hip_profile_names(char* buf, uint64_t cap);

/* ---- traversal statistics for benchmarks ----
 * As ursa_hip_bfs, also returning the TEPS numerator (the sum over all k
 * traversals of the forward degrees of the reached vertices) and choosing the
 * schedule: mode 0 = automatic (what ursa_hip_bfs does: 64-lane multi-source
 * BFS for k >= 4), 1 = one direction-optimizing traversal per source,
 * 2 = 64-lane multi-source BFS. All modes give identical levels. */
int ursa_hip_bfs_ex(ursa_hip_topo* t, const uint32_t* sources, uint32_t k, int direction,
                    int32_t* levels, uint64_t* reached_edges, int mode);
/* Traversal schedule autotuner. Times the BFS direction switch (alpha, beta in
 * {4, 15, 60} x {6, 18, 54}) on 8 single-source traversals and the multi-source
 * push rule (divisor in {5, 20, 80}) on two 64-source batches of this graph
 * (sources in a seeded random order over the vertices with forward edges), and
 * keeps the fastest for `direction` on the topology; a candidate displaces the
 * defaults (15, 18; 20) only when more than 5 % faster on each half of the
 * sample (4 traversals, one batch) separately. The choice is also
 * cached under a key of the forward degree distribution (log2 n, mean degree,
 * max/mean degree, share of zero-degree vertices); a later graph with the same
 * key takes it without timing unless `force`. *source: 1 measured, 2 from the
 * key cache. Results of every traversal are identical under any parameters. */
int ursa_hip_autotune(ursa_hip_topo* t, int direction, int force, double* alpha, double* beta,
                      uint32_t* push_div, int* source);

/* *out = ((0.0 + x[0]) + x[1]) + ... + x[n-1], rounded step by step as a
 * sequential f64 loop rounds (round-to-nearest-even), computed in parallel.
 * PageRank forms its dangling mass this way (the reference sums it with
 * `dangling.iter().map(|&u| rank[u]).sum()`, ursa-core/src/algo/pagerank.rs:78);
 * exported so the summation can be checked bit for bit on its own. x may be a
 * host or device pointer; n < 2^32. */
int ursa_hip_ordered_sum(const double* x, uint64_t n, double* out);

/* As ursa_hip_betweenness_sources with an explicit batch size (1..64 sources
 * advancing together; 1 = one traversal per launch sequence) and the TEPS
 * numerator (sum of out-degrees of reached vertices over all sources). */
int ursa_hip_betweenness_sources_ex(ursa_hip_topo* t, const uint32_t* sources, uint64_t k,
                                    int apply_scale, double scale, uint32_t batch, double* out,
                                    uint64_t* reached_edges);

/* The reference's seeded permutation of 0..n-1 (shuffled_order, rng.rs:64-73),
 * used to pick sampled sources. Host code; out is u32[n]. */
int ursa_hip_shuffled_order(uint64_t n, uint64_t seed, uint32_t* out);

/* Fixed launch sequences (8-iteration batches of PageRank and personalized
 * PageRank) are recorded once as HIP graphs and replayed (default on). With 0
 * they are launched one by one; results are identical either way. */
int ursa_hip_graphs_enable(int on);

/* Triangle-count tables (default on): owners of long oriented lists among the
 * 65536 highest-ranked vertices are staged as LDS bitmaps, the rest as LDS
 * hash sets. With 0, topologies whose triangle orientation is built afterwards
 * use hash sets only. Counts are identical either way. */
int ursa_hip_triangle_bitmaps(int on);
/* Work figures of triangle_count on t (builds its orientation if needed):
 * oriented edges |E+|, the sorted-merge intersection length
 * sum over (u,v) in E+ of d+(u) + d+(v), and the list entries the one-sided
 * walks read (each pair's shorter side, as the kernels walk it). */
int ursa_hip_triangle_work(ursa_hip_topo* t, uint64_t* oriented_edges, uint64_t* merge_steps,
                           uint64_t* walk_steps);

/* PageRank sweep schedule (default 0): 0 automatic (binned from 32 blocks of
 * 8192 nodes up), 1 pull (each in-row sums its sources' contributions gathered
 * in new-id order), 2 binned (terms streamed through a tiling of 8192-source
 * blocks by 4096-row bins and added into exact fixed-point row sums in LDS).
 * Both binned choices fall back to pull when the tiling would exceed 2^26
 * tiles (above about 47 M nodes). Every schedule is deterministic and within
 * 1e-12 relative of the reference per node; pull and binned differ in the last
 * bits. */
int ursa_hip_pagerank_schedule(int mode);

/* *out = sum of x[0..n) computed exactly in fixed point (each value truncated
 * below 2^-100) and rounded to nearest even: the accumulator of the binned
 * PageRank sweep, exported so it can be checked against exact rational sums.
 * Every |x[i]| < 4 (checked for host x) and the sum below 8 in magnitude
 * (integer parts wrap, so partial sums may exceed it); x host or device. */
int ursa_hip_exact_sum(const double* x, uint64_t n, double* out);

/* ---- mixed-precision PageRank ----
 * Same contract as ursa_hip_pagerank (pagerank.rs:98-169). The first sweeps
 * gather f32 contributions and sum each row piece with compensated (Neumaier)
 * f32 additions; once the L1 delta falls below 2e-6 or stops shrinking, f64
 * sweeps refine the same iterate until delta < tol. f32_iters_out (may be
 * NULL) = how many of *iters_out were f32 sweeps. */
int ursa_hip_pagerank_mixed(ursa_hip_topo* t, const double* weights, double damping,
                            uint32_t max_iter, double tol, double* out, uint32_t* iters_out,
                            uint32_t* f32_iters_out);

/* ---- personalized PageRank (serving extension; ursa-core has none) ----
 * pagerank.rs:63-124's iteration with the teleport and dangling mass sent to
 * the seed set (networkx personalization semantics): x0 = 1/n,
 * next[u] = d * sum_{v in in(u)} x[v]/od[v] + ((1-d) + d*dm) * p[u], p uniform
 * over the distinct in-range seeds, stop after the first iteration whose L1
 * delta < tol. f32 compensated sweeps first, then f64 sweeps refine to tol.
 * No seed in range (or max_iter 0): out = 1/n, *iters_out = 0. out is f64[n]. */
int ursa_hip_ppr(ursa_hip_topo* t, const uint32_t* seeds, uint64_t nseeds, double damping,
                 uint32_t max_iter, double tol, double* out, uint32_t* iters_out);
/* k queries on one graph solved together, up to 16 per device pass; query q's
 * seeds are seeds[seed_offsets[q] .. seed_offsets[q+1]), its result
 * out[q*n .. (q+1)*n) and its iteration count iters_out[q]. Each query's result
 * is bit-identical to ursa_hip_ppr on the same arguments. */
int ursa_hip_ppr_batch(ursa_hip_topo* t, uint32_t k, const uint64_t* seed_offsets,
                       const uint32_t* seeds, double damping, uint32_t max_iter, double tol,
                       double* out, uint32_t* iters_out);
/* One query on each of `count` graphs, solved in cross-graph passes: query i
 * runs on topos[i] with seeds seeds[seed_offsets[i] .. seed_offsets[i+1]) and
 * writes f64[n_i] to outs[i] (host or device memory) and its iteration count
 * to iters_out[i]. Graphs of at most 2^19 vertices on the pull schedule are
 * swept together (up to 256 queries and 8 Mi vertices per pass, one launch per
 * sweep phase for all of them); other graphs are solved one at a time. Every
 * result is bit-identical to ursa_hip_ppr on the same arguments. No
 * counterpart in ursa-core: the request server's cross-graph batching. */
int ursa_hip_ppr_multi(uint32_t count, ursa_hip_topo* const* topos, const uint64_t* seed_offsets,
                       const uint32_t* seeds, double damping, uint32_t max_iter, double tol,
                       double* const* outs, uint32_t* iters_out);
/* Personalized PageRank sweep schedule (default 0): 0 automatic (binned on
 * graphs of 32 or more blocks of 8192 vertices with no row above 256
 * in-edges and at least 8 edges per vertex, else pull), 1 pull, 2 binned (each lane's sweep streamed through
 * PageRank's source-block x row-bin tiling, exact fixed-point row sums). The
 * schedule never depends on the batch, so the bit-identity above holds under
 * every schedule; pull and binned differ in the last bits. */
int ursa_hip_ppr_schedule(int mode);

/* Streaming device-to-device copy of `bytes` (multiple of 16) by a plain
 * kernel: the reference point for hardware byte counters (k_copy_probe moves
 * exactly `bytes` in and `bytes` out). */
int ursa_hip_copy_probe(const void* src_device, void* dst_device, uint64_t bytes);

/* out[i] = src[idx[i]] for i < count (all device memory; every idx[i] <
 * src_len, which the caller guarantees): the reference point for hardware byte
 * counters under random 8-byte gathers. */
int ursa_hip_gather_probe(const double* src_device, uint64_t src_len, const uint32_t* idx_device, uint64_t count,
                          double* out_device);

/* ---- the formats either side of the kernels ---- */
/* Edge weights in adjacency-slot order: w_slots[k] = w_rows[edge_ids[k]] for
 * the out-CSR (which = 0) or the in-CSR (which = 1), the gather a weighted
 * kernel does per slot (ursa-plan/src/weight.rs:8-10, ursa-core/src/algo/
 * pagerank.rs:159-161). w_rows is evaluate_weight's one f64 per edge row
 * (ursa-plan/src/weight.rs:51-106); like evaluate_weight, any negative weight
 * is URSA_HIP_EINVAL ("weights must be non-negative"; w_slots is filled
 * anyway). Host or device pointers, m entries each. */
int ursa_hip_gather_edge_weights(ursa_hip_topo* t, int which, const double* w_rows, double* w_slots);

/* Pinned host memory for result buffers: algo_array moves a kernel's Vec into
 * an Arrow array without a copy (ursa-plan/src/result.rs:84-139); a buffer
 * from here can back that array (freed by ursa_hip_host_free when the array
 * is dropped), and any entry point writing a result into it moves it by DMA
 * at link speed. bytes == 0 gives NULL. */
int ursa_hip_host_alloc(uint64_t bytes, void** out);
int ursa_hip_host_free(void* p);

/* ---- device memory for callers without a HIP runtime binding ---- */
int ursa_hip_device_alloc(uint64_t bytes, void** out);
int ursa_hip_device_free(void* p);
int ursa_hip_memcpy_to_host(void* dst, const void* src_device, uint64_t bytes);
int ursa_hip_memcpy_to_device(void* dst_device, const void* src, uint64_t bytes);

/* ---- seeded workload synthesis on the device ----
 * R-MAT: edge_factor << scale edge draws; each draw's endpoints come from
 * `scale` quadrant choices (MSB first) with probabilities a, b, c, 1-a-b-c,
 * using splitmix64 in counter form (draw t of stream `seed` is
 * mix(seed + t * 0x9E3779B97F4A7C15)). Erdos-Renyi: edge_factor * n draws,
 * src = draw(2e+1) % n, dst = draw(2e+2) % n. Both drop self-loops and later
 * duplicates (first occurrence kept, row order = generation order);
 * symmetrize appends every kept (v,u) after the kept rows and dedups again.
 * Node ids are then interned first-seen, src then dst per row
 * (ursa-core/src/id_map.rs:220-292), so the dense numbering is the one ursa
 * would assign to the same int64 edge list. */
typedef struct ursa_hip_edges ursa_hip_edges;
int ursa_hip_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int symmetrize, ursa_hip_edges** out);
int ursa_hip_gen_er(uint64_t n, uint32_t edge_factor, uint64_t seed, int symmetrize,
                    ursa_hip_edges** out);
int ursa_hip_edges_size(const ursa_hip_edges* g, uint64_t* n, uint64_t* m);
/* Any pointer may be NULL; host or device buffers: user_src/user_dst int64[m]
 * (generated ids), src_dense/dst_dense u32[m], user_ids int64[n]. */
int ursa_hip_edges_export(ursa_hip_edges* g, int64_t* user_src, int64_t* user_dst,
                          uint32_t* src_dense, uint32_t* dst_dense, int64_t* user_ids);
/* Device addresses of the dense endpoint arrays (owned by the handle). */
int ursa_hip_edges_dense_ptrs(ursa_hip_edges* g, const uint32_t** src_dense,
                              const uint32_t** dst_dense);
/* Topology::build straight from the device-resident dense arrays. */
int ursa_hip_edges_topology(ursa_hip_edges* g, ursa_hip_topo** out);
int ursa_hip_edges_destroy(ursa_hip_edges* g);
/* f32-exact edge weights in (0,1]: w = ((x >> 40) + 1) / 2^24 with x drawn
 * from the stream seed ^ 0x5745494748540000 (draw e+1 for edge e). */
int ursa_hip_gen_weights(uint64_t m, uint64_t seed, double* out);

#ifdef __cplusplus
}
#endif
#endif /* URSA_HIP_EXT_H */
