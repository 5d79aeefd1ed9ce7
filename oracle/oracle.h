/*
 * oracle.h — CPU restatement of ursa-core's CSR hot-path kernels.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP product path
 * is compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Every function restates one reference function (cited as file:line relative
 * to the cldixon/ursa 0.2.0 tree) in plain C++, keeping the reference's
 * summation and traversal orders so f64 results agree to the last bits where
 * the reference's order is fixed, and integer results agree exactly.
 *
 * Conventions: status-returning functions return 0 on success, negative on a
 * bad argument. Variable-length outputs use a two-call pattern: pass cap=0 to
 * learn the length, then call again with a buffer of that size.
 */
#ifndef URSA_ORACLE_H
#define URSA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_topo orc_topo;

/* Direction: ursa-core/src/topology.rs:27-32 */
enum { ORC_OUT = 0, ORC_IN = 1, ORC_BOTH = 2 };

/* ---- topology (ursa-core/src/topology.rs) ---- */
orc_topo* orc_topo_create(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst);
void orc_topo_destroy(orc_topo* t);
uint64_t orc_topo_n(const orc_topo* t);
uint64_t orc_topo_m(const orc_topo* t);
/* which: 0 = out (topology.rs:276-298), 1 = incoming (topology.rs:343-348) */
void orc_topo_adjacency(orc_topo* t, int which, uint64_t* offsets, uint32_t* targets,
                        uint32_t* edge_ids);
/* undirected CSR (topology.rs:353-389); returns number of target slots */
uint64_t orc_topo_undirected(orc_topo* t, uint64_t* offsets, uint32_t* targets);

/* ---- id map (ursa-core/src/id_map.rs:220-319, int64 path) ---- */
/* Interns src/dst in first-seen order, src then dst per row. Writes dense ids
 * and the dense->user table (capacity 2*m). Returns n. */
uint64_t orc_intern_i64(uint64_t m, const int64_t* src, const int64_t* dst, uint32_t* src_dense,
                        uint32_t* dst_dense, int64_t* user_ids);

/* ---- rng (ursa-core/src/algo/rng.rs) ---- */
void orc_splitmix_stream(uint64_t seed, uint64_t count, uint64_t* out);
void orc_shuffled_order(uint64_t n, uint64_t seed, uint32_t* out);
/* returns k actually written (sorted ascending) */
uint64_t orc_sample_indices(uint64_t r, uint64_t k, int has_seed, uint64_t seed, uint64_t* out);
/* betweenness.rs:349-362; has_frac=0 means None. Returns k. */
uint64_t orc_sample_sources(uint64_t n, int has_frac, double frac, int has_seed, uint64_t seed,
                            uint32_t* out);
uint64_t orc_default_seed(void);

/* ---- kernels (ursa-core/src/algo/) ---- */
void orc_degree(orc_topo* t, int dir, uint32_t* out);
/* pagerank.rs:63-169. weights (length m, by original row) may be NULL.
 * iters_out: iterations executed; last_delta_out: L1 delta of the last one. */
void orc_pagerank(orc_topo* t, const double* weights, double damping, uint32_t max_iter,
                  double tol, int nthreads, double* out, uint32_t* iters_out,
                  double* last_delta_out);
/* Personalized PageRank (serving extension; no ursa-core counterpart): the
 * pull iteration of pagerank.rs:63-124 with teleport and dangling mass sent to
 * the seed set, p = uniform over the distinct seeds (networkx personalization
 * semantics): next[u] = d * acc[u] + ((1 - d) + d * dangling) * p[u].
 * Start 1/n, stop after the first iteration with L1 delta < tol. */
void orc_ppr(orc_topo* t, const uint32_t* seeds, uint64_t nseeds, double damping, uint32_t max_iter,
             double tol, double* out, uint32_t* iters_out);
void orc_bfs(orc_topo* t, uint32_t source, int dir, int32_t* out);
/* ursa-plan/src/stats.rs:80-168: strided stats source sample, bfs_source_scan
 * (sum over sources in order of per-source f64 sums over vertices in id order,
 * count, max), avg_path_length and diameter (out-edges). */
uint64_t orc_stats_sources(uint64_t n, int has_frac, double frac, uint32_t* out);
void orc_bfs_source_scan(orc_topo* t, const uint32_t* sources, uint64_t k, int dir, double* sum, uint64_t* count,
                         int32_t* max_dist);
double orc_avg_path_length(orc_topo* t, int has_sample, double sample);
int64_t orc_diameter(orc_topo* t, int approximate);
/* closeness.rs:69-137 (unweighted, Out): out[i] for sources[i]. */
void orc_closeness(orc_topo* t, const uint32_t* sources, uint64_t k, double* out);
/* closeness.rs:143-222 (weighted, Dijkstra over Out): w[m] per edge row. */
void orc_closeness_weighted(orc_topo* t, const double* w, const uint32_t* sources, uint64_t k, double* out);
/* clustering.rs:21-40 over the undirected view and u64 triangle counts. */
void orc_clustering(orc_topo* t, double* out);
/* hop.rs:52-130. Returns number of pairs; writes min(cap, count). */
uint64_t orc_k_hop(orc_topo* t, const uint32_t* seeds, uint64_t nseeds, uint32_t k, int dir,
                   uint32_t* out_src, uint32_t* out_dst, uint64_t cap);
/* bfs.rs:48-102. Returns path length (nodes), 0 for None. */
uint64_t orc_shortest_path(orc_topo* t, uint32_t s, uint32_t tgt, int dir, uint32_t* out_path,
                           uint64_t cap);
/* betweenness.rs:45-230 over an explicit source list; scale applied when
 * apply_scale != 0 (the sampled branch, betweenness.rs:171-176). */
void orc_betweenness_sources(orc_topo* t, const uint32_t* sources, uint64_t k, int apply_scale,
                             double scale, int nthreads, double* out);
/* betweenness.rs:153-177 with sample_sources */
void orc_betweenness(orc_topo* t, int has_frac, double frac, int has_seed, uint64_t seed,
                     int nthreads, double* out);
/* triangle.rs:23-63, accumulated in u64 */
void orc_triangle_count(orc_topo* t, int nthreads, uint64_t* out);
/* components.rs:14-68: raw union-find root labels */
void orc_cc_weak(orc_topo* t, uint32_t* out);
/* relabel each component by its minimum dense id */
void orc_canonicalize_labels(uint64_t n, const uint32_t* labels, uint32_t* out);

/* ---- workload generators (seeded synthetic inputs, identical to the GPU
 * generator's stream definition) ---- */
/* R-MAT: 16*2^scale draws by default (edges = ef << scale). Self-loops and
 * duplicates dropped keeping first occurrence; symmetrize appends reversed
 * kept edges and dedups again. Returns m; writes min(cap, m). */
uint64_t orc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int symmetrize, int64_t* src, int64_t* dst, uint64_t cap);
uint64_t orc_gen_er(uint64_t n, uint32_t edge_factor, uint64_t seed, int symmetrize, int64_t* src,
                    int64_t* dst, uint64_t cap);
/* f32-exact weights in (0,1]: w = ((x >> 40) + 1) / 2^24 from stream
 * seed ^ 0x5745494748540000 */
void orc_gen_weights(uint64_t m, uint64_t seed, double* out);

#ifdef __cplusplus
}
#endif
#endif
