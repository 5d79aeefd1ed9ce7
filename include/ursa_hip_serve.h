/* ursa_hip_serve.h — request server over resident graphs (serving extension).
 *
 * ursa-core has no server: its kernels are called once per query from
 * GraphAlgorithmExec (ursa-plan/src/physical.rs:94-113) on tokio blocking
 * threads, and the topology is shared read-only across concurrent queries
 * (ursa-plan/src/node.rs:21-26). This server keeps many device topologies
 * resident and answers four query kinds on them:
 *   personalized PageRank (ursa_hip_ppr), k-hop neighbourhood
 *   (ursa_hip_k_hop, hop.rs:52-130), unweighted shortest path
 *   (ursa_hip_shortest_path, bfs.rs:48-102) and per-node triangle counts
 *   (ursa_hip_triangle_count, triangle.rs:58-63).
 *
 * Continuous batching: `slots` worker threads each own a HIP stream. A free
 * slot takes the oldest waiting query, joins to it every waiting query of the
 * same kind on the same graph (PPR up to 16 lanes of one solve, k-hop and
 * shortest-path queries in one workspace, identical triangle queries served by
 * one count), reserves the pass's device workspace from the HBM budget and
 * runs it; queries that arrive meanwhile wait for the next free slot. Queries
 * on small graphs are also joined across graphs: PPR on graphs of at most 2^19
 * nodes (one launch per sweep phase for all of them, up to max_batch), k-hop
 * and shortest path on graphs of at most 2^16 nodes (one workgroup per query,
 * one launch per pass, up to cross_batch). A pass
 * whose workspace does not fit waits for running passes to release memory
 * (first come, first served). Passes on different slots overlap on the device.
 * Every query's result is identical to the plain entry point's on the same
 * arguments, alone or in any mix.
 *
 * Serial mode (config.serial = 1) is the plain loop the server is measured
 * against: one query at a time in arrival order, each through the plain entry
 * point (its own launches and allocations), on one thread.
 */
#ifndef URSA_HIP_SERVE_H
#define URSA_HIP_SERVE_H

#include <stdint.h>

#include "ursa_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ursa_hip_server ursa_hip_server;

enum {
    URSA_HIP_Q_PPR = 0,       /* seeds[0..nseeds), damping, max_iter, tol -> f64[n] */
    URSA_HIP_Q_KHOP = 1,      /* seeds[0..nseeds), k, direction -> pairs: u32 src[len], u32 dst[len] */
    URSA_HIP_Q_PATH = 2,      /* seeds[0] -> seeds[1], direction -> u32 path[len] (len 0 = None) */
    URSA_HIP_Q_TRIANGLES = 3  /* -> u64[n] */
};

typedef struct {
    uint32_t kind;
    uint32_t graph;     /* id returned by ursa_hip_server_add_graph */
    uint32_t nseeds;    /* PPR 1..16, KHOP 1..16, PATH 2 */
    uint32_t k;         /* KHOP depth */
    uint32_t seeds[16];
    int32_t direction;  /* KHOP / PATH: URSA_HIP_OUT, _IN, _BOTH */
    uint32_t max_iter;  /* PPR */
    double damping;     /* PPR */
    double tol;         /* PPR */
    double arrival_s;   /* trace runs: arrival time in seconds after the run starts */
} ursa_hip_query;

typedef struct {
    int32_t status;    /* URSA_HIP_OK or the error code the query failed with */
    uint32_t iters;    /* PPR iterations */
    uint64_t len;      /* PPR/TRIANGLES: n; KHOP: pair count; PATH: node count */
    uint64_t digest;   /* ursa_hip_digest of the result bytes */
    double arrival_s;  /* seconds after the run started */
    double start_s;    /* its device pass began */
    double done_s;     /* its result was in host memory */
    uint32_t batch;    /* queries in that device pass */
    uint32_t slot;     /* worker slot that ran it */
} ursa_hip_result;

typedef struct {
    uint64_t workspace_bytes; /* HBM budget for in-flight passes (0: 25% of free memory, at most 32 GiB) */
    uint64_t graph_bytes;     /* HBM budget for resident graphs (0: unlimited) */
    uint32_t slots;           /* concurrent passes, one stream each (0: 8) */
    uint32_t max_batch;       /* queries per pass (0: 16; PPR passes take at most 16; 1: no batching) */
    int32_t serial;           /* 1: plain serial loop (slots, max_batch, budget ignored) */
    int32_t keep_results;     /* keep every result for ursa_hip_server_result_data */
    uint32_t cross_batch;     /* k-hop/shortest-path queries per cross-graph pass (0: 256; at least max_batch) */
    uint32_t reserved;
} ursa_hip_server_config;

int ursa_hip_server_create(const ursa_hip_server_config* cfg, ursa_hip_server** out);
/* Makes t resident for serving: builds every cache the four query kinds read
 * (transpose, undirected view, triangle orientation, PPR row segments) and
 * books the topology's device bytes against graph_bytes (URSA_HIP_ENOMEM if it
 * does not fit; the graph is then not added). The caller keeps ownership of t
 * and must not destroy it before the server. */
int ursa_hip_server_add_graph(ursa_hip_server* s, ursa_hip_topo* t, uint32_t* graph_id);
/* Resident graph bytes booked so far. */
int ursa_hip_server_graph_bytes(const ursa_hip_server* s, uint64_t* bytes);
/* Runs nq queries and waits for all of them. With timed_arrivals each query is
 * submitted at its arrival_s; otherwise all are submitted at once. res[i]
 * receives query i's record (times relative to the start of this call). */
int ursa_hip_server_run(ursa_hip_server* s, const ursa_hip_query* qs, uint64_t nq, int timed_arrivals,
                        ursa_hip_result* res);
/* Copies the result bytes of query `index` of the last run (keep_results) into
 * dst; *bytes receives the size (dst may be NULL to query it). */
int ursa_hip_server_result_data(const ursa_hip_server* s, uint64_t index, void* dst, uint64_t cap,
                                 uint64_t* bytes);
int ursa_hip_server_destroy(ursa_hip_server* s);

/* 64-bit digest of a byte string: sum over 8-byte little-endian words w_i
 * (zero-padded) of splitmix64(w_i + i * 0xD1B54A32D192ED03), plus
 * splitmix64(byte length), modulo 2^64. */
uint64_t ursa_hip_digest(const void* data, uint64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* URSA_HIP_SERVE_H */
