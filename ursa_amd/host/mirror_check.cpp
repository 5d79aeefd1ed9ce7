// mirror_check.cpp — drives every function of the C++ mirror (ursa_core.hpp)
// on one edge list and writes the results into caller arrays, so a test can
// compare the mirror, end to end through the C ABI, with the oracle. Built as
// ursa_amd/host/libursa_mirror.so beside libursa_hip.so.
//
// Returns 0, or the status of the ursa_hip_error the mirror threw (its
// message copied into err).
#include <cstring>

#include "ursa_core.hpp"

namespace {

void copy_msg(const char* m, char* err, uint64_t cap) {
    if (!err || cap == 0) return;
    std::strncpy(err, m, cap - 1);
    err[cap - 1] = '\0';
}

}  // namespace

extern "C" {

// Outputs (n = n_nodes): deg_{out,in,both} u32[n]; pr, prw f64[n] (default
// PageRankParams, weights w f64[m]); bfs i32[n] from `source` (Out);
// hop_a/hop_b u32[hop_cap] with *hop_len (k_hop from {source}, depth 2, Out);
// path u32[n] with *path_len (shortest_path source -> goal, Out; 0 = none);
// bc f64[n] (exact betweenness); tri u64[n]; cc u32[n].
int ursa_mirror_run(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, const double* w,
                    uint32_t source, uint32_t goal, uint32_t* deg_out, uint32_t* deg_in, uint32_t* deg_both,
                    double* pr, double* prw, int32_t* bfs, uint32_t* hop_a, uint32_t* hop_b, uint64_t hop_cap,
                    uint64_t* hop_len, uint32_t* path, uint64_t* path_len, double* bc, uint64_t* tri, uint32_t* cc,
                    char* err, uint64_t err_cap) {
    using namespace ursa_core;
    try {
        std::vector<uint32_t> s(src, src + m), d(dst, dst + m);
        Topology t = Topology::build(n, s, d);
        auto put = [](const auto& v, auto* out) { std::copy(v.begin(), v.end(), out); };
        put(algo::degree(t, Direction::Out), deg_out);
        put(algo::degree(t, Direction::In), deg_in);
        put(algo::degree(t, Direction::Both), deg_both);
        put(algo::pagerank(t), pr);
        put(algo::pagerank_weighted(t, std::vector<double>(w, w + m)), prw);
        put(algo::bfs_distances(t, source, Direction::Out), bfs);
        auto hop = algo::k_hop(t, {source}, 2, Direction::Out);
        *hop_len = hop.first.size();
        if (hop.first.size() <= hop_cap) {
            put(hop.first, hop_a);
            put(hop.second, hop_b);
        }
        auto p = algo::shortest_path(t, source, goal, Direction::Out);
        *path_len = p ? p->size() : 0;
        if (p) put(*p, path);
        put(algo::betweenness(t, std::nullopt, std::nullopt), bc);
        put(algo::triangle_count(t), tri);
        put(algo::connected_components_weak(t), cc);
        return 0;
    } catch (const ursa_hip_error& e) {
        copy_msg(e.what(), err, err_cap);
        return e.code;
    }
}

// pagerank_weighted with `wlen` weights: the mirror rejects a length other
// than the edge count before any device call (pagerank.rs:140-144 panics).
int ursa_mirror_weights_len(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, uint64_t wlen,
                            char* err, uint64_t err_cap) {
    using namespace ursa_core;
    try {
        Topology t = Topology::build(n, std::vector<uint32_t>(src, src + m), std::vector<uint32_t>(dst, dst + m));
        algo::pagerank_weighted(t, std::vector<double>(wlen, 1.0));
        return 0;
    } catch (const ursa_hip_error& e) {
        copy_msg(e.what(), err, err_cap);
        return e.code;
    }
}

}  // extern "C"
