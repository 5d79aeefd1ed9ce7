// ursa_core.hpp — C++ host mirror of ursa-core's kernel API over the gfx950
// C ABI (include/ursa_hip.h). Header-only; link against libursa_hip.so.
//
// Names, argument meaning and edge cases follow ursa_core (ursa-core/src/lib.rs:
// 25-30, ursa-core/src/algo/mod.rs:35-50): Topology::build, Direction,
// PageRankParams, and one free function per kernel returning an owned vector.
// Where the reference panics on misuse (pagerank.rs:140-144, topology.rs:61-64)
// this mirror throws ursa_hip_error carrying the C status and message.
#pragma once

#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ursa_hip.h"

namespace ursa_core {

struct ursa_hip_error : std::runtime_error {
    int code;
    ursa_hip_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != URSA_HIP_OK) throw ursa_hip_error(rc, ursa_hip_last_error());
}

// topology.rs:27-32
enum class Direction : int { Out = URSA_HIP_OUT, In = URSA_HIP_IN, Both = URSA_HIP_BOTH };

// pagerank.rs:21-36
struct PageRankParams {
    double damping = 0.85;
    uint32_t max_iter = 30;
    double tol = 1e-6;
};

// topology.rs:251-404: device-resident out-CSR, lazy transpose and undirected view.
class Topology {
   public:
    // Topology::build (topology.rs:276-298)
    static Topology build(size_t n_nodes, const std::vector<uint32_t>& src_dense,
                          const std::vector<uint32_t>& dst_dense) {
        if (src_dense.size() != dst_dense.size()) throw ursa_hip_error(URSA_HIP_EINVAL, "src/dst length mismatch");
        ursa_hip_topo* h = nullptr;
        check(ursa_hip_topology_create(n_nodes, src_dense.size(), src_dense.data(), dst_dense.data(), &h));
        return Topology(h, n_nodes, src_dense.size());
    }
    size_t n_nodes() const { return n_; }
    size_t n_edges() const { return m_; }
    ursa_hip_topo* raw() const { return h_.get(); }

   private:
    struct Del {
        void operator()(ursa_hip_topo* t) const { ursa_hip_topology_destroy(t); }
    };
    // Shared like ursa's Arc<Topology>: copies refer to one device topology.
    Topology(ursa_hip_topo* h, size_t n, size_t m) : h_(h, Del{}), n_(n), m_(m) {}
    std::shared_ptr<ursa_hip_topo> h_;
    size_t n_ = 0, m_ = 0;
};

namespace algo {

// degree.rs:12-28
inline std::vector<uint32_t> degree(const Topology& t, Direction d) {
    std::vector<uint32_t> out(t.n_nodes());
    check(ursa_hip_degree(t.raw(), (int)d, out.data()));
    return out;
}

// pagerank.rs:98-124
inline std::vector<double> pagerank(const Topology& t, PageRankParams p = {}) {
    std::vector<double> out(t.n_nodes());
    check(ursa_hip_pagerank(t.raw(), nullptr, p.damping, p.max_iter, p.tol, out.data(), nullptr));
    return out;
}

// pagerank.rs:135-169 (weights by original edge row)
inline std::vector<double> pagerank_weighted(const Topology& t, const std::vector<double>& weights,
                                             PageRankParams p = {}) {
    if (weights.size() != t.n_edges())
        throw ursa_hip_error(URSA_HIP_EINVAL, "weights length must equal the edge count");
    std::vector<double> out(t.n_nodes());
    check(ursa_hip_pagerank(t.raw(), weights.data(), p.damping, p.max_iter, p.tol, out.data(), nullptr));
    return out;
}

// bfs.rs:18-42
inline std::vector<int32_t> bfs_distances(const Topology& t, uint32_t source, Direction d) {
    std::vector<int32_t> out(t.n_nodes());
    if (!out.empty()) check(ursa_hip_bfs(t.raw(), &source, 1, (int)d, out.data()));
    return out;
}

// hop.rs:52-75
inline std::pair<std::vector<uint32_t>, std::vector<uint32_t>> k_hop(const Topology& t,
                                                                     const std::vector<uint32_t>& seeds, uint32_t k,
                                                                     Direction d) {
    ursa_hip_pairs* p = nullptr;
    check(ursa_hip_k_hop(t.raw(), seeds.data(), seeds.size(), k, (int)d, &p));
    uint64_t cnt = 0;
    ursa_hip_pairs_size(p, &cnt);
    std::vector<uint32_t> a(cnt), b(cnt);
    int rc = ursa_hip_pairs_export(p, a.data(), b.data());
    ursa_hip_pairs_destroy(p);
    check(rc);
    return {std::move(a), std::move(b)};
}

// bfs.rs:48-102
inline std::optional<std::vector<uint32_t>> shortest_path(const Topology& t, uint32_t s, uint32_t g, Direction d) {
    std::vector<uint32_t> path(t.n_nodes() ? t.n_nodes() : 1);
    uint64_t len = 0;
    check(ursa_hip_shortest_path(t.raw(), s, g, (int)d, path.data(), path.size(), &len));
    if (len == 0) return std::nullopt;
    path.resize(len);
    return path;
}

// betweenness.rs:153-177
inline std::vector<double> betweenness(const Topology& t, std::optional<double> sample,
                                       std::optional<uint64_t> seed) {
    std::vector<double> out(t.n_nodes());
    check(ursa_hip_betweenness(t.raw(), sample.has_value(), sample.value_or(0.0), seed.has_value(),
                               seed.value_or(0), out.data()));
    return out;
}

// triangle.rs:58-63 (u64 counts)
inline std::vector<uint64_t> triangle_count(const Topology& t) {
    std::vector<uint64_t> out(t.n_nodes());
    check(ursa_hip_triangle_count(t.raw(), out.data()));
    return out;
}

// components.rs:58-68 (labels = component minimum dense id)
inline std::vector<uint32_t> connected_components_weak(const Topology& t) {
    std::vector<uint32_t> out(t.n_nodes());
    check(ursa_hip_cc_weak(t.raw(), out.data()));
    return out;
}

}  // namespace algo
}  // namespace ursa_core
