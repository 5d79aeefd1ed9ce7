// oracle.cpp — CPU restatement of ursa-core's CSR hot-path kernels.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path. Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Each function follows the cited reference function (paths relative to the
// cldixon/ursa 0.2.0 tree). Where the reference fixes a floating-point
// summation order (serial folds, 8192-element L1 chunks, 64-source Brandes
// chunks) this restatement keeps that order, so results are bit-identical to
// the reference at any thread count. Threads are used only where the reference
// uses rayon and the order is provably unaffected.
//
// Parity pins (what checks each function; see tests/test_oracle.py):
//   Topology build/incoming/undirected — reference unit tests topology.rs:443-513
//     plus a brute-force stable grouping check.
//   degree — degree.rs:35-41 golden triple.
//   pagerank — pagerank.rs:176-219 invariants, networkx.pagerank fixtures
//     (committed), a dense power-iteration closed form, star/cycle closed forms.
//   bfs / shortest_path / k_hop — bfs.rs:266-332, hop.rs:148-191 goldens,
//     scipy csgraph BFS-distance fixtures, path validity.
//   betweenness — betweenness.rs:369-460 goldens, networkx fixtures on simple
//     graphs, brute-force path enumeration on tiny graphs.
//   triangle_count — triangle.rs:70-99 goldens, networkx.triangles fixtures,
//     O(n^3) brute force.
//   cc_weak — components.rs:145-153 partition golden, scipy csgraph fixtures.
//   rng — rng.rs:80-121 properties plus a committed fixture of splitmix64
//     outputs checked against the published splitmix64 test values.
//   intern — id_map.rs:459-478 goldens.

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <queue>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

// ursa-core/src/topology.rs:35-47
struct Adj {
    std::vector<uint64_t> off;
    std::vector<uint32_t> tgt;
    std::vector<uint32_t> eid;
    uint32_t degree(uint32_t u) const { return (uint32_t)(off[u + 1] - off[u]); }
};

// ursa-core/src/topology.rs:79-107 — serial two-pass counting sort; the
// parallel build (:115-187) is byte-identical by construction (:109-114).
Adj build_adj(uint64_t n, const uint32_t* keys, const uint32_t* other, uint64_t m) {
    Adj a;
    a.off.assign(n + 1, 0);
    for (uint64_t i = 0; i < m; ++i) a.off[(uint64_t)keys[i] + 1] += 1;
    for (uint64_t i = 0; i < n; ++i) a.off[i + 1] += a.off[i];
    a.tgt.assign(m, 0);
    a.eid.assign(m, 0);
    std::vector<uint64_t> cursor(a.off.begin(), a.off.begin() + n);
    for (uint64_t row = 0; row < m; ++row) {
        uint64_t pos = cursor[keys[row]]++;
        a.tgt[pos] = other[row];
        a.eid[pos] = (uint32_t)row;
    }
    return a;
}

int resolve_threads(int nthreads) {
    if (nthreads > 0) return nthreads;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

// Run f(lo, hi) over [0, n) split into nthreads contiguous ranges.
template <class F>
void parallel_ranges(uint64_t n, int nthreads, F f) {
    nthreads = resolve_threads(nthreads);
    if (nthreads <= 1 || n < 4096) {
        f((uint64_t)0, n);
        return;
    }
    std::vector<std::thread> ts;
    uint64_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        uint64_t lo = (uint64_t)t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        ts.emplace_back([=, &f] { f(lo, hi); });
    }
    for (auto& th : ts) th.join();
}

}  // namespace

// ursa-core/src/topology.rs:251-261
struct orc_topo {
    uint64_t n = 0, m = 0;
    Adj out;
    bool has_inc = false;
    Adj inc;
    bool has_und = false;
    std::vector<uint64_t> uoff;
    std::vector<uint32_t> utgt;

    // topology.rs:343-348 via reconstruct_endpoints :307-321
    const Adj& incoming() {
        if (!has_inc) {
            std::vector<uint32_t> src_by_row(m), dst_by_row(m);
            for (uint64_t u = 0; u < n; ++u) {
                for (uint64_t k = out.off[u]; k < out.off[u + 1]; ++k) {
                    uint32_t row = out.eid[k];
                    src_by_row[row] = (uint32_t)u;
                    dst_by_row[row] = out.tgt[k];
                }
            }
            inc = build_adj(n, dst_by_row.data(), src_by_row.data(), m);
            has_inc = true;
        }
        return inc;
    }

    // topology.rs:360-389
    void undirected() {
        if (has_und) return;
        const Adj& in = incoming();
        uoff.assign(n + 1, 0);
        std::vector<std::vector<uint32_t>> lists(n);
        for (uint64_t u = 0; u < n; ++u) {
            auto& l = lists[u];
            for (uint64_t k = out.off[u]; k < out.off[u + 1]; ++k)
                if (out.tgt[k] != u) l.push_back(out.tgt[k]);
            for (uint64_t k = in.off[u]; k < in.off[u + 1]; ++k)
                if (in.tgt[k] != u) l.push_back(in.tgt[k]);
            std::sort(l.begin(), l.end());
            l.erase(std::unique(l.begin(), l.end()), l.end());
        }
        for (uint64_t u = 0; u < n; ++u) uoff[u + 1] = uoff[u] + lists[u].size();
        utgt.reserve(uoff[n]);
        for (auto& l : lists) utgt.insert(utgt.end(), l.begin(), l.end());
        has_und = true;
    }

    // topology.rs:395-404 — Out, In, or out-then-in for Both.
    template <class F>
    void for_each_neighbor(uint32_t u, int dir, F f) {
        if (dir == ORC_OUT || dir == ORC_BOTH)
            for (uint64_t k = out.off[u]; k < out.off[u + 1]; ++k) f(out.tgt[k]);
        if (dir == ORC_IN || dir == ORC_BOTH) {
            const Adj& in = incoming();
            for (uint64_t k = in.off[u]; k < in.off[u + 1]; ++k) f(in.tgt[k]);
        }
    }
};

extern "C" {

// topology.rs:276-298
orc_topo* orc_topo_create(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst) {
    for (uint64_t i = 0; i < m; ++i)
        if (src[i] >= n || dst[i] >= n) return nullptr;
    if (m > 0xFFFFFFFFull) return nullptr;  // topology.rs:61-64
    orc_topo* t = new orc_topo;
    t->n = n;
    t->m = m;
    t->out = build_adj(n, src, dst, m);
    return t;
}

void orc_topo_destroy(orc_topo* t) { delete t; }
uint64_t orc_topo_n(const orc_topo* t) { return t->n; }
uint64_t orc_topo_m(const orc_topo* t) { return t->m; }

void orc_topo_adjacency(orc_topo* t, int which, uint64_t* offsets, uint32_t* targets,
                        uint32_t* edge_ids) {
    const Adj& a = which == 0 ? t->out : t->incoming();
    if (offsets) std::memcpy(offsets, a.off.data(), (t->n + 1) * sizeof(uint64_t));
    if (targets && t->m) std::memcpy(targets, a.tgt.data(), t->m * sizeof(uint32_t));
    if (edge_ids && t->m) std::memcpy(edge_ids, a.eid.data(), t->m * sizeof(uint32_t));
}

uint64_t orc_topo_undirected(orc_topo* t, uint64_t* offsets, uint32_t* targets) {
    t->undirected();
    if (offsets) std::memcpy(offsets, t->uoff.data(), (t->n + 1) * sizeof(uint64_t));
    if (targets && !t->utgt.empty())
        std::memcpy(targets, t->utgt.data(), t->utgt.size() * sizeof(uint32_t));
    return t->utgt.size();
}

// id_map.rs:220-292 (int64 fast path :246-255) and intern_i64 :294-319:
// first-seen order, src then dst within each row.
uint64_t orc_intern_i64(uint64_t m, const int64_t* src, const int64_t* dst, uint32_t* src_dense,
                        uint32_t* dst_dense, int64_t* user_ids) {
    std::unordered_map<int64_t, uint32_t> to_dense;
    to_dense.reserve(m * 2 + 1);
    uint64_t next = 0;
    auto intern = [&](int64_t id) -> uint32_t {
        auto it = to_dense.find(id);
        if (it != to_dense.end()) return it->second;
        uint32_t idx = (uint32_t)next++;
        to_dense.emplace(id, idx);
        user_ids[idx] = id;
        return idx;
    };
    for (uint64_t i = 0; i < m; ++i) {
        src_dense[i] = intern(src[i]);
        dst_dense[i] = intern(dst[i]);
    }
    return next;
}

// ---- rng.rs ----
static const uint64_t kDefaultSeed = 0x5EED5EED5EED5EEDull;  // rng.rs:9
static const uint64_t kGamma = 0x9E3779B97F4A7C15ull;        // rng.rs:21

// rng.rs:20-28 — splitmix64 output function.
static inline uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct SplitMix64 {  // rng.rs:12-34
    uint64_t state;
    uint64_t next_u64() {
        state += kGamma;
        return splitmix_mix(state);
    }
    uint64_t below(uint64_t bound) { return next_u64() % bound; }
};

uint64_t orc_default_seed(void) { return kDefaultSeed; }

void orc_splitmix_stream(uint64_t seed, uint64_t count, uint64_t* out) {
    SplitMix64 r{seed};
    for (uint64_t i = 0; i < count; ++i) out[i] = r.next_u64();
}

// rng.rs:64-73
void orc_shuffled_order(uint64_t n, uint64_t seed, uint32_t* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = (uint32_t)i;
    SplitMix64 r{seed ^ 0xD1B54A32D192ED03ull};
    for (uint64_t i = n; i-- > 1;) {
        uint64_t j = r.below(i + 1);
        std::swap(out[i], out[j]);
    }
}

// rng.rs:45-59
uint64_t orc_sample_indices(uint64_t r, uint64_t k, int has_seed, uint64_t seed, uint64_t* out) {
    k = std::min(k, r);
    std::vector<uint64_t> pool(r);
    for (uint64_t i = 0; i < r; ++i) pool[i] = i;
    SplitMix64 rng{(has_seed ? seed : kDefaultSeed) ^ 0x2545F4914F6CDD1Dull};
    for (uint64_t i = 0; i < k; ++i) {
        uint64_t j = i + rng.below(r - i);
        std::swap(pool[i], pool[j]);
    }
    std::sort(pool.begin(), pool.begin() + k);
    std::memcpy(out, pool.data(), k * sizeof(uint64_t));
    return k;
}

// betweenness.rs:349-362
uint64_t orc_sample_sources(uint64_t n, int has_frac, double frac, int has_seed, uint64_t seed,
                            uint32_t* out) {
    if (!has_frac) {
        for (uint64_t i = 0; i < n; ++i) out[i] = (uint32_t)i;
        return n;
    }
    double f = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);  // f64::clamp
    double kf = std::ceil((double)n * f);
    // Rust `as usize` saturates: NaN -> 0, negative -> 0.
    uint64_t k = (kf != kf || kf <= 0.0) ? 0 : (kf >= 1.8e19 ? UINT64_MAX : (uint64_t)kf);
    k = std::max<uint64_t>(k, 1);
    k = std::min<uint64_t>(k, n);
    if (k == 0) return 0;
    std::vector<uint32_t> order(n);
    orc_shuffled_order(n, has_seed ? seed : kDefaultSeed, order.data());
    std::memcpy(out, order.data(), k * sizeof(uint32_t));
    return k;
}

// ---- degree.rs:12-28 ----
void orc_degree(orc_topo* t, int dir, uint32_t* out) {
    for (uint64_t u = 0; u < t->n; ++u) {
        uint32_t d = 0;
        if (dir == ORC_OUT || dir == ORC_BOTH) d += t->out.degree((uint32_t)u);
        if (dir == ORC_IN || dir == ORC_BOTH) d += t->incoming().degree((uint32_t)u);
        out[u] = d;
    }
}

// ---- pagerank.rs ----
// l1_delta, pagerank.rs:44-52: 8192-element chunks summed left to right, then
// the chunk partials summed left to right (Rust's f64 Sum folds from -0.0).
static double l1_delta(const std::vector<double>& a, const std::vector<double>& b) {
    const uint64_t CHUNK = 8192;
    const uint64_t n = a.size();
    double total = -0.0;
    for (uint64_t lo = 0; lo < n; lo += CHUNK) {
        uint64_t hi = std::min(n, lo + CHUNK);
        double s = -0.0;
        for (uint64_t i = lo; i < hi; ++i) s += std::fabs(a[i] - b[i]);
        total += s;
    }
    return total;
}

// pagerank.rs:98-124 (unweighted) and :135-169 (weighted), through the shared
// power_iterate :63-94.
void orc_pagerank(orc_topo* t, const double* weights, double damping, uint32_t max_iter,
                  double tol, int nthreads, double* out, uint32_t* iters_out,
                  double* last_delta_out) {
    const uint64_t n = t->n;
    if (iters_out) *iters_out = 0;
    if (last_delta_out) *last_delta_out = 0.0;
    if (n == 0) return;
    const Adj& out_adj = t->out;
    std::vector<uint32_t> out_deg(n);
    for (uint64_t u = 0; u < n; ++u) out_deg[u] = out_adj.degree((uint32_t)u);
    std::vector<double> out_str;
    if (weights) {  // pagerank.rs:148-150
        out_str.assign(n, 0.0);
        for (uint64_t v = 0; v < n; ++v) {
            double s = -0.0;
            for (uint64_t k = out_adj.off[v]; k < out_adj.off[v + 1]; ++k) s += weights[out_adj.eid[k]];
            out_str[v] = s;
        }
    }
    const Adj& inc = t->incoming();
    std::vector<uint64_t> dangling;  // pagerank.rs:111, :155
    for (uint64_t u = 0; u < n; ++u)
        if (weights ? !(out_str[u] > 0.0) : out_deg[u] == 0) dangling.push_back(u);

    const double nf = (double)n;
    const double d = damping;
    std::vector<double> rank(n, 1.0 / nf), next(n, 0.0);
    uint32_t iters = 0;
    double delta = 0.0;
    for (uint32_t it = 0; it < max_iter; ++it) {
        double dm = -0.0;
        for (uint64_t u : dangling) dm += rank[u];
        const double base = (1.0 - d) / nf + d * dm / nf;
        parallel_ranges(n, nthreads, [&](uint64_t lo, uint64_t hi) {
            for (uint64_t u = lo; u < hi; ++u) {
                double acc = 0.0;
                for (uint64_t k = inc.off[u]; k < inc.off[u + 1]; ++k) {
                    uint32_t v = inc.tgt[k];
                    if (weights) {  // pagerank.rs:161-166
                        double os = out_str[v];
                        if (os > 0.0) acc += weights[inc.eid[k]] * rank[v] / os;
                    } else {  // pagerank.rs:113-122
                        uint32_t od = out_deg[v];
                        if (od > 0) acc += rank[v] / (double)od;
                    }
                }
                next[u] = base + d * acc;
            }
        });
        delta = l1_delta(next, rank);
        std::swap(rank, next);
        iters = it + 1;
        if (delta < tol) break;
    }
    std::memcpy(out, rank.data(), n * sizeof(double));
    if (iters_out) *iters_out = iters;
    if (last_delta_out) *last_delta_out = delta;
}

// Personalized PageRank: pagerank.rs:63-124's loop with the personalization
// vector p in place of the uniform teleport (and dangling) distribution.
void orc_ppr(orc_topo* t, const uint32_t* seeds, uint64_t nseeds, double damping, uint32_t max_iter, double tol,
             double* out, uint32_t* iters_out) {
    const uint64_t n = t->n;
    if (iters_out) *iters_out = 0;
    if (n == 0) return;
    std::vector<double> p(n, 0.0);
    std::vector<uint32_t> uniq;
    for (uint64_t i = 0; i < nseeds; ++i)
        if (seeds[i] < n) uniq.push_back(seeds[i]);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    for (uint32_t u : uniq) p[u] = 1.0 / (double)uniq.size();
    const Adj& inc = t->incoming();
    std::vector<uint32_t> od(n);
    std::vector<uint64_t> dangling;
    for (uint64_t u = 0; u < n; ++u) {
        od[u] = t->out.degree((uint32_t)u);
        if (od[u] == 0) dangling.push_back(u);
    }
    const double d = damping;
    std::vector<double> rank(n, 1.0 / (double)n), next(n, 0.0);
    uint32_t iters = 0;
    for (uint32_t it = 0; it < max_iter && !uniq.empty(); ++it) {
        double dm = -0.0;
        for (uint64_t u : dangling) dm += rank[u];
        const double pb = (1.0 - d) + d * dm;
        for (uint64_t u = 0; u < n; ++u) {
            double acc = 0.0;
            for (uint64_t k = inc.off[u]; k < inc.off[u + 1]; ++k) acc += rank[inc.tgt[k]] / (double)od[inc.tgt[k]];
            next[u] = d * acc + pb * p[u];
        }
        const double delta = l1_delta(next, rank);
        std::swap(rank, next);
        iters = it + 1;
        if (delta < tol) break;
    }
    std::memcpy(out, rank.data(), n * sizeof(double));
    if (iters_out) *iters_out = iters;
}

// ---- ursa-plan/src/stats.rs:80-98 ----
uint64_t orc_stats_sources(uint64_t n, int has_frac, double frac, uint32_t* out) {
    if (!has_frac) {
        for (uint64_t i = 0; i < n; ++i) out[i] = (uint32_t)i;
        return n;
    }
    const double f = std::min(std::max(frac, 0.0), 1.0);
    uint64_t k = (uint64_t)std::max(std::round((double)n * f), 1.0);
    k = std::min<uint64_t>(k, n);
    if (k == 0) return 0;
    const double step = std::max((double)n / (double)k, 1.0);
    for (uint64_t i = 0; i < k; ++i)
        out[i] = (uint32_t)std::min<uint64_t>((uint64_t)std::floor((double)i * step), n - 1);
    return k;
}

// ---- ursa-plan/src/stats.rs:100-140: per-source partials folded in source order ----
void orc_bfs_source_scan(orc_topo* t, const uint32_t* sources, uint64_t k, int dir, double* sum, uint64_t* count,
                         int32_t* max_dist) {
    std::vector<int32_t> d(t->n);
    double S = 0.0;
    uint64_t C = 0;
    int32_t M = 0;
    for (uint64_t i = 0; i < k; ++i) {
        double s = 0.0;
        uint64_t c = 0;
        int32_t m = 0;
        if (t->n) orc_bfs(t, sources[i], dir, d.data());
        for (uint64_t v = 0; v < t->n; ++v)
            if (d[v] >= 1) {
                s += (double)d[v];
                ++c;
                if (d[v] > m) m = d[v];
            }
        S += s;
        C += c;
        if (m > M) M = m;
    }
    *sum = S;
    *count = C;
    *max_dist = M;
}

// ---- ursa-plan/src/stats.rs:142-152 ----
double orc_avg_path_length(orc_topo* t, int has_sample, double sample) {
    std::vector<uint32_t> src(std::max<uint64_t>(t->n, 1));
    const uint64_t k = orc_stats_sources(t->n, has_sample, sample, src.data());
    double s;
    uint64_t c;
    int32_t m;
    orc_bfs_source_scan(t, src.data(), k, ORC_OUT, &s, &c, &m);
    return c == 0 ? 0.0 : s / (double)c;
}

// ---- ursa-plan/src/stats.rs:154-168 ----
int64_t orc_diameter(orc_topo* t, int approximate) {
    const uint64_t n = t->n;
    std::vector<uint32_t> src(std::max<uint64_t>(n, 1));
    uint64_t k;
    if (approximate && n > 0)
        k = orc_stats_sources(n, 1, std::min(128.0 / (double)n, 1.0), src.data());
    else
        k = orc_stats_sources(n, 0, 0.0, src.data());
    double s;
    uint64_t c;
    int32_t m;
    orc_bfs_source_scan(t, src.data(), k, ORC_OUT, &s, &c, &m);
    return m;
}

// ---- closeness.rs:69-137: exact i64 sums, one f64 division ----
void orc_closeness(orc_topo* t, const uint32_t* sources, uint64_t k, double* out) {
    std::vector<int32_t> d(t->n);
    for (uint64_t i = 0; i < k; ++i) {
        out[i] = 0.0;
        if (sources[i] >= t->n) continue;
        orc_bfs(t, sources[i], ORC_OUT, d.data());
        int64_t sum = 0, reach = 0;
        for (uint64_t v = 0; v < t->n; ++v)
            if (d[v] >= 1) {
                sum += d[v];
                ++reach;
            }
        out[i] = sum > 0 ? (double)reach / (double)sum : 0.0;
    }
}

// ---- closeness.rs:143-222 (weighted, Out) ----
// Dijkstra from each source over out-edges with weight w[edge row]
// (for_each_weighted_neighbor, topology.rs), then the node-order (0..n) f64
// sum of the finite distances of the other nodes (closeness.rs:199-215).
void orc_closeness_weighted(orc_topo* t, const double* w, const uint32_t* sources, uint64_t k, double* out) {
    const uint64_t n = t->n;
    std::vector<double> dist(n);
    std::vector<char> settled(n);
    typedef std::pair<double, uint32_t> Item;  // (distance, node): the reference's heap key
    for (uint64_t i = 0; i < k; ++i) {
        out[i] = 0.0;
        const uint32_t u = sources[i];
        if (u >= n) continue;
        std::fill(dist.begin(), dist.end(), INFINITY);
        std::fill(settled.begin(), settled.end(), 0);
        std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
        dist[u] = 0.0;
        heap.push(Item(0.0, u));
        while (!heap.empty()) {
            const Item top = heap.top();
            heap.pop();
            const double du = top.first;
            const uint32_t v = top.second;
            if (settled[v]) continue;
            settled[v] = 1;
            for (uint64_t e = t->out.off[v]; e < t->out.off[v + 1]; ++e) {
                const uint32_t x = t->out.tgt[e];
                if (settled[x]) continue;
                const double nd = du + w[t->out.eid[e]];
                if (nd < dist[x]) {
                    dist[x] = nd;
                    heap.push(Item(nd, x));
                }
            }
        }
        double sum = 0.0;
        int64_t reach = 0;
        for (uint64_t v = 0; v < n; ++v)
            if (v != u && std::isfinite(dist[v])) {
                sum += dist[v];
                ++reach;
            }
        out[i] = sum > 0.0 ? (double)reach / sum : 0.0;
    }
}

// ---- clustering.rs:21-40 ----
void orc_clustering(orc_topo* t, double* out) {
    const uint64_t n = t->n;
    if (!n) return;
    std::vector<uint64_t> tri(n);
    orc_triangle_count(t, 1, tri.data());
    t->undirected();
    for (uint64_t u = 0; u < n; ++u) {
        const double k = (double)(t->uoff[u + 1] - t->uoff[u]);
        out[u] = k < 2.0 ? 0.0 : 2.0 * (double)tri[u] / (k * (k - 1.0));
    }
}

// ---- bfs.rs:18-42 ----
void orc_bfs(orc_topo* t, uint32_t source, int dir, int32_t* dist) {
    const uint64_t n = t->n;
    for (uint64_t i = 0; i < n; ++i) dist[i] = -1;
    if ((uint64_t)source >= n) return;
    dist[source] = 0;
    std::vector<uint32_t> frontier{source}, next;
    int32_t level = 1;
    while (!frontier.empty()) {
        next.clear();
        for (uint32_t u : frontier) {
            t->for_each_neighbor(u, dir, [&](uint32_t v) {
                if (dist[v] < 0) {
                    dist[v] = level;
                    next.push_back(v);
                }
            });
        }
        level += 1;
        std::swap(frontier, next);
    }
}

// ---- hop.rs:52-130 ----
uint64_t orc_k_hop(orc_topo* t, const uint32_t* seeds, uint64_t nseeds, uint32_t k, int dir,
                   uint32_t* out_src, uint32_t* out_dst, uint64_t cap) {
    const uint64_t n = t->n;
    if (k == 0 || n == 0) return 0;
    std::vector<char> visited(n, 0);
    std::vector<uint32_t> frontier, next, touched;
    uint64_t count = 0;
    for (uint64_t si = 0; si < nseeds; ++si) {
        uint32_t seed = seeds[si];
        if ((uint64_t)seed >= n) continue;
        visited[seed] = 1;
        frontier.assign(1, seed);
        touched.assign(1, seed);
        for (uint32_t level = 0; level < k; ++level) {
            next.clear();
            for (uint32_t u : frontier) {
                t->for_each_neighbor(u, dir, [&](uint32_t v) {
                    if (!visited[v]) {
                        visited[v] = 1;
                        touched.push_back(v);
                        next.push_back(v);
                        if (count < cap) {
                            out_src[count] = seed;
                            out_dst[count] = v;
                        }
                        ++count;
                    }
                });
            }
            if (next.empty()) break;
            std::swap(frontier, next);
        }
        for (uint32_t x : touched) visited[x] = 0;
    }
    return count;
}

// ---- bfs.rs:48-102 ----
uint64_t orc_shortest_path(orc_topo* t, uint32_t s, uint32_t tgt, int dir, uint32_t* out_path,
                           uint64_t cap) {
    const uint64_t n = t->n;
    if ((uint64_t)s >= n || (uint64_t)tgt >= n) return 0;
    if (s == tgt) {
        if (cap >= 1) out_path[0] = s;
        return 1;
    }
    std::vector<int64_t> parent(n, -1);
    parent[s] = s;
    std::vector<uint32_t> frontier{s}, next;
    bool found = false;
    while (!frontier.empty() && !found) {
        next.clear();
        for (uint32_t u : frontier) {
            bool hit = false;
            t->for_each_neighbor(u, dir, [&](uint32_t v) {
                if (parent[v] < 0) {
                    parent[v] = u;
                    if (v == tgt) hit = true;
                    next.push_back(v);
                }
            });
            if (hit) {
                found = true;
                break;
            }
        }
        std::swap(frontier, next);
    }
    if (!found) return 0;
    std::vector<uint32_t> path{tgt};
    uint32_t cur = tgt;
    while (cur != s) {
        cur = (uint32_t)parent[cur];
        path.push_back(cur);
    }
    std::reverse(path.begin(), path.end());
    for (uint64_t i = 0; i < path.size() && i < cap; ++i) out_path[i] = path[i];
    return path.size();
}

// ---- betweenness.rs ----
namespace {
struct BrandesScratch {  // betweenness.rs:104-124
    std::vector<int64_t> dist;
    std::vector<double> sigma, delta;
    std::vector<std::vector<uint32_t>> preds;
    std::vector<uint32_t> order;
    std::deque<uint32_t> queue;
    explicit BrandesScratch(uint64_t n) : dist(n, -1), sigma(n, 0.0), delta(n, 0.0), preds(n) {}
};

// betweenness.rs:185-230 with reverse_accumulate :81-98
void brandes_from(orc_topo* t, uint32_t s, double* bc, BrandesScratch& sc) {
    const Adj& out = t->out;
    sc.dist[s] = 0;
    sc.sigma[s] = 1.0;
    sc.queue.push_back(s);
    while (!sc.queue.empty()) {
        uint32_t v = sc.queue.front();
        sc.queue.pop_front();
        sc.order.push_back(v);
        int64_t dv = sc.dist[v];
        for (uint64_t k = out.off[v]; k < out.off[v + 1]; ++k) {
            uint32_t w = out.tgt[k];
            if (sc.dist[w] < 0) {
                sc.dist[w] = dv + 1;
                sc.queue.push_back(w);
            }
            if (sc.dist[w] == dv + 1) {
                sc.sigma[w] += sc.sigma[v];
                sc.preds[w].push_back(v);
            }
        }
    }
    for (auto it = sc.order.rbegin(); it != sc.order.rend(); ++it) {
        uint32_t w = *it;
        double coeff = (1.0 + sc.delta[w]) / sc.sigma[w];
        for (uint32_t v : sc.preds[w]) sc.delta[v] += sc.sigma[v] * coeff;
        if (w != s) bc[w] += sc.delta[w];
    }
    for (uint32_t w : sc.order) {
        sc.dist[w] = -1;
        sc.sigma[w] = 0.0;
        sc.delta[w] = 0.0;
        sc.preds[w].clear();
    }
    sc.order.clear();
}
}  // namespace

// accumulate_sources, betweenness.rs:45-70: 64-source chunks accumulated in
// source order, chunk partials combined in chunk order. Chunks run on threads;
// the combine order is fixed, so the result is thread-count independent.
void orc_betweenness_sources(orc_topo* t, const uint32_t* sources, uint64_t k, int apply_scale,
                             double scale, int nthreads, double* out) {
    const uint64_t n = t->n;
    const uint64_t SRC_CHUNK = 64;
    uint64_t nchunks = (k + SRC_CHUNK - 1) / SRC_CHUNK;
    std::vector<std::vector<double>> partials(nchunks);
    int nt = std::max(1, std::min<int>(resolve_threads(nthreads), (int)std::max<uint64_t>(nchunks, 1)));
    std::vector<std::thread> ts;
    for (int w = 0; w < nt; ++w) {
        ts.emplace_back([&, w] {
            BrandesScratch sc(n);
            for (uint64_t c = w; c < nchunks; c += nt) {
                std::vector<double> acc(n, 0.0);
                for (uint64_t i = c * SRC_CHUNK; i < std::min(k, (c + 1) * SRC_CHUNK); ++i) {
                    if ((uint64_t)sources[i] < n) brandes_from(t, sources[i], acc.data(), sc);
                }
                partials[c] = std::move(acc);
            }
        });
    }
    for (auto& th : ts) th.join();
    for (uint64_t i = 0; i < n; ++i) out[i] = 0.0;
    for (auto& p : partials)
        for (uint64_t i = 0; i < n; ++i) out[i] += p[i];
    if (apply_scale)
        for (uint64_t i = 0; i < n; ++i) out[i] = out[i] * scale;
}

// betweenness.rs:153-177
void orc_betweenness(orc_topo* t, int has_frac, double frac, int has_seed, uint64_t seed,
                     int nthreads, double* out) {
    const uint64_t n = t->n;
    if (n == 0) return;
    std::vector<uint32_t> sources(n);
    uint64_t k = orc_sample_sources(n, has_frac, frac, has_seed, seed, sources.data());
    if (k == 0) {
        for (uint64_t i = 0; i < n; ++i) out[i] = 0.0;
        return;
    }
    orc_betweenness_sources(t, sources.data(), k, has_frac, (double)n / (double)k, nthreads, out);
}

// ---- triangle.rs:23-63 (u64 accumulation) ----
void orc_triangle_count(orc_topo* t, int nthreads, uint64_t* out) {
    const uint64_t n = t->n;
    if (n == 0) return;
    t->undirected();
    const uint64_t* off = t->uoff.data();
    const uint32_t* tg = t->utgt.data();
    parallel_ranges(n, nthreads, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t u = lo; u < hi; ++u) {
            uint64_t count = 0;
            for (uint64_t i = off[u]; i < off[u + 1]; ++i) {
                uint32_t a = tg[i];
                // intersection_count(nu[i+1..], N(a)) — triangle.rs:23-37
                uint64_t p = i + 1, pe = off[u + 1], q = off[a], qe = off[a + 1];
                while (p < pe && q < qe) {
                    if (tg[p] < tg[q]) ++p;
                    else if (tg[p] > tg[q]) ++q;
                    else { ++count; ++p; ++q; }
                }
            }
            out[u] = count;
        }
    });
}

// ---- components.rs:14-68 ----
void orc_cc_weak(orc_topo* t, uint32_t* out) {
    const uint64_t n = t->n;
    std::vector<uint32_t> parent(n);
    std::vector<uint8_t> rnk(n, 0);
    for (uint64_t i = 0; i < n; ++i) parent[i] = (uint32_t)i;
    auto find = [&](uint32_t x) {  // path halving, components.rs:26-34
        while (parent[x] != x) {
            uint32_t gp = parent[parent[x]];
            parent[x] = gp;
            x = gp;
        }
        return x;
    };
    for (uint64_t u = 0; u < n; ++u) {
        for (uint64_t k = t->out.off[u]; k < t->out.off[u + 1]; ++k) {
            uint32_t ra = find((uint32_t)u), rb = find(t->out.tgt[k]);  // union :36-52
            if (ra == rb) continue;
            uint8_t rar = rnk[ra], rbr = rnk[rb];
            uint32_t small = rar < rbr ? ra : rb, big = rar < rbr ? rb : ra;
            parent[small] = big;
            if (rar == rbr) rnk[big] += 1;
        }
    }
    for (uint64_t u = 0; u < n; ++u) out[u] = find((uint32_t)u);
}

void orc_canonicalize_labels(uint64_t n, const uint32_t* labels, uint32_t* out) {
    std::unordered_map<uint32_t, uint32_t> minid;
    for (uint64_t u = 0; u < n; ++u) {
        auto it = minid.find(labels[u]);
        if (it == minid.end()) minid.emplace(labels[u], (uint32_t)u);  // ascending u: first is min
    }
    for (uint64_t u = 0; u < n; ++u) out[u] = minid[labels[u]];
}

// ---- workload generators ----
// Counter form of splitmix64: the t-th output (1-based) of a stream seeded
// with `seed` is mix(seed + t * gamma), so any draw is computable directly.
static inline uint64_t stream_at(uint64_t seed, uint64_t t) { return splitmix_mix(seed + t * kGamma); }

static uint64_t dedup_first(std::vector<int64_t>& s, std::vector<int64_t>& d) {
    std::unordered_set<uint64_t> seen;
    seen.reserve(s.size() * 2 + 1);
    uint64_t w = 0;
    for (uint64_t i = 0; i < s.size(); ++i) {
        uint64_t key = ((uint64_t)s[i] << 32) | (uint64_t)(uint32_t)d[i];
        if (seen.insert(key).second) {
            s[w] = s[i];
            d[w] = d[i];
            ++w;
        }
    }
    s.resize(w);
    d.resize(w);
    return w;
}

static uint64_t finish_edges(std::vector<int64_t>& s, std::vector<int64_t>& d, int symmetrize,
                             int64_t* src, int64_t* dst, uint64_t cap) {
    dedup_first(s, d);
    if (symmetrize) {
        uint64_t m = s.size();
        s.reserve(2 * m);
        d.reserve(2 * m);
        for (uint64_t i = 0; i < m; ++i) {
            s.push_back(d[i]);
            d.push_back(s[i]);
        }
        dedup_first(s, d);
    }
    uint64_t m = s.size();
    for (uint64_t i = 0; i < m && i < cap; ++i) {
        src[i] = s[i];
        dst[i] = d[i];
    }
    return m;
}

uint64_t orc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int symmetrize, int64_t* src, int64_t* dst, uint64_t cap) {
    const uint64_t draws = (uint64_t)edge_factor << scale;
    const double ab = a + b, abc = a + b + c;
    std::vector<int64_t> s, d;
    s.reserve(draws);
    d.reserve(draws);
    for (uint64_t e = 0; e < draws; ++e) {
        uint64_t u = 0, v = 0;
        for (uint32_t j = 0; j < scale; ++j) {
            uint64_t x = stream_at(seed, e * scale + j + 1);
            double r = (double)(x >> 11) * 0x1.0p-53;
            uint64_t bu = 0, bv = 0;
            if (r < a) {
            } else if (r < ab) {
                bv = 1;
            } else if (r < abc) {
                bu = 1;
            } else {
                bu = 1;
                bv = 1;
            }
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        if (u == v) continue;  // self-loop dropped
        s.push_back((int64_t)u);
        d.push_back((int64_t)v);
    }
    return finish_edges(s, d, symmetrize, src, dst, cap);
}

uint64_t orc_gen_er(uint64_t n, uint32_t edge_factor, uint64_t seed, int symmetrize, int64_t* src,
                    int64_t* dst, uint64_t cap) {
    const uint64_t draws = (uint64_t)edge_factor * n;
    std::vector<int64_t> s, d;
    for (uint64_t e = 0; e < draws && n > 0; ++e) {
        uint64_t u = stream_at(seed, 2 * e + 1) % n;
        uint64_t v = stream_at(seed, 2 * e + 2) % n;
        if (u == v) continue;
        s.push_back((int64_t)u);
        d.push_back((int64_t)v);
    }
    return finish_edges(s, d, symmetrize, src, dst, cap);
}

void orc_gen_weights(uint64_t m, uint64_t seed, double* out) {
    const uint64_t ws = seed ^ 0x5745494748540000ull;
    for (uint64_t e = 0; e < m; ++e) {
        uint64_t x = stream_at(ws, e + 1);
        out[e] = (double)((x >> 40) + 1) * 0x1.0p-24;
    }
}

}  // extern "C"
