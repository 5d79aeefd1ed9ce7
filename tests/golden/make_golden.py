"""Generate the committed golden fixtures in this directory.

Run from the repository root:  python tests/golden/make_golden.py

The expected values come from independent implementations — networkx 3.4.2 (the
reference's own cross-check oracle, tests/test_networkx_reference.py:33-46) and
scipy.sparse.csgraph — never from this project's oracle or product code. Each
graph is stored as its int64 edge list in row order (the order fixes ursa's
first-seen dense numbering, ursa-core/src/id_map.rs:220-292) together with the
expected per-node values keyed by user id.

Also writes the first outputs of the splitmix64 stream for a set of seeds from a
direct transcription of the published splitmix64 (Steele, Lea & Flood 2014;
Vigna's reference C), cross-checked in the test against the widely published
seed-1234567 vector.
"""
from __future__ import annotations

import os

import networkx as nx
import numpy as np
import scipy.sparse as sp
from scipy.sparse import csgraph

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> list[int]:
    out, s = [], seed & M64
    for _ in range(count):
        s = (s + 0x9E3779B97F4A7C15) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def edge_arrays(edges):
    s = np.array([int(a) for a, _ in edges], dtype=np.int64)
    d = np.array([int(b) for _, b in edges], dtype=np.int64)
    return s, d


def keyed(d: dict):
    keys = np.array(sorted(d), dtype=np.int64)
    vals = np.array([d[k] for k in keys])
    return keys, vals


def bfs_levels(s, d, nodes, sources, directed=True):
    """scipy BFS hop distances, -1 for unreachable, over nodes in `nodes` order."""
    idx = {int(v): i for i, v in enumerate(nodes)}
    r = np.array([idx[int(x)] for x in s])
    c = np.array([idx[int(x)] for x in d])
    n = len(nodes)
    a = sp.csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    out = []
    for src in sources:
        dist = csgraph.shortest_path(a, directed=directed, unweighted=True, indices=idx[int(src)])
        lv = np.where(np.isinf(dist), -1, dist).astype(np.int32)
        out.append(lv)
    return np.stack(out)


def main():
    fx = {}

    # 1) The reference's own cross-check graph (tests/test_networkx_reference.py:33-46).
    g = nx.gnp_random_graph(100, 0.05, seed=42, directed=True)
    edges = list(g.edges())
    s, d = edge_arrays(edges)
    gd = nx.DiGraph(); gd.add_edges_from(edges)
    gu = nx.Graph(); gu.add_edges_from(edges)
    fx["gnp100_src"], fx["gnp100_dst"] = s, d
    fx["gnp100_pr_ids"], fx["gnp100_pr"] = keyed(nx.pagerank(gd, alpha=0.85, tol=1e-14, max_iter=10000))
    fx["gnp100_bc_ids"], fx["gnp100_bc"] = keyed(nx.betweenness_centrality(gd, normalized=False, endpoints=False))
    fx["gnp100_tri_ids"], fx["gnp100_tri"] = keyed(nx.triangles(gu))
    comp = {}
    for ci, cset in enumerate(sorted(nx.weakly_connected_components(gd), key=min)):
        for v in cset:
            comp[v] = ci
    fx["gnp100_wcc_ids"], fx["gnp100_wcc"] = keyed(comp)
    nodes = np.array(sorted(gd.nodes()), dtype=np.int64)
    srcs = nodes[[0, 7, 33, 50]]
    fx["gnp100_bfs_nodes"] = nodes
    fx["gnp100_bfs_sources"] = srcs
    fx["gnp100_bfs_out"] = bfs_levels(s, d, nodes, srcs, directed=True)
    fx["gnp100_bfs_both"] = bfs_levels(s, d, nodes, srcs, directed=False)

    # 2) Weighted PageRank: nx uses weight / weighted out-degree and uniform dangling
    #    redistribution, the same definition as pagerank.rs:135-169.
    g2 = nx.gnp_random_graph(300, 0.02, seed=7, directed=True)
    e2 = list(g2.edges())
    rng = np.random.default_rng(11)
    w2 = rng.uniform(0.05, 1.0, size=len(e2))
    gw = nx.DiGraph()
    for (a, b), w in zip(e2, w2):
        gw.add_edge(a, b, weight=float(w))
    s2, d2 = edge_arrays(e2)
    fx["gnp300w_src"], fx["gnp300w_dst"], fx["gnp300w_w"] = s2, d2, w2
    fx["gnp300w_pr_ids"], fx["gnp300w_pr"] = keyed(nx.pagerank(gw, alpha=0.85, tol=1e-14, max_iter=10000, weight="weight"))
    fx["gnp300w_upr_ids"], fx["gnp300w_upr"] = keyed(nx.pagerank(gw, alpha=0.85, tol=1e-14, max_iter=10000, weight=None))

    # 3) Sparse directed graph with many dangling nodes and several components.
    g3 = nx.gnp_random_graph(400, 0.004, seed=5, directed=True)
    e3 = list(g3.edges())
    s3, d3 = edge_arrays(e3)
    g3d = nx.DiGraph(); g3d.add_edges_from(e3)
    fx["sparse400_src"], fx["sparse400_dst"] = s3, d3
    fx["sparse400_pr_ids"], fx["sparse400_pr"] = keyed(nx.pagerank(g3d, alpha=0.85, tol=1e-14, max_iter=10000))
    comp = {}
    for ci, cset in enumerate(sorted(nx.weakly_connected_components(g3d), key=min)):
        for v in cset:
            comp[v] = ci
    fx["sparse400_wcc_ids"], fx["sparse400_wcc"] = keyed(comp)
    fx["sparse400_bc_ids"], fx["sparse400_bc"] = keyed(nx.betweenness_centrality(g3d, normalized=False, endpoints=False))

    # 4) Undirected power-law graph (both directions stored as rows) for triangles.
    g4 = nx.barabasi_albert_graph(500, 4, seed=3)
    e4 = []
    for a, b in g4.edges():
        e4.append((a, b))
        e4.append((b, a))
    s4, d4 = edge_arrays(e4)
    fx["ba500_src"], fx["ba500_dst"] = s4, d4
    fx["ba500_tri_ids"], fx["ba500_tri"] = keyed(nx.triangles(g4))
    g4d = nx.DiGraph(); g4d.add_edges_from(e4)
    fx["ba500_bc_ids"], fx["ba500_bc"] = keyed(nx.betweenness_centrality(g4d, normalized=False, endpoints=False))

    # 5) Personalized PageRank (networkx personalization: teleport and dangling
    #    mass to the seed set) on the dangling-heavy sparse graph and gnp100.
    for name, gg in (("sparse400", g3d), ("gnp100", gd)):
        nodes_sorted = sorted(gg.nodes())
        for qi, seeds in enumerate(([nodes_sorted[0]], nodes_sorted[3:9:2], nodes_sorted[-5:])):
            pers = {v: 1.0 for v in seeds}
            ref = nx.pagerank(gg, alpha=0.85, personalization=pers, tol=1e-14, max_iter=10000)
            fx[f"{name}_ppr{qi}_seeds"] = np.array(seeds, dtype=np.int64)
            fx[f"{name}_ppr{qi}_ids"], fx[f"{name}_ppr{qi}"] = keyed(ref)

    # 6) Local clustering (nx.clustering on the simple undirected graph) and
    #    closeness over out-edges: nx measures distances *to* a node, so the
    #    reversed graph gives distances from it; wf_improved=False is
    #    reachable / sum of distances (closeness.rs:61-68).
    fx["gnp100_clust_ids"], fx["gnp100_clust"] = keyed(nx.clustering(gu))
    fx["ba500_clust_ids"], fx["ba500_clust"] = keyed(nx.clustering(g4))
    fx["gnp100_close_ids"], fx["gnp100_close"] = keyed(nx.closeness_centrality(gd.reverse(), wf_improved=False))
    fx["sparse400_close_ids"], fx["sparse400_close"] = keyed(
        nx.closeness_centrality(g3d.reverse(), wf_improved=False))
    # weighted closeness (closeness.rs:143-222): Dijkstra distances over
    # seeded per-row weights, same reversal and wf_improved=False
    for name, gg, ss, dd in (("gnp100", gd, s, d), ("sparse400", g3d, s3, d3)):
        wr = np.random.default_rng(77).random(len(ss)) + 0.05
        gw2 = nx.DiGraph()
        gw2.add_nodes_from(gg.nodes())
        for a_, b_, x_ in zip(ss.tolist(), dd.tolist(), wr.tolist()):
            gw2.add_edge(a_, b_, weight=x_)
        fx[f"{name}_wclose_w"] = wr
        fx[f"{name}_wclose_ids"], fx[f"{name}_wclose"] = keyed(
            nx.closeness_centrality(gw2.reverse(), distance="weight", wf_improved=False))

    # splitmix64 streams
    seeds = np.array([0, 4, 7, 42, 0x5EED5EED5EED5EED, 1234567], dtype=np.uint64)
    fx["splitmix_seeds"] = seeds
    fx["splitmix_out"] = np.array([splitmix64(int(x), 16) for x in seeds], dtype=np.uint64)

    np.savez_compressed(os.path.join(HERE, "nx_fixtures.npz"), **fx)
    print("wrote", len(fx), "arrays")


if __name__ == "__main__":
    main()
