"""Pins the CPU oracle before anything is compared against it.

Each check is independent of the oracle's own code: the reference's unit-test
golden values (cited file:line), networkx / scipy fixtures committed under
tests/golden/ (generated by tests/golden/make_golden.py), closed forms, and brute
force on tiny graphs.
"""
import itertools
import os

import numpy as np
import pytest

import oracle as orc

FX = np.load(os.path.join(os.path.dirname(__file__), "golden", "nx_fixtures.npz"))


def topo_from_user(src, dst):
    n, sd, dd, users = orc.intern_i64(src, dst)
    return orc.Topology(n, sd, dd), users


def by_user(users, values):
    return {int(u): v for u, v in zip(users, values)}


# ---------------------------------------------------------------- topology
def test_out_adjacency_groups_by_source():  # topology.rs:472-479
    t = orc.Topology(3, [0, 0, 1, 2], [1, 2, 2, 0])
    off, tg, ei = t.adjacency("out")
    assert list(off) == [0, 2, 3, 4]
    assert list(tg[0:2]) == [1, 2] and list(tg[2:3]) == [2] and list(tg[3:4]) == [0]
    assert sorted(ei[0:2]) == [0, 1]  # topology.rs:492-498


def test_transpose_matches_golden():  # topology.rs:481-489, :516-532
    t = orc.Topology(3, [0, 0, 1, 2], [1, 2, 2, 0])
    off, tg, _ = t.adjacency("in")
    nb = lambda u: list(tg[off[u]:off[u + 1]])
    assert sorted(nb(2)) == [0, 1] and nb(0) == [2] and nb(1) == [0]


def test_undirected_sorted_dedup_no_self_loops():  # topology.rs:501-513
    t = orc.Topology(3, [0, 0, 1, 2, 1, 0], [1, 2, 2, 0, 1, 1])
    off, tg = t.undirected()
    nb = lambda u: list(tg[off[u]:off[u + 1]])
    assert nb(0) == [1, 2] and nb(1) == [0, 2] and nb(2) == [0, 1]


@pytest.mark.parametrize("n,m,seed", [(50, 2000, 1), (2000, 30000, 2), (16, 8000, 3)])
def test_counting_sort_is_stable_grouping(n, m, seed):
    """CSR == stable sort of rows by key (the byte-identity contract, topology.rs:109-114)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, n, m, dtype=np.uint32)
    o = rng.integers(0, n, m, dtype=np.uint32)
    t = orc.Topology(n, k, o)
    off, tg, ei = t.adjacency("out")
    perm = np.argsort(k, kind="stable")
    assert np.array_equal(ei, perm.astype(np.uint32))
    assert np.array_equal(tg, o[perm])
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(k, minlength=n))]).astype(np.uint64))
    ioff, itg, iei = t.adjacency("in")
    perm_in = np.argsort(o, kind="stable")
    assert np.array_equal(iei, perm_in.astype(np.uint32))
    assert np.array_equal(itg, k[perm_in])
    uoff, utg = t.undirected()
    for u in range(0, n, max(1, n // 40)):
        ref = sorted((set(o[k == u].tolist()) | set(k[o == u].tolist())) - {u})
        assert list(utg[uoff[u]:uoff[u + 1]]) == ref


# ---------------------------------------------------------------- degree
def test_degree_golden():  # degree.rs:35-41
    t = orc.Topology(3, [0, 0, 1, 2], [1, 2, 2, 0])
    assert list(orc.degree(t, "out")) == [2, 1, 1]
    assert list(orc.degree(t, "in")) == [1, 1, 2]
    assert list(orc.degree(t, "both")) == [3, 2, 3]


# ---------------------------------------------------------------- pagerank
def test_pagerank_sums_to_one_hub_highest():  # pagerank.rs:176-185
    t = orc.Topology(4, [0, 1, 3, 2], [2, 2, 2, 0])
    pr, _, _ = orc.pagerank(t)
    assert abs(pr.sum() - 1.0) < 1e-6
    assert all(pr[i] < pr[2] for i in range(4) if i != 2)


def test_pagerank_weighted_goldens():  # pagerank.rs:194-219
    t = orc.Topology(3, [0, 0, 1, 2], [1, 2, 0, 0])
    u, _, _ = orc.pagerank(t, weights=[1.0, 1.0, 1.0, 1.0])
    assert abs(u[1] - u[2]) < 1e-9
    w, _, _ = orc.pagerank(t, weights=[1.0, 9.0, 1.0, 1.0])
    assert w[2] > w[1] and abs(w.sum() - 1.0) < 1e-6
    t2 = orc.Topology(2, [0], [1])
    z, _, _ = orc.pagerank(t2, weights=[0.0])
    assert abs(z.sum() - 1.0) < 1e-6


def test_pagerank_empty_and_zero_iter():
    t = orc.Topology(0, [], [])
    assert orc.pagerank(t)[0].size == 0
    t = orc.Topology(3, [0, 1], [1, 2])
    pr, it, _ = orc.pagerank(t, max_iter=0)
    assert it == 0 and np.all(pr == 1.0 / 3.0)


def test_pagerank_cycle_closed_form():
    n = 37
    t = orc.Topology(n, list(range(n)), [(i + 1) % n for i in range(n)])
    pr, it, _ = orc.pagerank(t, max_iter=100, tol=1e-15)
    assert np.allclose(pr, 1.0 / n, rtol=0, atol=1e-15)


def test_pagerank_star_closed_form():
    # k leaves -> hub, hub dangling. Fixed point: leaf = (1-d)/n + d*h/n,
    # hub = leaf + d*k*leaf  with n = k+1 nodes and h = hub rank.
    k, d = 9, 0.85
    n = k + 1
    t = orc.Topology(n, list(range(1, n)), [0] * k)
    pr, _, _ = orc.pagerank(t, damping=d, max_iter=1000, tol=1e-16)
    # Solve the 2x2 linear system for (leaf, hub).
    A = np.array([[1.0, -d / n], [-(1 + d * k), 1.0 - 0.0]])
    # hub = leaf*(1 + d*k) exactly since hub's own base equals a leaf's.
    leaf = (1 - d) / n / (1 - d / n * (1 + d * k))
    hub = leaf * (1 + d * k)
    assert abs(pr[0] - hub) < 1e-13 and np.allclose(pr[1:], leaf, atol=1e-13)
    del A


@pytest.mark.parametrize("name", ["gnp100", "sparse400"])
def test_pagerank_vs_networkx(name):
    t, users = topo_from_user(FX[f"{name}_src"], FX[f"{name}_dst"])
    pr, it, _ = orc.pagerank(t, max_iter=10000, tol=1e-15)
    got = by_user(users, pr)
    ref = dict(zip(FX[f"{name}_pr_ids"].tolist(), FX[f"{name}_pr"].tolist()))
    assert set(got) == set(ref)
    for k, v in ref.items():
        assert abs(got[k] - v) < 1e-11


def test_pagerank_weighted_vs_networkx():
    t, users = topo_from_user(FX["gnp300w_src"], FX["gnp300w_dst"])
    pr, _, _ = orc.pagerank(t, max_iter=10000, tol=1e-15, weights=FX["gnp300w_w"])
    got = by_user(users, pr)
    for k, v in zip(FX["gnp300w_pr_ids"].tolist(), FX["gnp300w_pr"].tolist()):
        assert abs(got[k] - v) < 1e-11
    pr, _, _ = orc.pagerank(t, max_iter=10000, tol=1e-15)
    got = by_user(users, pr)
    for k, v in zip(FX["gnp300w_upr_ids"].tolist(), FX["gnp300w_upr"].tolist()):
        assert abs(got[k] - v) < 1e-11


def test_pagerank_thread_count_independent():  # ursa-core/tests/determinism_threads.rs:51-66
    rng = np.random.default_rng(5)
    s = rng.integers(0, 3000, 40000)
    d = rng.integers(0, 3000, 40000)
    t, _ = topo_from_user(s, d)
    ref = orc.pagerank(t, max_iter=100, tol=1e-12, nthreads=1)
    for th in (2, 3, 8):
        got = orc.pagerank(t, max_iter=100, tol=1e-12, nthreads=th)
        assert np.array_equal(ref[0], got[0]) and ref[1] == got[1]


# ---------------------------------------------------------------- bfs / paths
def path5():
    return orc.Topology(5, [0, 1, 2, 1], [1, 2, 3, 4])  # bfs.rs:262-264


def test_bfs_goldens():  # bfs.rs:266-298
    assert list(orc.bfs(path5(), 0, "out")) == [0, 1, 2, 3, 2]
    d = orc.bfs(path5(), 3, "out")
    assert d[3] == 0 and d[0] == -1 and d[1] == -1
    d = orc.bfs(path5(), 3, "both")
    assert d[0] == 3 and d[4] == 3
    assert list(orc.bfs(path5(), 99, "out")) == [-1] * 5


@pytest.mark.parametrize("direction,key", [("out", "gnp100_bfs_out"), ("both", "gnp100_bfs_both")])
def test_bfs_vs_scipy(direction, key):
    t, users = topo_from_user(FX["gnp100_src"], FX["gnp100_dst"])
    dense = {int(u): i for i, u in enumerate(users)}
    nodes = FX["gnp100_bfs_nodes"]
    for srow, src in enumerate(FX["gnp100_bfs_sources"]):
        got = orc.bfs(t, dense[int(src)], direction)
        exp = FX[key][srow]
        for j, v in enumerate(nodes):
            assert got[dense[int(v)]] == exp[j]


def test_shortest_path_goldens():  # bfs.rs:301-332
    t = path5()
    assert list(orc.shortest_path(t, 0, 3)) == [0, 1, 2, 3]
    assert list(orc.shortest_path(t, 0, 4)) == [0, 1, 4]
    assert list(orc.shortest_path(t, 2, 2)) == [2]
    assert orc.shortest_path(t, 3, 0) is None
    assert orc.shortest_path(t, 0, 99) is None
    assert list(orc.shortest_path(t, 3, 0, "in")) == [3, 2, 1, 0]
    t2 = orc.Topology(4, [0, 1, 0], [1, 3, 3])  # bfs.rs:336-339
    assert list(orc.shortest_path(t2, 0, 3)) == [0, 3]


def test_shortest_path_is_a_shortest_walk():
    t, _ = topo_from_user(FX["gnp100_src"], FX["gnp100_dst"])
    off, tg, _ = t.adjacency("out")
    edges = set()
    for u in range(t.n):
        for v in tg[off[u]:off[u + 1]]:
            edges.add((u, int(v)))
    for s in range(0, t.n, 9):
        lv = orc.bfs(t, s)
        for g in range(0, t.n, 7):
            p = orc.shortest_path(t, s, g)
            if lv[g] < 0:
                assert p is None
                continue
            assert len(p) == lv[g] + 1 and p[0] == s and p[-1] == g
            assert all((int(a), int(b)) in edges for a, b in zip(p[:-1], p[1:]))


def test_k_hop_goldens():  # hop.rs:148-191
    t = path5()
    pairs = lambda r: sorted(zip(r[0].tolist(), r[1].tolist()))
    assert pairs(orc.k_hop(t, [0], 1)) == [(0, 1)]
    assert pairs(orc.k_hop(t, [0], 2)) == [(0, 1), (0, 2), (0, 4)]
    assert pairs(orc.k_hop(t, [0], 10)) == [(0, 1), (0, 2), (0, 3), (0, 4)]
    assert pairs(orc.k_hop(t, [3], 2, "in")) == [(3, 1), (3, 2)]
    assert pairs(orc.k_hop(t, [0, 2], 1)) == [(0, 1), (2, 3)]
    assert orc.k_hop(t, [99], 3)[0].size == 0 and orc.k_hop(t, [0], 0)[0].size == 0


def test_k_hop_equals_bfs_ball():
    t, _ = topo_from_user(FX["sparse400_src"], FX["sparse400_dst"])
    seeds = [0, 5, 5, 17, 100]
    a, b = orc.k_hop(t, seeds, 3, "both")
    pos = 0
    for s in seeds:
        lv = orc.bfs(t, s, "both")
        ball = sorted(np.nonzero((lv >= 1) & (lv <= 3))[0].tolist())
        cnt = len(ball)
        assert list(a[pos:pos + cnt]) == [s] * cnt
        assert sorted(b[pos:pos + cnt].tolist()) == ball
        # discovery order is by level
        assert list(lv[b[pos:pos + cnt]]) == sorted(lv[b[pos:pos + cnt]])
        pos += cnt
    assert pos == a.size


# ---------------------------------------------------------------- betweenness
def test_betweenness_goldens():  # betweenness.rs:369-387, :421-460
    t = orc.Topology(5, [0, 1, 2, 3], [1, 2, 3, 4])
    assert list(orc.betweenness(t)) == [0.0, 3.0, 4.0, 3.0, 0.0]
    t = orc.Topology(4, [0, 0, 1, 2], [1, 2, 3, 3])
    bc = orc.betweenness(t)
    assert bc[0] == 0 and bc[3] == 0 and abs(bc[1] - 0.5) < 1e-12 and abs(bc[2] - 0.5) < 1e-12
    n = 80
    t = orc.Topology(n, list(range(n)), [(i + 1) % n for i in range(n)])
    bc = orc.betweenness(t)
    assert np.all(np.abs(bc - bc[0]) < 1e-9) and bc[0] > 0
    # closed form for a directed n-cycle: each node carries sum_{L=2}^{n-1} (L-1)
    assert abs(bc[0] - sum(L - 1 for L in range(2, n))) < 1e-9
    t = orc.Topology(7, [0, 1, 2, 3, 4, 5, 6, 0, 2], [1, 2, 3, 4, 5, 6, 0, 3, 5])
    assert np.array_equal(orc.betweenness(t, 0.5, 7), orc.betweenness(t, 0.5, 7))
    assert not np.array_equal(orc.sample_sources(7, 0.5, 1), orc.sample_sources(7, 0.5, 2))
    bc = orc.betweenness(orc.Topology(5, [0, 1, 2, 3], [1, 2, 3, 4]), 0.5, 7)
    assert np.all(bc >= 0) and bc[2] >= bc[0]


@pytest.mark.parametrize("name", ["gnp100", "sparse400", "ba500"])
def test_betweenness_vs_networkx(name):
    t, users = topo_from_user(FX[f"{name}_src"], FX[f"{name}_dst"])
    got = by_user(users, orc.betweenness(t))
    for k, v in zip(FX[f"{name}_bc_ids"].tolist(), FX[f"{name}_bc"].tolist()):
        assert abs(got[k] - v) <= 1e-9 * max(1.0, abs(v))


def test_betweenness_multigraph_counts_parallel_edges():
    # 0->1 twice, 0->2 once, 1->3, 2->3: sigma(3) = 3 paths via CSR slots, so node 1
    # carries 2/3 and node 2 carries 1/3 of the (0,3) pair (betweenness.rs:9-13).
    t = orc.Topology(4, [0, 0, 0, 1, 2], [1, 1, 2, 3, 3])
    bc = orc.betweenness(t)
    assert abs(bc[1] - 2.0 / 3.0) < 1e-15 and abs(bc[2] - 1.0 / 3.0) < 1e-15


def test_betweenness_thread_count_independent():  # determinism_threads.rs:69-94
    rng = np.random.default_rng(9)
    t, _ = topo_from_user(rng.integers(0, 300, 2000), rng.integers(0, 300, 2000))
    ref = orc.betweenness(t, nthreads=1)
    for th in (2, 3, 8):
        assert np.array_equal(ref, orc.betweenness(t, nthreads=th))
    ref = orc.betweenness(t, 0.3, 11, nthreads=1)
    assert np.array_equal(ref, orc.betweenness(t, 0.3, 11, nthreads=4))


# ---------------------------------------------------------------- triangles
def test_triangle_goldens():  # triangle.rs:70-99
    assert list(orc.triangle_count(orc.Topology(3, [0, 0, 1, 2], [1, 2, 2, 0]))) == [1, 1, 1]
    assert list(orc.triangle_count(orc.Topology(3, [0, 1], [1, 2]))) == [0, 0, 0]
    assert list(orc.triangle_count(orc.Topology(4, [0, 0, 0, 1, 1, 2], [1, 2, 3, 2, 3, 3]))) == [3, 3, 3, 3]
    assert list(orc.triangle_count(orc.Topology(3, [0, 0, 1, 2, 0, 0], [1, 2, 2, 0, 0, 1]))) == [1, 1, 1]


@pytest.mark.parametrize("name", ["gnp100", "ba500"])
def test_triangles_vs_networkx(name):
    t, users = topo_from_user(FX[f"{name}_src"], FX[f"{name}_dst"])
    got = by_user(users, orc.triangle_count(t))
    for k, v in zip(FX[f"{name}_tri_ids"].tolist(), FX[f"{name}_tri"].tolist()):
        assert got[k] == v


def test_triangles_clique_with_satellites_closed_form():
    # K_60 plus 400 satellites, each joined to k in {1, 2, 5, 10} distinct
    # clique vertices: a satellite is in C(k, 2) triangles, a clique vertex in
    # C(59, 2) plus k - 1 per satellite joined to it.
    rng = np.random.default_rng(5)
    c, ns = 60, 400
    iu, ju = np.triu_indices(c, 1)
    src, dst = [iu], [ju]
    want = np.zeros(c + ns, np.uint64)
    want[:c] = (c - 1) * (c - 2) // 2
    for x in range(ns):
        k = int(rng.choice([1, 2, 5, 10]))
        nb = rng.choice(c, size=k, replace=False)
        # both orientations of the satellite edges, so the undirected view dedups them
        src += [np.full(k, c + x), nb]
        dst += [nb, np.full(k, c + x)]
        want[c + x] = k * (k - 1) // 2
        want[nb] += k - 1
    s = np.concatenate(src).astype(np.uint32)
    d = np.concatenate(dst).astype(np.uint32)
    assert np.array_equal(orc.triangle_count(orc.Topology(c + ns, s, d)), want)


def test_triangles_brute_force():
    rng = np.random.default_rng(3)
    n = 40
    s = rng.integers(0, n, 300)
    d = rng.integers(0, n, 300)
    t = orc.Topology(n, s, d)
    adj = np.zeros((n, n), bool)
    adj[s, d] = True
    adj[d, s] = True
    np.fill_diagonal(adj, False)
    exp = np.zeros(n, np.uint64)
    for a, b, c in itertools.combinations(range(n), 3):
        if adj[a, b] and adj[b, c] and adj[a, c]:
            exp[[a, b, c]] += 1
    assert np.array_equal(orc.triangle_count(t), exp)
    assert int(exp.sum()) % 3 == 0


# ---------------------------------------------------------------- components
def test_cc_golden_partition():  # components.rs:145-153
    cc = orc.cc_weak(orc.Topology(5, [0, 1, 3], [1, 2, 4]))
    assert cc[0] == cc[1] == cc[2] and cc[3] == cc[4] and cc[0] != cc[3]
    assert list(orc.canonicalize(cc)) == [0, 0, 0, 3, 3]


@pytest.mark.parametrize("name", ["gnp100", "sparse400"])
def test_cc_vs_networkx(name):
    t, users = topo_from_user(FX[f"{name}_src"], FX[f"{name}_dst"])
    lab = orc.canonicalize(orc.cc_weak(t))
    ref = dict(zip(FX[f"{name}_wcc_ids"].tolist(), FX[f"{name}_wcc"].tolist()))
    dense_of = {int(u): i for i, u in enumerate(users)}
    # same partition: label equality <=> reference component equality
    groups = {}
    for u, c in ref.items():
        groups.setdefault(c, []).append(dense_of[u])
    for members in groups.values():
        assert len({int(lab[x]) for x in members}) == 1
        assert int(lab[members[0]]) == min(members)
    assert len(set(lab.tolist())) == len(groups)


# ---------------------------------------------------------------- rng / ids
def test_splitmix_published_vector():
    # Published splitmix64 outputs for seed 1234567 (Vigna's splitmix64.c).
    got = orc.splitmix_stream(1234567, 5).tolist()
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_splitmix_fixture_streams():
    for seed, row in zip(FX["splitmix_seeds"].tolist(), FX["splitmix_out"]):
        assert np.array_equal(orc.splitmix_stream(int(seed), 16), row)


def test_shuffle_and_sample_properties():  # rng.rs:80-121
    o = orc.shuffled_order(100, 42)
    assert sorted(o.tolist()) == list(range(100))
    assert np.array_equal(orc.shuffled_order(50, 7), orc.shuffled_order(50, 7))
    assert not np.array_equal(orc.shuffled_order(50, 7), orc.shuffled_order(50, 8))
    s = orc.sample_indices(100, 10, 7)
    assert s.size == 10 and np.all(np.diff(s.astype(np.int64)) > 0) and s.max() < 100
    assert np.array_equal(orc.sample_indices(100, 10, None), orc.sample_indices(100, 10, None))
    assert list(orc.sample_indices(4, 1000, 1)) == [0, 1, 2, 3]
    assert orc.sample_indices(0, 5, 1).size == 0
    assert orc.default_seed() == 0x5EED5EED5EED5EED


def test_shuffle_is_fisher_yates_over_the_stream():
    # Fisher-Yates from i=n-1 down with j = next % (i+1), stream seeded seed^fold
    # (rng.rs:64-73), replayed here over the published-vector-checked stream.
    n, seed = 23, 4
    stream = orc.splitmix_stream(seed ^ 0xD1B54A32D192ED03, n).tolist()
    perm = list(range(n))
    for step, i in enumerate(range(n - 1, 0, -1)):
        j = stream[step] % (i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    assert orc.shuffled_order(n, seed).tolist() == perm


def test_sample_sources_count():  # betweenness.rs:349-362
    assert orc.sample_sources(10, None).tolist() == list(range(10))
    assert orc.sample_sources(10, 0.25, 3).size == 3  # ceil(2.5)
    assert orc.sample_sources(10, 0.0, 3).size == 1
    assert orc.sample_sources(10, 5.0, 3).size == 10
    s = orc.sample_sources(1000, 0.256, 4)
    assert np.array_equal(s, orc.shuffled_order(1000, 4)[:256])


def test_intern_goldens():  # id_map.rs:459-478
    n, sd, dd, users = orc.intern_i64([10, 10, 20, 30], [20, 30, 30, 10])
    assert n == 3 and list(sd) == [0, 0, 1, 2] and list(dd) == [1, 2, 2, 0]
    n, sd, dd, users = orc.intern_i64([100, 100], [7, 42])
    assert list(users) == [100, 7, 42]


def test_intern_batches_goldens():  # id_map.rs:529-550, 569-574
    n, sd, dd, users = orc.intern_i64_batches([([10, 10], [20, 30]), ([20, 30], [30, 10])])
    assert n == 3 and list(sd) == [0, 0, 1, 2] and list(dd) == [1, 2, 2, 0] and list(users) == [10, 20, 30]
    n, sd, dd, users = orc.intern_i64_batches([])
    assert n == 0 and sd.size == 0 and dd.size == 0


# ---------------------------------------------------------------- generators
def test_rmat_generator_contract():
    s, d = orc.gen_rmat(10, 16, seed=1)
    assert s.size > 0.6 * (16 << 10)
    assert np.all(s != d) and s.max() < 1024 and d.max() < 1024
    keys = s * 1024 + d
    assert np.unique(keys).size == keys.size
    s2, d2 = orc.gen_rmat(10, 16, seed=1)
    assert np.array_equal(s, s2) and np.array_equal(d, d2)
    # first edge = replay of the stream by hand
    st = orc.splitmix_stream(1, 10).tolist()
    u = v = 0
    for x in st:
        r = (x >> 11) * 2.0 ** -53
        bu, bv = (0, 0) if r < 0.57 else (0, 1) if r < 0.76 else (1, 0) if r < 0.95 else (1, 1)
        u, v = 2 * u + bu, 2 * v + bv
    if u != v:
        assert (s[0], d[0]) == (u, v)
    # quadrant skew: top bit of src is 1 with probability c + d = 0.24
    frac = np.mean(s >= 512)
    assert 0.18 < frac < 0.30
    ss, sd_ = orc.gen_rmat(10, 16, seed=1, symmetrize=True)
    fwd = set(zip(ss.tolist(), sd_.tolist()))
    assert all((b, a) in fwd for a, b in fwd)
    assert len(fwd) == ss.size


def test_er_generator_and_weights():
    s, d = orc.gen_er(500, 8, 1000)
    assert np.all(s != d) and s.max() < 500
    w = orc.gen_weights(1000, 3)
    assert np.all(w > 0) and np.all(w <= 1.0)
    assert np.array_equal(w.astype(np.float32).astype(np.float64), w)


# ---------------------------------------------------------------- personalized PageRank
@pytest.mark.parametrize("name", ["sparse400", "gnp100"])
@pytest.mark.parametrize("qi", [0, 1, 2])
def test_ppr_vs_networkx_personalization(name, qi):
    t, users = topo_from_user(FX[f"{name}_src"], FX[f"{name}_dst"])
    dense = {int(u): i for i, u in enumerate(users)}
    seeds = [dense[int(x)] for x in FX[f"{name}_ppr{qi}_seeds"]]
    pr, _ = orc.ppr(t, seeds, 0.85, 10000, 1e-15)
    got = by_user(users, pr)
    for k, v in zip(FX[f"{name}_ppr{qi}_ids"].tolist(), FX[f"{name}_ppr{qi}"].tolist()):
        assert abs(got[k] - v) < 1e-11


def test_ppr_with_all_nodes_as_seeds_is_pagerank():
    # p uniform over every node: teleport and dangling terms match ursa's
    # global PageRank algebraically (within rounding).
    t, _ = topo_from_user(FX["sparse400_src"], FX["sparse400_dst"])
    pr, _, _ = orc.pagerank(t, 0.85, 10000, 1e-15)
    pp, _ = orc.ppr(t, list(range(t.n)), 0.85, 10000, 1e-15)
    assert np.allclose(pr, pp, rtol=1e-12, atol=0)


# ---------------------------------------------------------------- stats.rs (bfs_source_scan & co)
def test_stats_goldens():  # ursa-plan/src/stats.rs:223-259
    tri = orc.Topology(3, [0, 1, 2], [1, 2, 0])
    assert abs(orc.avg_path_length(tri) - 1.5) < 1e-12
    assert orc.diameter(tri, approximate=False) == 2
    assert orc.diameter(tri, approximate=True) <= 2
    line = orc.Topology(4, [0, 1, 2], [1, 2, 3])
    assert orc.diameter(line, approximate=False) == 3
    two = orc.Topology(4, [0, 2], [1, 3])
    assert abs(orc.avg_path_length(two) - 1.0) < 1e-12
    assert orc.diameter(two, approximate=False) == 1
    assert orc.avg_path_length(orc.Topology(2, [], [])) == 0.0


def test_stats_sources_stride():  # stats.rs:80-98
    assert orc.stats_sources(5).tolist() == [0, 1, 2, 3, 4]
    assert orc.stats_sources(10, 0.3).tolist() == [0, 3, 6]
    assert orc.stats_sources(10, 0.0).tolist() == [0]
    assert orc.stats_sources(7, 5.0).tolist() == list(range(7))
    assert orc.stats_sources(1000, 128 / 1000).shape[0] == 128


def test_bfs_source_scan_vs_scipy_fixture():
    # The committed scipy BFS levels of gnp100 from four sources fix the fold.
    t, users = topo_from_user(FX["gnp100_src"], FX["gnp100_dst"])
    dense = {int(u): i for i, u in enumerate(users)}
    srcs = [dense[int(x)] for x in FX["gnp100_bfs_sources"]]
    lv = FX["gnp100_bfs_out"]
    want_sum = float(lv[lv >= 1].sum())
    want_cnt = int((lv >= 1).sum())
    s, c, m = orc.bfs_source_scan(t, srcs + [10**6])
    assert (s, c, m) == (want_sum, want_cnt, int(lv.max()))


# ---------------------------------------------------------------- closeness / clustering
def test_closeness_goldens():  # closeness.rs:233-275
    c = orc.closeness(orc.Topology(3, [0, 1, 2], [1, 2, 0]))
    assert np.allclose(c, 2.0 / 3.0, rtol=0, atol=1e-12)
    c = orc.closeness(orc.Topology(4, [0, 1, 2], [1, 2, 3]))
    assert abs(c[0] - 0.5) < 1e-12 and abs(c[1] - 2 / 3) < 1e-12 and abs(c[2] - 1.0) < 1e-12 and c[3] == 0.0
    assert orc.closeness(orc.Topology(0, [], [])).size == 0
    n = 100
    c = orc.closeness(orc.Topology(n, np.arange(n), (np.arange(n) + 1) % n))
    assert np.all(np.abs(c - c[0]) < 1e-12) and c[0] > 0


def test_closeness_weighted_goldens():  # closeness.rs:277-309
    c = orc.closeness_weighted(orc.Topology(3, [0, 1, 0], [1, 2, 2]), [1.0, 1.0, 5.0])
    assert abs(c[0] - 2.0 / 3.0) < 1e-12 and abs(c[1] - 1.0) < 1e-12
    c = orc.closeness_weighted(orc.Topology(3, [0, 1], [1, 2]), [0.0, 1.0])  # zero-cost node counted
    assert abs(c[0] - 2.0) < 1e-12 and abs(c[1] - 1.0) < 1e-12
    assert orc.closeness_weighted(orc.Topology(3, [0, 0], [1, 2]), [0.0, 0.0])[0] == 0.0
    # unit weights reduce to the hop-count closeness exactly (integer distances)
    n = 300
    rng = np.random.default_rng(4)
    s, d = rng.integers(0, n, 1500), rng.integers(0, n, 1500)
    t = orc.Topology(n, s, d)
    assert np.array_equal(orc.closeness_weighted(t, np.ones(1500)), orc.closeness(t))
    # scaling every weight by 2 halves every value exactly
    w = rng.random(1500)
    assert np.array_equal(orc.closeness_weighted(t, 2.0 * w) * 2.0, orc.closeness_weighted(t, w))


@pytest.mark.parametrize("name", ["gnp100", "sparse400"])
def test_closeness_weighted_vs_networkx(name):
    t, users = topo_from_user(FX[f"{name}_src"], FX[f"{name}_dst"])
    got = by_user(users, orc.closeness_weighted(t, FX[f"{name}_wclose_w"]))
    for k, v in zip(FX[f"{name}_wclose_ids"].tolist(), FX[f"{name}_wclose"].tolist()):
        assert abs(got[k] - v) <= 1e-14 * max(abs(v), 1e-300)


def test_clustering_goldens():  # clustering.rs:46-78
    assert orc.clustering(orc.Topology(3, [0, 0, 1, 2], [1, 2, 2, 0])).tolist() == [1.0, 1.0, 1.0]
    assert orc.clustering(orc.Topology(3, [0, 1], [1, 2])).tolist() == [0.0, 0.0, 0.0]
    cc = orc.clustering(orc.Topology(4, [0, 0, 0], [1, 2, 3]))
    assert cc[0] == 0.0 and cc[1] == 0.0
    cc = orc.clustering(orc.Topology(4, [0, 0, 0, 1], [1, 2, 3, 2]))
    assert abs(cc[0] - 1 / 3) < 1e-12


@pytest.mark.parametrize("name", ["gnp100", "ba500"])
def test_clustering_vs_networkx(name):
    t, users = topo_from_user(FX[f"{name}_src"], FX[f"{name}_dst"])
    got = by_user(users, orc.clustering(t))
    for k, v in zip(FX[f"{name}_clust_ids"].tolist(), FX[f"{name}_clust"].tolist()):
        assert abs(got[k] - v) < 1e-12


@pytest.mark.parametrize("name", ["gnp100", "sparse400"])
def test_closeness_vs_networkx(name):
    t, users = topo_from_user(FX[f"{name}_src"], FX[f"{name}_dst"])
    got = by_user(users, orc.closeness(t))
    for k, v in zip(FX[f"{name}_close_ids"].tolist(), FX[f"{name}_close"].tolist()):
        assert abs(got[k] - v) < 1e-12
