"""GPU parity for the traversal and intersection kernels, through the C ABI,
against the CPU oracle on the same inputs.

Bars: BFS levels, k_hop pairs (order included), shortest paths, triangle counts
and canonical component labels bit-exact; betweenness within 1e-12 relative per
node.
"""
import contextlib

import numpy as np
import pytest

import oracle as orc
from graphs import cases, rmat_dense

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-12
GRAPHS = cases()
IDS = [g[0] for g in GRAPHS]
DIRS = ("out", "in", "both")


@pytest.fixture(scope="module")
def ua():
    import ursa_amd

    ursa_amd.init(0)
    return ursa_amd


@pytest.fixture(scope="module")
def rmat14(ua):
    n, s, d = rmat_dense(14, 5)
    return n, s, d, ua.Topology.build(n, s, d), orc.Topology(n, s, d)


@pytest.fixture(scope="module")
def rmat12sym(ua):
    n, s, d = rmat_dense(12, 4, symmetrize=True)
    return n, s, d, ua.Topology.build(n, s, d), orc.Topology(n, s, d)


def sources_for(n):
    return sorted({0, n // 2, max(n - 1, 0), n // 7}) + [n + 5]


# ---------------------------------------------------------------- BFS
@pytest.mark.parametrize("mode", [1, 2], ids=["per_source", "multi_source"])
@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_bfs_levels_match(ua, case, mode):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    src = sources_for(n)
    for direction in DIRS:
        got = ua.bfs_multi(t, src, direction, mode=mode)
        for i, x in enumerate(src):
            assert np.array_equal(got[i], orc.bfs(o, x, direction)), (name, direction, x)


@pytest.mark.parametrize("mode", [1, 2], ids=["per_source", "multi_source"])
def test_bfs_rmat_direction_optimizing(ua, rmat14, mode):
    n, _, _, t, o = rmat14
    rng = np.random.default_rng(1)
    src = rng.integers(0, n, 12)
    for direction in DIRS:
        got = ua.bfs_multi(t, src, direction, mode=mode)
        for i, x in enumerate(src):
            assert np.array_equal(got[i], orc.bfs(o, int(x), direction)), (direction, int(x))


def test_msbfs_directions_any_order(ua):
    # The pull lists are built once per direction on first use; the order in
    # which directions are first asked for must not change any level.
    n, s, d = rmat_dense(14, 5)
    o = orc.Topology(n, s, d)
    src = np.random.default_rng(3).integers(0, n, 40)
    want = {x: [orc.bfs(o, int(v), x) for v in src] for x in DIRS}
    t = ua.Topology.build(n, s, d)
    for direction in ("both", "in", "out", "out", "in", "both"):
        got = ua.bfs_multi(t, src, direction, mode=2)
        for i in range(len(src)):
            assert np.array_equal(got[i], want[direction][i]), (direction, i)


def test_msbfs_two_batches_with_duplicates(ua, rmat14):
    n, _, _, t, o = rmat14
    rng = np.random.default_rng(2)
    src = np.concatenate([rng.integers(0, n, 97), [5, 5, n + 3]])
    got = ua.bfs_multi(t, src, "out", mode=2)
    for i, x in enumerate(src.tolist()):
        assert np.array_equal(got[i], orc.bfs(o, int(x), "out")), (i, int(x))


def test_bfs_hub_frontier(ua):
    # source -> hub with 20000 out-edges: the top-down level runs the grid-wide path
    hub_deg = 20000
    s = np.concatenate([[0], np.ones(hub_deg, np.int64), np.arange(2, 2 + hub_deg)])
    d = np.concatenate([[1], np.arange(2, 2 + hub_deg), np.arange(3, 3 + hub_deg) % (2 + hub_deg)])
    n = 2 + hub_deg
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    for mode in (1, 2):
        assert np.array_equal(ua.bfs_multi(t, [0], "out", mode=mode)[0], orc.bfs(o, 0, "out"))


def test_bfs_goldens(ua):  # bfs.rs:266-298
    t = ua.Topology.build(5, [0, 1, 2, 1], [1, 2, 3, 4])
    assert list(ua.bfs_distances(t, 0)) == [0, 1, 2, 3, 2]
    d = ua.bfs_distances(t, 3, "both")
    assert d[0] == 3 and d[4] == 3
    assert list(ua.bfs_distances(t, 99)) == [-1] * 5


# ---------------------------------------------------------------- k_hop / path
@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_k_hop_matches_in_order(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    seeds = [0, 1, 1, n // 3, n + 7] if n > 1 else [0, n + 7]
    for direction in DIRS:
        for k in (0, 1, 2, 3):
            ga, gb = ua.k_hop(t, seeds, k, direction)
            ra, rb = orc.k_hop(o, seeds, k, direction)
            assert np.array_equal(ga, ra) and np.array_equal(gb, rb), (name, direction, k)


def test_k_hop_rmat(ua, rmat14):
    n, _, _, t, o = rmat14
    seeds = [3, 77, 1000, 5]
    for direction in DIRS:
        ga, gb = ua.k_hop(t, seeds, 3, direction)
        ra, rb = orc.k_hop(o, seeds, 3, direction)
        assert np.array_equal(ga, ra) and np.array_equal(gb, rb), direction


def test_k_hop_goldens(ua):  # hop.rs:148-191
    t = ua.Topology.build(5, [0, 1, 2, 1], [1, 2, 3, 4])
    pairs = lambda r: sorted(zip(r[0].tolist(), r[1].tolist()))
    assert pairs(ua.k_hop(t, [0], 2)) == [(0, 1), (0, 2), (0, 4)]
    assert pairs(ua.k_hop(t, [3], 2, "in")) == [(3, 1), (3, 2)]
    assert ua.k_hop(t, [99], 3)[0].size == 0 and ua.k_hop(t, [0], 0)[0].size == 0


@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_shortest_path_matches(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    rng = np.random.default_rng(3)
    pairs = [(0, n - 1), (n // 2, 0), (1 % max(n, 1), 1 % max(n, 1)), (0, n + 3)]
    pairs += [tuple(int(x) for x in rng.integers(0, n, 2)) for _ in range(6)] if n else []
    for direction in DIRS:
        for a, b in pairs:
            g = ua.shortest_path(t, a, b, direction)
            r = orc.shortest_path(o, a, b, direction)
            if r is None:
                assert g is None, (name, direction, a, b)
            else:
                assert g is not None and np.array_equal(g, r), (name, direction, a, b)


def test_shortest_path_rmat(ua, rmat14):
    n, _, _, t, o = rmat14
    rng = np.random.default_rng(9)
    for _ in range(10):
        a, b = (int(x) for x in rng.integers(0, n, 2))
        for direction in DIRS:
            g, r = ua.shortest_path(t, a, b, direction), orc.shortest_path(o, a, b, direction)
            assert (g is None and r is None) or np.array_equal(g, r), (a, b, direction)


def test_shortest_path_goldens(ua):  # bfs.rs:301-332
    t = ua.Topology.build(5, [0, 1, 2, 1], [1, 2, 3, 4])
    assert list(ua.shortest_path(t, 0, 3)) == [0, 1, 2, 3]
    assert list(ua.shortest_path(t, 2, 2)) == [2]
    assert ua.shortest_path(t, 3, 0) is None and ua.shortest_path(t, 0, 99) is None
    assert list(ua.shortest_path(t, 3, 0, "in")) == [3, 2, 1, 0]


# ---------------------------------------------------------------- components
@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_cc_canonical_labels_match(ua, case):
    name, n, s, d = case
    o = orc.Topology(n, s, d)
    ref = orc.canonicalize(orc.cc_weak(o))
    t = ua.Topology.build(n, s, d)
    assert np.array_equal(ua.connected_components_weak(t), ref), name  # no transpose: full sweep
    t.incoming()
    assert np.array_equal(ua.connected_components_weak(t), ref), name  # giant-component skip path


def test_cc_rmat(ua, rmat14):
    n, _, _, t, o = rmat14
    t.incoming()
    assert np.array_equal(ua.connected_components_weak(t), orc.canonicalize(orc.cc_weak(o)))


def test_cc_many_small_components(ua):
    # 20000 disjoint 3-paths plus isolated pairs: many roots, no giant component
    k = 20000
    base = np.arange(k) * 3
    s = np.concatenate([base, base + 1, np.arange(3 * k, 3 * k + 1000, 2)])
    d = np.concatenate([base + 1, base + 2, np.arange(3 * k + 1, 3 * k + 1001, 2)])
    n = 3 * k + 1000
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    assert np.array_equal(ua.connected_components_weak(t), orc.canonicalize(orc.cc_weak(o)))


def _cc_shapes():
    rng = np.random.default_rng(33)
    out = []
    # A long path over shuffled ids: every hook round links along a chain, and
    # the round-1 root links race along it.
    n = 200000
    perm = rng.permutation(n)
    out.append(("shuffled_path", n, perm[:-1], perm[1:]))
    # Stars whose centres have the largest ids: every leaf's first neighbour is
    # larger than the leaf, so round 0 hooks them all by compare-and-swap.
    k, leaves = 300, 400
    centres = np.arange(k * leaves, k * leaves + k)
    s = np.repeat(centres, leaves)
    d = np.arange(k * leaves)
    out.append(("stars_high_centres", k * leaves + k, s, d))
    # Two dense clusters joined by one edge whose endpoints carry no other
    # link in their first two slots.
    a = rng.integers(0, 5000, 60000)
    b = rng.integers(0, 5000, 60000)
    c = rng.integers(5000, 10000, 60000)
    e = rng.integers(5000, 10000, 60000)
    s = np.concatenate([a, c, [4999]])
    d = np.concatenate([b, e, [5000]])
    out.append(("barbell", 10000, s, d))
    return out


@pytest.mark.parametrize("shape", _cc_shapes(), ids=lambda x: x[0])
def test_cc_store_rounds(ua, shape):
    # The two sampling rounds link roots by plain stores, which may drop a link
    # when they race; the final sweep (all out-slots from 1, and all in-edges
    # when the transpose exists) must make every dropped link exactly.
    name, n, s, d = shape
    o = orc.Topology(n, s, d)
    ref = orc.canonicalize(orc.cc_weak(o))
    t = ua.Topology.build(n, s, d)
    for _ in range(3):
        assert np.array_equal(ua.connected_components_weak(t), ref), name  # no transpose
    t.incoming()
    for _ in range(3):
        assert np.array_equal(ua.connected_components_weak(t), ref), name  # giant-component skip


# ---------------------------------------------------------------- triangles
@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_triangles_match(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    got = ua.triangle_count(t)
    assert np.array_equal(got, orc.triangle_count(o)), name
    assert int(got.sum()) % 3 == 0


@contextlib.contextmanager
def tri_tables(ua, bitmaps):
    """Triangle tables for topologies built inside: bitmaps for eligible owners
    (the default) or hash sets only."""
    lib = ua.load()
    ua.check(lib.ursa_hip_triangle_bitmaps(int(bitmaps)))
    try:
        yield
    finally:
        ua.check(lib.ursa_hip_triangle_bitmaps(1))


def test_triangles_rmat_symmetric(ua, rmat12sym):
    n, s, d, t, o = rmat12sym
    want = orc.triangle_count(o, nthreads=8)
    assert np.array_equal(ua.triangle_count(t), want)
    with tri_tables(ua, False):
        assert np.array_equal(ua.triangle_count(ua.Topology.build(n, s, d)), want)


def test_triangles_dense_rmat_both_tables(ua):
    # R-MAT scale 12, edge factor 400, symmetrised: 506 oriented lists longer
    # than 256 (up to 407), staged as bitmaps or as per-workgroup hash sets.
    n, s, d = rmat_dense(12, 9, symmetrize=True, edge_factor=400)
    want = orc.triangle_count(orc.Topology(n, s, d), nthreads=8)
    for bitmaps in (True, False):
        with tri_tables(ua, bitmaps):
            t = ua.Topology.build(n, s, d)
            for _ in range(2):
                assert np.array_equal(ua.triangle_count(t), want), bitmaps


def test_triangles_clique_with_satellites_closed_form(ua):
    # K_300 plus 3000 satellites, each joined to k in {1, 2, 20, 40} distinct
    # clique vertices. Oriented by (degree, id), the 43 lowest-ranked clique
    # vertices have more than 256 out-entries, so satellite pairs into them
    # are walked from the hub's side (only the satellite's entries before the
    # hub), pairs at position 0 have no triangle, and the rest from the
    # satellite's side. Closed form: a satellite with k neighbours is in
    # C(k, 2) triangles, a clique vertex in C(299, 2) plus k - 1 for every
    # satellite joined to it.
    rng = np.random.default_rng(11)
    c, ns = 300, 3000
    iu, ju = np.triu_indices(c, 1)
    src, dst = [iu], [ju]
    want = np.zeros(c + ns, np.uint64)
    want[:c] = (c - 1) * (c - 2) // 2
    for x in range(ns):
        k = int(rng.choice([1, 2, 20, 40]))
        nb = rng.choice(c, size=k, replace=False)
        src.append(np.full(k, c + x))
        dst.append(nb)
        want[c + x] = k * (k - 1) // 2
        want[nb] += k - 1
    s = np.concatenate(src).astype(np.uint32)
    d = np.concatenate(dst).astype(np.uint32)
    n = c + ns
    assert np.array_equal(orc.triangle_count(orc.Topology(n, s, d), nthreads=8), want)
    for bitmaps in (True, False):
        with tri_tables(ua, bitmaps):
            t = ua.Topology.build(n, s, d)
            assert np.array_equal(ua.triangle_count(t), want), bitmaps


def test_triangles_goldens(ua):  # triangle.rs:70-99
    t = ua.Topology.build(4, [0, 0, 0, 1, 1, 2], [1, 2, 3, 2, 3, 3])
    assert list(ua.triangle_count(t)) == [3, 3, 3, 3]
    t = ua.Topology.build(3, [0, 0, 1, 2, 0, 0], [1, 2, 2, 0, 0, 1])
    assert list(ua.triangle_count(t)) == [1, 1, 1]


@pytest.mark.parametrize("bitmaps", [True, False])
def test_triangles_dense_clique(ua, bitmaps):
    with tri_tables(ua, bitmaps):
        # K_300: every node in C(299, 2) triangles; oriented lists exceed one LDS stage
        k = 300
        a, b = np.triu_indices(k, 1)
        t = ua.Topology.build(k, a, b)
        assert np.all(ua.triangle_count(t) == (k - 1) * (k - 2) // 2)
        # K_1100: oriented lists up to 1099 entries, staged per workgroup
        k = 1100
        a, b = np.triu_indices(k, 1)
        t = ua.Topology.build(k, a, b)
        assert np.all(ua.triangle_count(t) == (k - 1) * (k - 2) // 2)


def test_triangles_disjoint_cliques(ua):
    # 300 disjoint K_300 (90,000 nodes, all of degree 299). New ids run in
    # decreasing id, so a clique's lowest node (|N+| = 299) has new id
    # 89,999 - 300 c: the first 82 cliques' are beyond the bitmap range and use
    # the per-workgroup hash table, the others the bitmap. Closed form: every
    # node in C(299, 2) triangles.
    k, c = 300, 300
    a, b = np.triu_indices(k, 1)
    off = (np.arange(c) * k)[:, None]
    t = ua.Topology.build(k * c, (a[None, :] + off).ravel(), (b[None, :] + off).ravel())
    assert np.all(ua.triangle_count(t) == (k - 1) * (k - 2) // 2)


@pytest.mark.parametrize("side", [1000, 4200])
def test_triangles_bipartite_with_chords(ua, side):
    # K_{side,side} plus chords a0-a1, a1-a2, a5-a6 inside side A. Closed form:
    # a chord's endpoints each gain `side` triangles (one per B vertex), every B
    # vertex gains one per chord. With equal degrees the higher id ranks higher,
    # so A vertices orient all `side` B edges outwards: |N+(a)| = side, the
    # workgroup-staged path at 1000 and the unstaged path at 4200.
    A = np.repeat(np.arange(side), side)
    B = side + np.tile(np.arange(side), side)
    chords = np.array([[0, 1], [1, 2], [5, 6]])
    src = np.concatenate([A, chords[:, 0]])
    dst = np.concatenate([B, chords[:, 1]])
    t = ua.Topology.build(2 * side, src, dst)
    want = np.zeros(2 * side, np.uint64)
    for x, y in chords:
        want[x] += side
        want[y] += side
    want[side:] = len(chords)
    assert np.array_equal(ua.triangle_count(t), want)


def _bucket0_ids(k, nbuckets):
    """k new ids >= 1 whose home bucket in an LDS table of `nbuckets` buckets is
    0 (bits 7.. of id * 0x9E3779B1, the kernel's hash): staged together they
    overflow one bucket. Ids colliding for nbuckets also collide for every
    smaller power of two."""
    ids, x = [], 1
    while len(ids) < k:
        if ((((x * 0x9E3779B1) & 0xFFFFFFFF) >> 7) & (nbuckets - 1)) == 0:
            ids.append(x)
        x += 1
    return np.array(ids, np.int64)


@pytest.mark.parametrize("k,nb,bitmaps", [(14, 4, True), (40, 16, True), (300, 128, False), (300, 128, True)])
def test_triangles_colliding_keys(ua, k, nb, bitmaps):
    # An apex joined to k nodes w that form a clique. The triangle kernel hashes
    # new ids (the rank order: decreasing degree, then decreasing id), so the
    # graph pins them: P = max(C) + 1 nodes rank above the apex, all of degree
    # k, and with equal degrees new id p is caller id P - p. The w take the new
    # ids C (all homed in bucket 0), star centres with k private leaves take
    # the other positions below P, the apex is caller id 0 (new id P) and
    # N+(apex) holds exactly C. The bucket fills and the remaining keys spill:
    # 6 buckets at k = 14, 32 at k = 40 (per-wave tables), 292 at k = 300 in a
    # per-workgroup hash table (hash sets only) or, with bitmaps, the bitmap
    # table. Closed form: the apex closes C(k, 2) triangles, each w
    # C(k-1, 2) + (k-1), centres and leaves none.
    cpos = _bucket0_ids(k, nb)
    P = int(cpos.max()) + 1
    w = P - cpos  # caller ids of the clique
    fill = np.setdiff1d(np.arange(P), cpos)
    centres = P - fill
    leaves = P + 1 + np.arange(fill.size * k)
    a, b = np.triu_indices(k, 1)
    src = np.concatenate([np.zeros(k, np.int64), w[a], np.repeat(centres, k)])
    dst = np.concatenate([w, w[b], leaves])
    n = P + 1 + leaves.size
    with tri_tables(ua, bitmaps):
        t = ua.Topology.build(n, src, dst)
        got = ua.triangle_count(t)
    want = np.zeros(n, np.uint64)
    want[0] = k * (k - 1) // 2
    want[w] = (k - 1) * (k - 2) // 2 + (k - 1)
    assert np.array_equal(got, want)
    if k < 100:
        assert np.array_equal(got, orc.triangle_count(orc.Topology(n, src, dst)))


# ---------------------------------------------------------------- betweenness
def assert_bc_close(got, ref, name):
    err = np.abs(got - ref)
    bad = err > BC_RTOL * np.abs(ref)
    assert not bad.any(), f"{name}: {bad.sum()} nodes off, worst abs {err.max():.3e}"


@pytest.mark.parametrize("case", GRAPHS[:7], ids=IDS[:7])
def test_betweenness_exact_matches(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    assert_bc_close(ua.betweenness(t), orc.betweenness(o, nthreads=8), name)


def test_betweenness_sampled_and_batching(ua, rmat12sym):
    n, _, _, t, o = rmat12sym
    ref = orc.betweenness(o, 0.05, 7, nthreads=8)
    assert_bc_close(ua.betweenness(t, 0.05, 7), ref, "sampled")
    src = orc.shuffled_order(n, 4)[:256]
    ref = orc.betweenness_sources(o, src, scale=n / 256, nthreads=8)
    for batch in (64, 16, 1):
        assert_bc_close(ua.betweenness_sources(t, src, scale=n / 256, batch=batch), ref, f"batch{batch}")


def _lattice(w):
    """w x w grid, both directions: corner-to-corner path counts are C(2w-2, w-1)."""
    idx = np.arange(w * w).reshape(w, w)
    a = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    b = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    return w * w, np.concatenate([a, b]), np.concatenate([b, a])


def _layered(width, layers):
    """Layers of `width` vertices, every vertex joined to every vertex of the next
    layer (one direction): width^(layers-1) shortest paths end to end."""
    s, d = [], []
    for L in range(layers - 1):
        for i in range(width):
            for j in range(width):
                s.append(L * width + i)
                d.append((L + 1) * width + j)
    return width * layers, np.array(s), np.array(d)


@pytest.mark.parametrize("shape", ["lattice64", "layered4x40"])
def test_betweenness_past_2_53_paths(ua, shape):
    """Path counts far above 2^53 (C(126, 63) ~ 6e36 corner to corner; 4^39 end to
    end): sums round, and the ordered recount keeps results bit-identical run to run
    and within the relative bar of the oracle (betweenness.rs:200-215,
    ursa-core/tests/determinism_threads.rs:69-94)."""
    n, s, d = _lattice(64) if shape == "lattice64" else _layered(4, 40)
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    src = np.arange(0, n, 7, dtype=np.uint32)[:256]
    ref = orc.betweenness_sources(o, src, nthreads=8)
    for batch in (64, 1):
        runs = [ua.betweenness_sources(t, src, batch=batch) for _ in range(3)]
        assert_bc_close(runs[0], ref, f"{shape} batch{batch}")
        for r in runs[1:]:
            assert np.array_equal(r.view(np.uint64), runs[0].view(np.uint64)), (shape, batch)


def test_betweenness_goldens(ua):  # betweenness.rs:369-439
    t = ua.Topology.build(5, [0, 1, 2, 3], [1, 2, 3, 4])
    assert list(ua.betweenness(t)) == [0.0, 3.0, 4.0, 3.0, 0.0]
    n = 80
    t = ua.Topology.build(n, list(range(n)), [(i + 1) % n for i in range(n)])
    bc = ua.betweenness(t)
    assert np.all(np.abs(bc - bc[0]) < 1e-9) and abs(bc[0] - sum(L - 1 for L in range(2, n))) < 1e-9
    t = ua.Topology.build(4, [0, 0, 0, 1, 2], [1, 1, 2, 3, 3])  # parallel edge counts twice
    bc = ua.betweenness(t)
    assert abs(bc[1] - 2.0 / 3.0) < 1e-15 and abs(bc[2] - 1.0 / 3.0) < 1e-15


# ---------------------------------------------------------------- autotuner
def test_autotune_changes_schedule_not_results(ua):
    n, s, d = rmat_dense(13, 9)
    t = ua.Topology.build(n, s, d)
    srcs = np.arange(0, n, max(1, n // 70))[:70]
    o = orc.Topology(n, s, d)
    for direction in ("out", "both"):
        before1 = ua.bfs_multi(t, srcs[:3], direction, mode=1)
        before2 = ua.bfs_multi(t, srcs, direction, mode=2)
        p = ua.autotune(t, direction, force=True)
        assert p["source"] == "measured"
        assert p["alpha"] in (4.0, 15.0, 60.0) and p["beta"] in (6.0, 18.0, 54.0)
        assert p["push_div"] in (5, 20, 80)
        assert np.array_equal(ua.bfs_multi(t, srcs[:3], direction, mode=1), before1)
        assert np.array_equal(ua.bfs_multi(t, srcs, direction, mode=2), before2)
        assert np.array_equal(before1[0], orc.bfs(o, int(srcs[0]), direction))
        # A second graph with the same degree distribution reuses the choice.
        t2 = ua.Topology.build(n, s, d)
        q = ua.autotune(t2, direction)
        assert q["source"] == "cached" and (q["alpha"], q["beta"], q["push_div"]) == (
            p["alpha"], p["beta"], p["push_div"])
        assert np.array_equal(ua.bfs_multi(t2, srcs, direction, mode=2), before2)


# ---------------------------------------------------------------- stats.rs on the device
@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_bfs_source_scan_matches(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    srcs = list(range(0, n, max(1, n // 150)))[:150] + [n + 2]
    for direction in DIRS:
        assert ua.bfs_source_scan(t, srcs, direction) == orc.bfs_source_scan(o, srcs, direction), (name, direction)
    if n <= 3000:
        assert ua.avg_path_length(t) == orc.avg_path_length(o)
        assert ua.diameter(t, approximate=False) == orc.diameter(o, approximate=False)
    assert ua.avg_path_length(t, 0.05) == orc.avg_path_length(o, 0.05)
    assert ua.diameter(t) == orc.diameter(o)


def test_stats_goldens_device(ua):  # ursa-plan/src/stats.rs:223-259
    tri = ua.Topology.build(3, [0, 1, 2], [1, 2, 0])
    assert ua.avg_path_length(tri) == 1.5 and ua.diameter(tri, approximate=False) == 2
    assert ua.diameter(ua.Topology.build(4, [0, 1, 2], [1, 2, 3]), approximate=False) == 3
    two = ua.Topology.build(4, [0, 2], [1, 3])
    assert ua.avg_path_length(two) == 1.0 and ua.diameter(two, approximate=False) == 1


# ---------------------------------------------------------------- closeness / clustering on the device
@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_closeness_and_clustering_match(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    srcs = np.array(list(range(0, n, max(1, n // 100)))[:100] + [n + 4])
    assert np.array_equal(ua.closeness(t, srcs), orc.closeness(o, srcs)), name
    if n <= 3000:
        assert np.array_equal(ua.closeness(t), orc.closeness(o)), name
    assert np.array_equal(ua.clustering_coefficient(t), orc.clustering(o)), name


def test_closeness_clustering_goldens(ua):  # closeness.rs:233-275, clustering.rs:46-78
    c = ua.closeness(ua.Topology.build(4, [0, 1, 2], [1, 2, 3]))
    assert c.tolist() == [0.5, 2.0 / 3.0, 1.0, 0.0]
    cc = ua.clustering_coefficient(ua.Topology.build(4, [0, 0, 0, 1], [1, 2, 3, 2]))
    assert abs(cc[0] - 1 / 3) < 1e-15 and cc[3] == 0.0


# ---------------------------------------------------------------- weighted closeness (closeness.rs:143-222)
@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_closeness_weighted_matches_oracle(ua, case):
    """Dijkstra costs (min over paths of the left-to-right rounded cost) and the
    node-order sum are reproduced exactly, so the values are bit-identical."""
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    w = np.random.default_rng(21).random(len(s)) + 0.01
    w[::9] = 0.0  # zero-cost edges: reachable nodes at cost 0 still count
    srcs = np.array(list(range(0, n, max(1, n // 70)))[:70] + [n + 4])
    assert np.array_equal(ua.closeness_weighted(t, w, srcs), orc.closeness_weighted(o, w, srcs)), name
    if n <= 600:
        assert np.array_equal(ua.closeness_weighted(t, w), orc.closeness_weighted(o, w)), name


def test_closeness_weighted_rmat_and_edge_cases(ua):
    from ursa_amd import gen

    e = gen.rmat(12, 8, seed=6)
    _, _, sd, dd, _ = e.export()
    t, o = e.topology(), orc.Topology(e.n, sd, dd)
    w = np.random.default_rng(2).random(e.m)
    srcs = np.arange(0, e.n, 37)[:130]
    assert np.array_equal(ua.closeness_weighted(t, w, srcs), orc.closeness_weighted(o, w, srcs))
    # unit weights give the hop-count closeness exactly
    assert np.array_equal(ua.closeness_weighted(t, np.ones(e.m), srcs), ua.closeness(t, srcs))
    # reference goldens (closeness.rs:277-309)
    c = ua.closeness_weighted(ua.Topology.build(3, [0, 1, 0], [1, 2, 2]), [1.0, 1.0, 5.0])
    assert abs(c[0] - 2.0 / 3.0) < 1e-15 and abs(c[1] - 1.0) < 1e-15
    c = ua.closeness_weighted(ua.Topology.build(3, [0, 1], [1, 2]), [0.0, 1.0])
    assert c[0] == 2.0 and c[1] == 1.0
    assert ua.closeness_weighted(ua.Topology.build(3, [0, 0], [1, 2]), [0.0, 0.0])[0] == 0.0
    assert ua.closeness_weighted(ua.Topology.build(0, [], []), []).size == 0
    with pytest.raises(ua.UrsaHipError):
        ua.closeness_weighted(ua.Topology.build(3, [0, 1], [1, 2]), [1.0, -1.0])
    with pytest.raises(ValueError):
        ua.closeness_weighted(ua.Topology.build(3, [0, 1], [1, 2]), [1.0])
