"""GPU parity for personalized PageRank and the request server, through the C ABI.

Bars:
  * PPR vs the oracle: L1 <= 1e-10 and max-abs <= 1e-11 at tol 1e-12 (both sides
    stop at delta < tol, so each is within tol*d/(1-d) ~ 6e-12 L1 of the f64 fixed
    point); L1 <= 2e-7 at the serving tolerance 1e-8.
  * PPR lanes: every query of a batch is bit-identical to the same query alone.
  * Server: every query's result bytes (digest) equal the plain serial loop's, the
    same query run alone through the server, and the oracle's (k-hop pairs, paths
    and triangle counts bit-exact; PPR within the bar above).
"""
import numpy as np
import pytest

import oracle as orc
from graphs import cases, rmat_dense

pytestmark = pytest.mark.gpu

GRAPHS = [g for g in cases() if g[1] > 0]
IDS = [g[0] for g in GRAPHS]


@pytest.fixture(scope="module")
def ua():
    import ursa_amd

    ursa_amd.init(0)
    return ursa_amd


@pytest.fixture(scope="module")
def sv(ua):
    from ursa_amd import serve

    return serve


def seed_sets(n):
    return [[0], [n // 2, n // 3, n - 1], list(range(0, n, max(1, n // 16)))[:16], [n - 1, n - 1, 0]]


@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_ppr_matches_oracle(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    for seeds in seed_sets(n):
        got, it = ua.ppr(t, seeds, 0.85, 1000, 1e-12)
        ref, rit = orc.ppr(o, seeds, 0.85, 1000, 1e-12)
        assert np.abs(got - ref).sum() <= 1e-10, (name, seeds)
        assert np.abs(got - ref).max() <= 1e-11, (name, seeds)
        assert it > 0 and rit > 0, name


def test_ppr_serving_tolerance_rmat(ua):
    n, s, d = rmat_dense(13, 6)
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    rng = np.random.default_rng(2)
    for k in (1, 5, 16):
        seeds = rng.integers(0, n, k)
        got, _ = ua.ppr(t, seeds, 0.85, 100, 1e-8)
        ref, _ = orc.ppr(o, seeds, 0.85, 100, 1e-8)
        assert np.abs(got - ref).sum() <= 2e-7
        # Mass: an iterate off by eps from sum 1 moves by (1-d)*eps per sweep,
        # so delta < tol bounds the excess by tol/(1-d) ~ 6.7e-8.
        assert abs(got.sum() - 1.0) < 1e-8 / 0.15


@pytest.mark.parametrize("max_iter", [1, 2, 3, 100])
@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_ppr_f64_lanes_follow_the_oracle(ua, max_iter, tol):
    """Lanes with a coarse tol (>= 2e-6) or max_iter <= 16 run f64 sweeps from the
    uniform start: same iteration count as the oracle, values within 1e-11 relative."""
    n, s, d = rmat_dense(12, 3)
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    for seeds in seed_sets(n):
        got, it = ua.ppr(t, seeds, 0.85, max_iter, tol)
        ref, rit = orc.ppr(o, seeds, 0.85, max_iter, tol)
        if tol < 2e-6 and max_iter > 16:  # f32 sweeps, then f64 refinement
            assert np.abs(got - ref).sum() <= 2e-7, seeds
            continue
        assert it == rit, (seeds, it, rit)
        assert np.all(np.abs(got - ref) <= 1e-11 * np.abs(ref)), seeds


def test_ppr_max_iter_inside_the_f32_phase_ends_in_f64(ua):
    """max_iter reached while the f32 phase would still run: the last sweep is an f64
    one, so the result is an f64 iterate; it carries the f32 phase's error (about 1e-7
    of the mass, damped by d per f64 sweep), not the f64 solve's."""
    n, s, d = rmat_dense(12, 3)
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    for mi in (17, 18, 20):
        got, it = ua.ppr(t, [1, 5, 9], 0.85, mi, 1e-15)
        ref, rit = orc.ppr(o, [1, 5, 9], 0.85, mi, 1e-15)
        assert it == rit == mi
        assert np.abs(got - ref).sum() <= 1e-6 and abs(got.sum() - ref.sum()) <= 1e-7


def test_ppr_networkx_fixtures(ua):
    from graphs import FX

    for name in ("sparse400", "gnp100"):
        n, s, d, users = orc.intern_i64(FX[f"{name}_src"], FX[f"{name}_dst"])
        dense = {int(u): i for i, u in enumerate(users)}
        t = ua.Topology.build(n, s, d)
        for qi in range(3):
            seeds = [dense[int(x)] for x in FX[f"{name}_ppr{qi}_seeds"]]
            got, _ = ua.ppr(t, seeds, 0.85, 10000, 1e-15)
            want = dict(zip(FX[f"{name}_ppr{qi}_ids"].tolist(), FX[f"{name}_ppr{qi}"].tolist()))
            for u, i in dense.items():
                assert abs(got[i] - want[u]) < 1e-11, (name, qi, u)


def test_ppr_batch_lanes_are_bit_identical_to_alone(ua):
    n, s, d = rmat_dense(12, 3)
    t = ua.Topology.build(n, s, d)
    rng = np.random.default_rng(5)
    lists = [rng.integers(0, n, int(rng.integers(1, 17))).tolist() for _ in range(21)]
    lists[4] = [n + 3]  # no seed in range
    lists[7] = lists[2]  # duplicate query
    tols = 1e-8
    alone = [ua.ppr(t, sl, 0.85, 100, tols) for sl in lists]
    for group in (lists, lists[:3], lists[5:13][::-1]):
        out, iters = ua.ppr_batch(t, group, 0.85, 100, tols)
        for q, sl in enumerate(group):
            ref, rit = alone[lists.index(sl)]
            assert np.array_equal(out[q], ref) and iters[q] == rit, q
    x, it = alone[4]
    assert it == 0 and np.all(x == 1.0 / n)


def multi_graphs(ua):
    """Graphs for the cross-graph passes: every shared case (hub rows cut into
    segments, isolated nodes, multigraph), R-MAT graphs of several sizes and a
    graph used twice."""
    out = [ua.Topology.build(n, s, d) for (_, n, s, d) in GRAPHS]
    for scale, seed, ef in [(9, 1, 4), (12, 3, 16), (14, 7, 8), (11, 2, 24)]:
        n, s, d = rmat_dense(scale, seed, edge_factor=ef)
        out.append(ua.Topology.build(n, s, d))
    out.append(out[-3])
    return out


@pytest.mark.parametrize("max_iter,tol", [(100, 1e-8), (1, 1e-8), (2, 1e-8), (3, 1e-8), (17, 1e-8),
                                          (60, 1e-4), (1000, 1e-12)])
def test_ppr_multi_is_bit_identical_to_alone(ua, max_iter, tol):
    """One query per graph swept in cross-graph passes gives, for every query,
    the bits and iteration count of the same query alone (f32 and f64 phases,
    stops inside the f32 phase, coarse tolerances, no seed in range)."""
    topos = multi_graphs(ua)
    rng = np.random.default_rng(max_iter)
    lists = []
    for i, t in enumerate(topos):
        n = t.n_nodes()
        lists.append(rng.integers(0, n, int(rng.integers(1, 17))).tolist())
    lists[1] = [topos[1].n_nodes() + 5]  # no seed in range
    outs, iters = ua.ppr_multi(topos, lists, 0.85, max_iter, tol)
    for i, (t, sl) in enumerate(zip(topos, lists)):
        ref, rit = ua.ppr(t, sl, 0.85, max_iter, tol)
        assert np.array_equal(outs[i], ref) and iters[i] == rit, (i, max_iter, tol)
    assert iters[1] == 0 and np.all(outs[1] == 1.0 / topos[1].n_nodes())


def test_ppr_multi_passes_split_and_fall_back(ua):
    """More queries than one pass holds (several passes), device outputs, and
    graphs that sweep binned (forced schedule) solved one at a time: every
    result still equals the query alone."""
    import ctypes as C

    lib = ua.load()
    small = [ua.Topology.build(*rmat_dense(6 + (i % 5), 100 + i, edge_factor=4 + i % 9)) for i in range(40)]
    topos = small * 7  # 280 queries: more than the 256 of one pass
    rng = np.random.default_rng(9)
    lists = [rng.integers(0, t.n_nodes(), int(rng.integers(1, 5))).tolist() for t in topos]
    outs, iters = ua.ppr_multi(topos, lists, 0.85, 100, 1e-8)
    for i in range(0, len(topos), 3):
        ref, rit = ua.ppr(topos[i], lists[i], 0.85, 100, 1e-8)
        assert np.array_equal(outs[i], ref) and iters[i] == rit, i
    # Device outputs.
    k = 6
    dev = []
    for t in topos[:k]:
        p = C.c_void_p()
        ua.check(lib.ursa_hip_device_alloc(8 * t.n_nodes(), C.byref(p)))
        dev.append(p)
    try:
        offs = np.zeros(k + 1, np.uint64)
        for i in range(k):
            offs[i + 1] = offs[i] + len(lists[i])
        flat = np.concatenate([np.asarray(x, np.uint32) for x in lists[:k]])
        handles = (C.c_void_p * k)(*[t._h for t in topos[:k]])
        optrs = (C.c_void_p * k)(*[p.value for p in dev])
        it = np.zeros(k, np.uint32)
        ua.check(lib.ursa_hip_ppr_multi(k, handles, offs.ctypes.data_as(C.c_void_p), flat.ctypes.data_as(C.c_void_p),
                                        0.85, 100, 1e-8, optrs, it.ctypes.data_as(C.c_void_p)))
        for i in range(k):
            got = np.zeros(topos[i].n_nodes(), np.float64)
            ua.check(lib.ursa_hip_memcpy_to_host(got.ctypes.data_as(C.c_void_p), dev[i], got.nbytes))
            assert np.array_equal(got, outs[i]) and it[i] == iters[i], i
    finally:
        for p in dev:
            lib.ursa_hip_device_free(p)
    # Forced binned schedule: no graph qualifies, each is solved alone.
    ua.check(lib.ursa_hip_ppr_schedule(2))
    try:
        outs_b, iters_b = ua.ppr_multi(topos[:8], lists[:8], 0.85, 100, 1e-8)
        for i in range(8):
            ref, rit = ua.ppr(topos[i], lists[i], 0.85, 100, 1e-8)
            assert np.array_equal(outs_b[i], ref) and iters_b[i] == rit, i
    finally:
        ua.check(lib.ursa_hip_ppr_schedule(0))


def test_ppr_edge_cases(ua):
    t = ua.Topology.build(4, [0, 1, 2], [1, 2, 0])  # node 3 isolated and dangling
    o = orc.Topology(4, [0, 1, 2], [1, 2, 0])
    x, it = ua.ppr(t, [3], 0.85, 1000, 1e-13)
    r, _ = orc.ppr(o, [3], 0.85, 1000, 1e-13)
    assert np.abs(x - r).max() < 1e-12 and it > 0
    x, it = ua.ppr(t, [1], 0.85, 0, 1e-8)
    assert it == 0 and np.all(x == 0.25)
    empty = ua.Topology.build(0, [], [])
    x, it = ua.ppr(empty, [0], 0.85, 10, 1e-8)
    assert x.size == 0 and it == 0


# ------------------------------------------------- binned PPR schedule
@pytest.fixture
def ppr_binned(ua):
    """Forces the binned sweep (ursa_hip_ppr_schedule 2) for one test."""
    lib = ua.load()
    ua.check(lib.ursa_hip_ppr_schedule(2))
    yield
    ua.check(lib.ursa_hip_ppr_schedule(0))


@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_ppr_binned_matches_oracle(ua, case, ppr_binned):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    for seeds in seed_sets(n):
        got, it = ua.ppr(t, seeds, 0.85, 1000, 1e-12)
        ref, rit = orc.ppr(o, seeds, 0.85, 1000, 1e-12)
        assert np.abs(got - ref).sum() <= 1e-10, (name, seeds)
        assert np.abs(got - ref).max() <= 1e-11, (name, seeds)
        assert it > 0 and rit > 0, name


@pytest.mark.parametrize("max_iter", [1, 3, 100])
@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_ppr_binned_f64_lanes_follow_the_oracle(ua, max_iter, tol, ppr_binned):
    """As the pull schedule: f64 lanes keep the oracle's iteration count and values
    within 1e-11 relative (their row sums are exact, then rounded once)."""
    n, s, d = rmat_dense(12, 3)
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    for seeds in seed_sets(n):
        got, it = ua.ppr(t, seeds, 0.85, max_iter, tol)
        ref, rit = orc.ppr(o, seeds, 0.85, max_iter, tol)
        if tol < 2e-6 and max_iter > 16:
            assert np.abs(got - ref).sum() <= 2e-7, seeds
            continue
        assert it == rit, (seeds, it, rit)
        assert np.all(np.abs(got - ref) <= 1e-11 * np.abs(ref)), seeds


def test_ppr_binned_serving_tolerance_and_batches(ua, ppr_binned):
    """Serving tolerance against the oracle, and every lane of a batch bit-identical
    to the query alone: each lane runs its own binned passes."""
    n, s, d = rmat_dense(13, 6)
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    rng = np.random.default_rng(2)
    lists = [rng.integers(0, n, int(rng.integers(1, 17))).tolist() for _ in range(18)]
    lists[5] = [n + 3]
    alone = [ua.ppr(t, sl, 0.85, 100, 1e-8) for sl in lists]
    for sl, (got, _) in list(zip(lists, alone))[:4]:
        ref, _ = orc.ppr(o, sl, 0.85, 100, 1e-8)
        assert np.abs(got - ref).sum() <= 2e-7
        assert abs(got.sum() - 1.0) < 1e-8 / 0.15
    for group in (lists[:16], lists[2:5], lists[16:]):
        out, iters = ua.ppr_batch(t, group, 0.85, 100, 1e-8)
        for q, sl in enumerate(group):
            ref, rit = alone[lists.index(sl)]
            assert np.array_equal(out[q], ref) and iters[q] == rit, q
    again = ua.ppr(t, lists[0], 0.85, 100, 1e-8)
    assert np.array_equal(again[0], alone[0][0])


def test_ppr_binned_networkx_fixtures_and_edge_cases(ua, ppr_binned):
    from graphs import FX

    for name in ("sparse400", "gnp100"):
        n, s, d, users = orc.intern_i64(FX[f"{name}_src"], FX[f"{name}_dst"])
        dense = {int(u): i for i, u in enumerate(users)}
        t = ua.Topology.build(n, s, d)
        for qi in range(3):
            seeds = [dense[int(x)] for x in FX[f"{name}_ppr{qi}_seeds"]]
            got, _ = ua.ppr(t, seeds, 0.85, 10000, 1e-15)
            want = dict(zip(FX[f"{name}_ppr{qi}_ids"].tolist(), FX[f"{name}_ppr{qi}"].tolist()))
            for u, i in dense.items():
                assert abs(got[i] - want[u]) < 1e-11, (name, qi, u)
    t = ua.Topology.build(4, [0, 1, 2], [1, 2, 0])
    o = orc.Topology(4, [0, 1, 2], [1, 2, 0])
    x, it = ua.ppr(t, [3], 0.85, 1000, 1e-13)
    r, _ = orc.ppr(o, [3], 0.85, 1000, 1e-13)
    assert np.abs(x - r).max() < 1e-12 and it > 0
    x, it = ua.ppr(ua.Topology.build(0, [], []), [0], 0.85, 10, 1e-8)
    assert x.size == 0 and it == 0


def test_ppr_automatic_schedule_bins_large_sparse_graphs(ua):
    """An Erdos-Renyi graph of 40 blocks of 8192 vertices (no row above 256
    in-edges, 12 per vertex on average) takes the binned sweep automatically
    (an edge factor below 8 would stay pull): the automatic result equals the
    forced binned one bit for bit, stays within the bar of the oracle, and the
    kernel profile names the binned sweep; an R-MAT graph (hub rows) stays pull."""
    from ursa_amd import gen

    lib = ua.load()
    e = gen.erdos_renyi(40 * 8192, 12, 77)
    t = e.topology()
    _, _, sd, dd, _ = e.export()
    seeds = [3, 5000, 200000]
    ua.profile.enable(True)
    ua.profile.reset()
    try:
        auto, it = ua.ppr(t, seeds, 0.85, 1000, 1e-12)
        names = ua.profile.names()
    finally:
        ua.profile.enable(False)
    assert "ppr_binned" in names and "ppr_rows" not in names
    ua.check(lib.ursa_hip_ppr_schedule(2))
    try:
        forced, fit = ua.ppr(t, seeds, 0.85, 1000, 1e-12)
    finally:
        ua.check(lib.ursa_hip_ppr_schedule(0))
    assert np.array_equal(auto, forced) and it == fit
    o = orc.Topology(e.n, sd, dd)
    ref, _ = orc.ppr(o, seeds, 0.85, 1000, 1e-12)
    assert np.abs(auto - ref).sum() <= 1e-10 and np.abs(auto - ref).max() <= 1e-11
    r = gen.rmat(19, 8, seed=9).topology()
    ua.profile.enable(True)
    ua.profile.reset()
    try:
        ua.ppr(r, [1, 2], 0.85, 30, 1e-8)
        names = ua.profile.names()
    finally:
        ua.profile.enable(False)
    assert "ppr_rows" in names and "ppr_binned" not in names


# ---------------------------------------------------------------- server
@pytest.fixture(scope="module")
def pool(ua, sv):
    specs = []
    for i, (kind, a, ef) in enumerate([("rmat", 11, 8), ("er", 3000, 6), ("rmat", 12, 16), ("er", 1500, 20),
                                        ("rmat", 10, 4), ("rmat", 13, 12)]):
        sp = sv.GraphSpec(i, kind, (1 << a) if kind == "rmat" else a, ef, 1000 + i, a if kind == "rmat" else 0)
        specs.append(sp)
    topos = [sv.build_serving_graph(sp) for sp in specs]
    dense = []
    from ursa_amd import gen

    for sp in specs:
        e = gen.rmat(sp.scale, sp.edge_factor, seed=sp.seed) if sp.kind == "rmat" else \
            gen.erdos_renyi(sp.n_target, sp.edge_factor, sp.seed)
        _, _, sd, dd, ids = e.export()
        dense.append((ids.shape[0], sd, dd))
    return topos, dense


def oracle_result(o, q, sv):
    seeds = list(q.seeds)[: q.nseeds]
    dirn = ("out", "in", "both")[q.direction]
    if q.kind == sv.PPR:
        return orc.ppr(o, seeds, q.damping, q.max_iter, q.tol)[0]
    if q.kind == sv.KHOP:
        return orc.k_hop(o, seeds, q.k, dirn)
    if q.kind == sv.PATH:
        return orc.shortest_path(o, seeds[0], seeds[1], dirn)
    return orc.triangle_count(o)


def same(kind, a, b, sv):
    if kind == sv.PPR:
        return np.abs(a - b).sum() <= 2e-7
    if kind == sv.KHOP:
        return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    if kind == sv.PATH:
        return (a is None and b is None) or (a is not None and b is not None and np.array_equal(a, b))
    return np.array_equal(a, b)


def test_server_mix_equals_serial_alone_and_oracle(ua, sv, pool):
    topos, dense = pool
    ns = [t.n_nodes() for t in topos]
    trace = sv.serving_trace(ns, 240, qps=0, seed=11)
    # Pile several queries onto one graph so batches form.
    trace += [sv.ppr_query(2, [i, i + 1]) for i in range(20)] + [sv.triangles_query(1) for _ in range(3)]
    trace += [sv.khop_query(5, [i * 7]) for i in range(10)] + [sv.path_query(0, i, i * 3 + 1) for i in range(10)]

    with sv.Server(serial=True) as ser:
        for t in topos:
            ser.add_graph(t)
        rs = ser.run(trace)
    with sv.Server(slots=4, max_batch=16, keep_results=True) as srv:
        for t in topos:
            srv.add_graph(t)
        rm = srv.run(trace)
        assert max(r.batch for r in rm) > 1
        orcs = [orc.Topology(n, s, d) for (n, s, d) in dense]
        for i, q in enumerate(trace):
            assert rs[i].status == 0 and rm[i].status == 0, (i, rs[i].status, rm[i].status)
            assert rs[i].digest == rm[i].digest and rs[i].len == rm[i].len, (i, q.kind)
            if q.kind == sv.PPR:
                assert rs[i].iters == rm[i].iters
            got = srv.result(i, q)
            assert sv.digest(np.frombuffer(srv.result_bytes(i), np.uint8)) == rm[i].digest
            assert same(q.kind, got, oracle_result(orcs[q.graph], q, sv), sv), (i, q.kind)
        # Alone: each query as the only one in a run.
        for i in range(0, len(trace), 7):
            (r1,) = srv.run([trace[i]])
            assert r1.digest == rm[i].digest and r1.batch == 1, i


def test_server_cross_graph_traversals_equal_plain_and_oracle(ua, sv, pool):
    """k-hop and shortest-path queries on small graphs waiting together run as
    cross-graph passes (one workgroup per query): every result equals the plain
    serial loop's and the oracle's, over all three directions, depths 0-4,
    several seeds per k-hop query (one unknown), unreachable, equal and
    out-of-range path endpoints."""
    topos, dense = pool
    ns = [t.n_nodes() for t in topos]
    rng = np.random.default_rng(31)
    qs = []
    for i in range(160):
        g = int(rng.integers(0, len(topos)))
        n = ns[g]
        d = int(rng.integers(0, 3))
        if i % 2 == 0:
            seeds = rng.integers(0, n, int(rng.integers(1, 5))).tolist()
            if i % 10 == 0:
                seeds.append(n + 7)  # unknown seed (hop.rs:91-93)
            qs.append(sv.khop_query(g, seeds, k=int(rng.integers(0, 5)), direction=d))
        else:
            a, b = (int(x) for x in rng.integers(0, n, 2))
            if i % 9 == 1:
                b = a
            if i % 13 == 1:
                b = n + 3
            qs.append(sv.path_query(g, a, b, direction=d))
    with sv.Server(serial=True) as ser:
        for t in topos:
            ser.add_graph(t)
        rs = ser.run(qs)
    with sv.Server(slots=4, max_batch=16, keep_results=True) as srv:
        for t in topos:
            srv.add_graph(t)
        rm = srv.run(qs)
        assert max(r.batch for r in rm) > 1
        orcs = [orc.Topology(n, s, d) for (n, s, d) in dense]
        for i, q in enumerate(qs):
            assert rs[i].status == 0 and rm[i].status == 0, (i, rs[i].status, rm[i].status)
            assert rs[i].digest == rm[i].digest and rs[i].len == rm[i].len, (i, q.kind)
            assert same(q.kind, srv.result(i, q), oracle_result(orcs[q.graph], q, sv), sv), (i, q.kind)


def test_server_timed_arrivals_and_latency_fields(ua, sv, pool):
    topos, _ = pool
    ns = [t.n_nodes() for t in topos]
    trace = sv.serving_trace(ns, 60, qps=400, seed=3)
    with sv.Server(slots=4) as srv:
        for t in topos:
            srv.add_graph(t)
        res = srv.run(trace, timed_arrivals=True)
    for q, r in zip(trace, res):
        assert r.status == 0
        assert abs(r.arrival_s - q.arrival_s) < 1e-6
        assert r.arrival_s <= r.start_s + 1e-6 <= r.done_s + 2e-6
    summ = sv.latency_summary(res)
    assert summ["failed"] == 0 and summ["p99_ms"] >= summ["p50_ms"] > 0


def test_server_admission_by_workspace_budget(ua, sv, pool):
    topos, _ = pool
    big = 5  # R-MAT scale 13
    # 4 MiB holds a k-hop/path pass on the scale-10 graph (about 2.1 MiB) but
    # not a 16-lane PPR pass on the scale-13 graph (about 7 MiB).
    with sv.Server(slots=4, workspace_bytes=4 << 20) as srv:
        for t in topos:
            srv.add_graph(t)
        qs = [sv.ppr_query(big, [i]) for i in range(16)] + [sv.path_query(4, 0, 5)]
        res = srv.run(qs)
        assert all(r.status == ua._lib.ENOMEM for r in res[:16])
        assert res[16].status == 0
    # One pass at a time fits: unbatched passes queue for memory and all finish.
    with sv.Server(slots=4, max_batch=1, workspace_bytes=4 << 20) as srv:
        for t in topos:
            srv.add_graph(t)
        qs = [sv.khop_query(4, [i]) for i in range(12)]
        res = srv.run(qs)
        assert all(r.status == 0 and r.batch == 1 for r in res)
        spans = sorted((r.start_s, r.done_s) for r in res)
        assert all(b[0] >= a[1] - 1e-6 for a, b in zip(spans, spans[1:]))


def test_server_large_graphs_fit_their_reservations(ua, sv):
    """Graphs whose n-sized temporaries exceed a megabyte (R-MAT scale 18, about
    180K nodes; Erdos-Renyi with 300K): every kind of pass, alone and batched,
    must fit the reservation the server admits it with, and give the plain
    serial loop's bytes."""
    specs = [sv.GraphSpec(0, "rmat", 1 << 18, 8, 2001, 18), sv.GraphSpec(1, "er", 300000, 4, 2002, 0)]
    topos = [sv.build_serving_graph(sp) for sp in specs]
    ns = [t.n_nodes() for t in topos]
    assert min(ns) > (1 << 17)
    trace = []
    for g in range(2):
        trace += [sv.triangles_query(g), sv.triangles_query(g)]
        trace += [sv.ppr_query(g, [i, ns[g] - 1 - i]) for i in range(3)]
        trace += [sv.khop_query(g, [i * 11]) for i in range(3)] + [sv.path_query(g, i, ns[g] - 1 - i) for i in range(3)]
    with sv.Server(serial=True) as ser:
        for t in topos:
            ser.add_graph(t)
        rs = ser.run(trace)
    with sv.Server(slots=4, max_batch=16) as srv:
        for t in topos:
            srv.add_graph(t)
        rm = srv.run(trace)
        alone = [srv.run([q])[0] for q in trace if q.kind == sv.TRIANGLES]
    for i, q in enumerate(trace):
        assert rs[i].status == 0 and rm[i].status == 0, (i, q.kind, rs[i].status, rm[i].status)
        assert rs[i].digest == rm[i].digest and rs[i].len == rm[i].len, (i, q.kind)
    assert all(r.status == 0 for r in alone)


def test_server_rejects_bad_queries(ua, sv, pool):
    topos, _ = pool
    with sv.Server(slots=2) as srv:
        srv.add_graph(topos[0])
        bad = sv.khop_query(0, [1])
        bad.direction = 7
        res = srv.run([sv.ppr_query(3, [1]), bad, sv.triangles_query(0)])
        assert res[0].status == ua._lib.EINVAL and res[1].status == ua._lib.EINVAL and res[2].status == 0


def test_graph_replay_matches_direct_launches(ua):
    # The 8-sweep batches of PageRank and personalized PageRank are replayed as
    # HIP graphs; each replay must equal the directly launched sequence, with
    # other eager work in between.
    n, s, d = rmat_dense(12, 8)
    t = ua.Topology.build(n, s, d)
    lib = ua.load()
    params = ua.PageRankParams(0.85, 100, 1e-12)
    seeds = [[1, 2, 3], [n - 1], list(range(0, n, n // 9))]
    try:
        lib.ursa_hip_graphs_enable(0)
        pr_ref = ua.pagerank(t, params)
        ppr_ref, it_ref = ua.ppr_batch(t, seeds, 0.85, 100, 1e-10)
        lib.ursa_hip_graphs_enable(1)
        for _ in range(3):
            assert np.array_equal(ua.pagerank(t, params), pr_ref)
            ua.bfs_distances(t, 0)
            out, it = ua.ppr_batch(t, seeds, 0.85, 100, 1e-10)
            assert np.array_equal(out, ppr_ref) and np.array_equal(it, it_ref)
            ua.triangle_count(t)
    finally:
        lib.ursa_hip_graphs_enable(1)
