"""GPU parity: device topology, degree, PageRank and the device graph generator,
each through the C ABI, against the CPU oracle on the same seeded inputs.

Bars: integer outputs bit-exact; f64 PageRank within 1e-12 relative per node with
the same iteration count (the reference stops at the first delta < tol,
pagerank.rs:87-90).
"""
import numpy as np
import pytest

import oracle as orc
from graphs import cases, rmat_dense

pytestmark = pytest.mark.gpu

PR_RTOL = 1e-12


@pytest.fixture(scope="module")
def ua():
    import ursa_amd

    ursa_amd.init(0)
    return ursa_amd


GRAPHS = cases()
IDS = [g[0] for g in GRAPHS]


@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_topology_matches_oracle(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    for which in ("out", "in"):
        got = t.out() if which == "out" else t.incoming()
        ref = o.adjacency(which)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r), (name, which)
    gu, ru = t.undirected(), o.undirected()
    assert np.array_equal(gu[0], ru[0]) and np.array_equal(gu[1], ru[1]), name


@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_degree_matches_oracle(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    for direction in ("out", "in", "both"):
        assert np.array_equal(ua.degree(t, direction), orc.degree(o, direction)), (name, direction)


def assert_pr_close(got, ref, name):
    assert got.shape == ref.shape
    err = np.abs(got - ref)
    bad = err > PR_RTOL * np.abs(ref)
    assert not bad.any(), f"{name}: {bad.sum()} nodes beyond 1e-12 rel, worst {np.max(err / np.abs(ref)):.3e}"


@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
@pytest.mark.parametrize("params", [(0.85, 30, 1e-6), (0.85, 100, 1e-10), (0.5, 7, 0.0)])
def test_pagerank_matches_oracle(ua, case, params):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    got, it = ua._pagerank(t, None, ua.PageRankParams(*params))
    ref, rit, rdelta = orc.pagerank(o, *params)
    assert it == rit, (name, it, rit, rdelta)
    assert_pr_close(got, ref, name)


@pytest.mark.parametrize("case", GRAPHS[:7], ids=IDS[:7])
def test_pagerank_weighted_matches_oracle(ua, case):
    name, n, s, d = case
    m = len(s)
    w = orc.gen_weights(m, 3)  # f32-exact weights: the f32 edge-weight path
    w[::7] = 0.0
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    params = (0.85, 100, 1e-10)
    got, it = ua._pagerank(t, w, ua.PageRankParams(*params))
    ref, rit, _ = orc.pagerank(o, *params, weights=w)
    assert it == rit
    assert_pr_close(got, ref, name)
    # weights not exact in f32 take the f64 edge-weight path
    w64 = w * (1.0 + 1e-9)
    got, it = ua._pagerank(t, w64, ua.PageRankParams(*params))
    ref, rit, _ = orc.pagerank(o, *params, weights=w64)
    assert it == rit
    assert_pr_close(got, ref, name)


def test_pagerank_weighted_rmat16_and_out_hubs(ua):
    """Weighted PageRank (pagerank.rs:135-169) on R-MAT scale 16 and on a graph whose
    hubs have tens of thousands of out-edges (out-strength summed by a whole wave):
    same iteration count as the oracle and 1e-12 relative per node, f32-exact and
    f64 weights."""
    n16, s16, d16 = rmat_dense(16, 3)
    rng = np.random.default_rng(8)
    hubs = np.concatenate([np.zeros(40000, np.int64), np.ones(25000, np.int64)])
    sh = np.concatenate([hubs, rng.integers(2, 20000, 60000)])
    dh = np.concatenate([rng.integers(2, 20000, 65000), rng.integers(0, 20000, 60000)])
    for name, n, s, d in (("rmat16", n16, s16, d16), ("out_hubs", 20000, sh, dh)):
        t = ua.Topology.build(n, s, d)
        o = orc.Topology(n, s, d)
        w = orc.gen_weights(len(s), 3)
        for ww in (w, w * (1.0 + 1e-9)):
            got, it = ua._pagerank(t, ww, ua.PageRankParams(0.85, 100, 1e-10))
            ref, rit, _ = orc.pagerank(o, 0.85, 100, 1e-10, weights=ww)
            assert it == rit, (name, it, rit)
            assert_pr_close(got, ref, name)


def test_pagerank_edge_cases(ua):
    t = ua.Topology.build(0, [], [])
    assert ua.pagerank(t).size == 0
    t = ua.Topology.build(3, [0, 1], [1, 2])
    pr, it = ua._pagerank(t, None, ua.PageRankParams(0.85, 0, 1e-6))
    assert it == 0 and np.all(pr == 1.0 / 3.0)
    with pytest.raises(ValueError):
        ua.pagerank_weighted(t, [1.0], ua.PageRankParams())
    # reference goldens (pagerank.rs:176-219)
    t = ua.Topology.build(4, [0, 1, 3, 2], [2, 2, 2, 0])
    pr = ua.pagerank(t)
    assert abs(pr.sum() - 1.0) < 1e-6 and all(pr[i] < pr[2] for i in (0, 1, 3))


def test_pagerank_is_deterministic(ua):
    rng = np.random.default_rng(2)
    t = ua.Topology.build(20000, rng.integers(0, 20000, 300000), rng.integers(0, 20000, 300000))
    a = ua.pagerank(t, ua.PageRankParams(0.85, 50, 1e-12))
    b = ua.pagerank(t, ua.PageRankParams(0.85, 50, 1e-12))
    assert np.array_equal(a, b)


def test_out_of_range_endpoint_is_einval(ua):
    with pytest.raises(ua.UrsaHipError) as e:
        ua.Topology.build(3, [0, 5], [1, 2])
    assert e.value.code == -22


def test_generator_matches_oracle_generator(ua):
    from ursa_amd import gen

    for kw in (dict(scale=12, seed=1), dict(scale=11, seed=4, symmetrize=True)):
        e = gen.rmat(kw["scale"], 16, seed=kw["seed"], symmetrize=kw.get("symmetrize", False))
        us, ud, sd, dd, ids = e.export()
        rs, rd = orc.gen_rmat(kw["scale"], 16, seed=kw["seed"], symmetrize=kw.get("symmetrize", False))
        assert np.array_equal(us, rs) and np.array_equal(ud, rd), kw
        n, osd, odd, oids = orc.intern_i64(rs, rd)
        assert n == e.n and np.array_equal(osd, sd) and np.array_equal(odd, dd) and np.array_equal(oids, ids)
    e = gen.erdos_renyi(5000, 8, 1007)
    us, ud, _, _, _ = e.export()
    rs, rd = orc.gen_er(5000, 8, 1007)
    assert np.array_equal(us, rs) and np.array_equal(ud, rd)
    assert np.array_equal(gen.weights(1000, 3), orc.gen_weights(1000, 3))


def test_config1_pagerank_rmat_s16(ua):
    """R-MAT scale 16, seed 1, damping 0.85, tol 1e-6, max_iter 100, f64."""
    from ursa_amd import gen

    e = gen.rmat(16, 16, seed=1)
    _, _, sd, dd, _ = e.export()
    t = e.topology()
    o = orc.Topology(e.n, sd, dd)
    got, it = ua._pagerank(t, None, ua.PageRankParams(0.85, 100, 1e-6))
    ref, rit, rdelta = orc.pagerank(o, 0.85, 100, 1e-6)
    assert it == rit, (it, rit, rdelta)
    assert_pr_close(got, ref, "rmat16")


@pytest.mark.parametrize("case", GRAPHS[:8], ids=IDS[:8])
def test_pagerank_mixed_within_1e9_l1(ua, case):
    """f32 compensated sweeps + f64 refinement: L1 distance to the f64 oracle < 1e-9."""
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    params = ua.PageRankParams(0.85, 100, 1e-10)
    got, it, it32 = ua.pagerank_mixed(t, params)
    ref, rit, _ = orc.pagerank(o, 0.85, 100, 1e-10)
    assert np.abs(got - ref).sum() < 1e-9, (name, it, it32, np.abs(got - ref).sum())
    w = orc.gen_weights(len(s), 5)
    got, it, it32 = ua.pagerank_mixed(t, params, weights=w)
    ref, _, _ = orc.pagerank(o, 0.85, 100, 1e-10, weights=w)
    assert np.abs(got - ref).sum() < 1e-9, (name, "weighted", np.abs(got - ref).sum())


def test_pagerank_mixed_rmat(ua):
    from ursa_amd import gen

    e = gen.rmat(16, 16, seed=3)
    _, _, sd, dd, _ = e.export()
    t = e.topology()
    o = orc.Topology(e.n, sd, dd)
    got, it, it32 = ua.pagerank_mixed(t, ua.PageRankParams(0.85, 100, 1e-10))
    ref, rit, _ = orc.pagerank(o, 0.85, 100, 1e-10)
    assert it32 >= 1 and np.abs(got - ref).sum() < 1e-9
    top_g = np.argsort(-got, kind="stable")[:1000]
    top_r = np.argsort(-ref, kind="stable")[:1000]
    assert len(set(top_g) & set(top_r)) >= 995


# ---------------------------------------------------------------- from_csr / caller stream
def test_topology_from_csr_matches_build(ua):
    n, s, d = rmat_dense(11, 12)
    t = ua.Topology.build(n, s, d)
    off, tg, ei = t.out()
    t2 = ua.Topology.from_csr(n, off, tg, ei)
    for a, b in zip(t.out(), t2.out()):
        assert np.array_equal(a, b)
    for a, b in zip(t.incoming(), t2.incoming()):
        assert np.array_equal(a, b)
    for a, b in zip(t.undirected(), t2.undirected()):
        assert np.array_equal(a, b)
    p = ua.PageRankParams(0.85, 100, 1e-10)
    assert np.array_equal(ua.pagerank(t, p), ua.pagerank(t2, p))
    assert np.array_equal(ua.triangle_count(t), ua.triangle_count(t2))


def test_topology_from_csr_rejects_bad_csr(ua):
    off = np.array([0, 2, 1, 3], np.uint64)  # decreasing
    with pytest.raises(ua.UrsaHipError) as e:
        ua.Topology.from_csr(3, off, [1, 2, 0], [0, 1, 2])
    assert e.value.code == ua._lib.EINVAL
    with pytest.raises(ua.UrsaHipError):  # duplicate edge id
        ua.Topology.from_csr(3, [0, 1, 2, 3], [1, 2, 0], [0, 0, 2])
    with pytest.raises(ua.UrsaHipError):  # target out of range
        ua.Topology.from_csr(3, [0, 1, 2, 3], [1, 5, 0], [0, 1, 2])


def test_caller_stream(ua):
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")  # the HIP runtime the library already uses
    n, s, d = rmat_dense(10, 13)
    t = ua.Topology.build(n, s, d)
    ref = ua.ppr(t, [1, 2], 0.85, 100, 1e-10)[0]
    lib = ua.load()
    st = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(st), 1) == 0  # hipStreamNonBlocking
    try:
        ua.check(lib.ursa_hip_set_stream(st))
        got = ua.ppr(t, [1, 2], 0.85, 100, 1e-10)[0]
        lv = ua.bfs_distances(t, 0)
    finally:
        ua.check(lib.ursa_hip_set_stream(None))
        hip.hipStreamDestroy(st)
    assert np.array_equal(got, ref)
    assert np.array_equal(lv, ua.bfs_distances(t, 0))


def test_device_outputs_written_on_return(ua):
    # Calls return only when their results are written, also when the output is
    # device memory and the call's temporaries were kept for reuse instead of
    # freed: a plain hipMemcpy, which does not wait for the library's
    # non-blocking stream, must read the final values right after each call.
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n, s, d = rmat_dense(16, 31)
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    lib = ua.load()
    srcs = np.array([0, 5, 9, 17], np.uint32)
    want_lv = np.stack([orc.bfs(o, int(x)) for x in srcs])
    dev = C.c_void_p()
    ua.check(lib.ursa_hip_device_alloc(8 * n + 4 * len(srcs) * n, C.byref(dev)))
    try:
        dtri = dev.value
        dlv = dev.value + 8 * n
        reached = C.c_uint64(0)
        for _ in range(3):  # later calls reuse the blocks the first one kept
            ua.check(lib.ursa_hip_triangle_count(t._h, C.c_void_p(dtri)))
            tri = np.zeros(n, np.uint64)
            assert hip.hipMemcpy(tri.ctypes.data_as(C.c_void_p), C.c_void_p(dtri), 8 * n, 2) == 0
            assert np.array_equal(tri, ua.triangle_count(t))
            ua.check(lib.ursa_hip_bfs_ex(t._h, srcs.ctypes.data_as(C.c_void_p), len(srcs), 0,
                                         C.c_void_p(dlv), C.byref(reached), 0))
            lv = np.zeros((len(srcs), n), np.int32)
            assert hip.hipMemcpy(lv.ctypes.data_as(C.c_void_p), C.c_void_p(dlv), 4 * len(srcs) * n, 2) == 0
            assert np.array_equal(lv, want_lv)
    finally:
        ua.check(lib.ursa_hip_device_free(dev))


# ---------------------------------------------------------------- device ingest
def test_intern_i64_matches_oracle(ua):
    rng = np.random.default_rng(21)
    inputs = [
        (np.array([10, 20, 30, 10]), np.array([20, 30, 10, 40])),  # id_map.rs:459-478 shape
        (rng.integers(-50, 50, 1000), rng.integers(-50, 50, 1000)),  # negative ids, many repeats
        (rng.integers(0, 2**62, 20000, dtype=np.int64), rng.integers(0, 2**62, 20000, dtype=np.int64)),
        (np.full(7, 5), np.full(7, 5)),  # one id, self-loops
    ]
    us, ud = orc.gen_rmat(12, 16, seed=3)
    inputs.append((us, ud))
    for s, d in inputs:
        n, sd, dd, users = ua.intern_i64(s, d)
        rn, rs, rd, ru = orc.intern_i64(s, d)
        assert n == rn and np.array_equal(sd, rs) and np.array_equal(dd, rd) and np.array_equal(users, ru)
    n, sd, dd, users = ua.intern_i64([], [])
    assert n == 0 and sd.size == 0


def test_intern_i64_table_edge_cases(ua):
    """The device hash table marks empty slots with the id -2^63, which is still a
    valid id (counted beside the table, first seen anywhere in the rows); and more
    distinct ids than the first 2^23-slot table holds (12 M) force its rebuild."""
    lo = np.int64(-2**63)
    for s, d in [(np.array([lo, 5, 7, lo]), np.array([5, lo, lo, 9])),
                 (np.array([3, 4, 5]), np.array([lo, 4, 3])),
                 (np.array([lo]), np.array([lo]))]:
        n, sd, dd, users = ua.intern_i64(s, d)
        rn, rs, rd, ru = orc.intern_i64(s, d)
        assert n == rn and np.array_equal(sd, rs) and np.array_equal(dd, rd) and np.array_equal(users, ru)
    rng = np.random.default_rng(8)
    m = 6_000_000
    s = rng.integers(-2**62, 2**62, m, dtype=np.int64)
    d = rng.integers(-2**62, 2**62, m, dtype=np.int64)
    d[::97] = s[::97]  # some repeats
    n, sd, dd, users = ua.intern_i64(s, d)
    rn, rs, rd, ru = orc.intern_i64(s, d)
    assert n == rn and n > (1 << 23)
    assert np.array_equal(sd, rs) and np.array_equal(dd, rd) and np.array_equal(users, ru)


def test_intern_i64_batches_matches_oracle(ua):
    """Interning across batches (id_map.rs:220-292): the two-chunk golden of
    id_map.rs:529-550, the empty batch list, empty batches between full ones, and an
    R-MAT edge list cut at ragged points, each equal to the oracle's interning of the
    rows laid end to end."""
    n, sd, dd, users = ua.intern_i64_batches([([10, 10], [20, 30]), ([20, 30], [30, 10])])
    assert n == 3 and list(sd) == [0, 0, 1, 2] and list(dd) == [1, 2, 2, 0] and list(users) == [10, 20, 30]
    n, sd, dd, users = ua.intern_i64_batches([])
    assert n == 0 and sd.size == 0 and users.size == 0
    n, sd, dd, users = ua.intern_i64_batches([([], []), ([], [])])
    assert n == 0 and sd.size == 0
    us, ud = orc.gen_rmat(14, 16, seed=5)
    rng = np.random.default_rng(6)
    cuts = np.sort(rng.integers(0, us.size, 9))
    cuts = np.concatenate([[0], cuts, [cuts[-1]], [us.size]])  # one empty batch
    batches = [(us[a:b], ud[a:b]) for a, b in zip(cuts[:-1], cuts[1:])]
    got = ua.intern_i64_batches(batches)
    want = orc.intern_i64(us, ud)
    assert got[0] == want[0]
    for g, w in zip(got[1:], want[1:]):
        assert np.array_equal(g, w)
    with pytest.raises(ValueError):
        ua.intern_i64_batches([([1, 2], [3])])


# ---- binned PageRank sweep (exact fixed-point row sums) ----
import contextlib


@contextlib.contextmanager
def pr_schedule(ua, mode):
    """0 automatic, 1 pull, 2 binned (ursa_hip_pagerank_schedule)."""
    lib = ua.load()
    ua.check(lib.ursa_hip_pagerank_schedule(mode))
    try:
        yield
    finally:
        ua.check(lib.ursa_hip_pagerank_schedule(0))


@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_pagerank_binned_matches_oracle(ua, case):
    """Binned sweep on every fixture graph (one tile up to a few): the oracle's
    iteration count and 1e-12 relative per node, plain and weighted (f32-exact
    and f64 weights)."""
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    with pr_schedule(ua, 2):
        for params in ((0.85, 30, 1e-6), (0.85, 100, 1e-10), (0.5, 7, 0.0)):
            got, it = ua._pagerank(t, None, ua.PageRankParams(*params))
            ref, rit, rdelta = orc.pagerank(o, *params)
            assert it == rit, (name, params, it, rit, rdelta)
            assert_pr_close(got, ref, name)
        if len(s):
            w = orc.gen_weights(len(s), 3)
            w[::7] = 0.0
            for ww in (w, w * (1.0 + 1e-9)):
                got, it = ua._pagerank(t, ww, ua.PageRankParams(0.85, 100, 1e-10))
                ref, rit, _ = orc.pagerank(o, 0.85, 100, 1e-10, weights=ww)
                assert it == rit, (name, "weighted")
                assert_pr_close(got, ref, name)


def test_pagerank_binned_split_bins_rmat(ua):
    """R-MAT scale 18 (22 source blocks, 3.9 M in-slots): the hub rows fill bins
    of fewer than 4096 rows; plain, weighted and mixed, bit-identical across runs
    and within 1e-13 relative of the pull sweep."""
    from ursa_amd import gen

    e = gen.rmat(18, 16, seed=3)
    _, _, sd, dd, _ = e.export()
    t = e.topology()
    o = orc.Topology(e.n, sd, dd)
    params = ua.PageRankParams(0.85, 100, 1e-10)
    ref, rit, _ = orc.pagerank(o, 0.85, 100, 1e-10, nthreads=8)
    with pr_schedule(ua, 1):
        pull, pit = ua._pagerank(t, None, params)
    with pr_schedule(ua, 2):
        a, it = ua._pagerank(t, None, params)
        b, _ = ua._pagerank(t, None, params)
        assert it == rit == pit
        assert np.array_equal(a, b)
        assert_pr_close(a, ref, "rmat18")
        assert np.max(np.abs(a - pull) / np.abs(pull)) <= 1e-13
        w = orc.gen_weights(len(sd), 4)
        gw, wit = ua._pagerank(t, w, params)
        rw, rwit, _ = orc.pagerank(o, 0.85, 100, 1e-10, weights=w, nthreads=8)
        assert wit == rwit
        assert_pr_close(gw, rw, "rmat18 weighted")
        mx, mit, m32 = ua.pagerank_mixed(t, params)
        assert m32 >= 1 and np.abs(mx - ref).sum() < 1e-9
        mx2, _, _ = ua.pagerank_mixed(t, params)
        assert np.array_equal(mx, mx2)


def test_pagerank_automatic_schedule_is_binned_from_32_blocks(ua):
    """The automatic schedule takes the binned sweep from 32 source blocks up:
    its result equals the forced binned result bit for bit."""
    from ursa_amd import gen

    e = gen.rmat(19, 8, seed=9)
    t = e.topology()
    auto = ua.pagerank(t, ua.PageRankParams(0.85, 60, 1e-11))
    with pr_schedule(ua, 2):
        forced = ua.pagerank(t, ua.PageRankParams(0.85, 60, 1e-11))
    assert e.n > 31 * 8192
    assert np.array_equal(auto, forced)


def test_exact_sum_matches_rational_arithmetic(ua):
    """ursa_hip_exact_sum (the binned sweep's accumulator on the device) against
    Fraction sums of the terms truncated to multiples of 2^-100 (sums below 8
    in magnitude, the accumulator's range)."""
    from fractions import Fraction
    import ctypes as C
    import math

    lib = ua.load()
    rng = np.random.default_rng(3)
    cases = [np.zeros(0), np.array([1.0, 2.0 ** -53]), np.array([1.0, 2.0 ** -53, 2.0 ** -100]),
             rng.uniform(-1, 1, 5000) * 10.0 ** rng.uniform(-30, 0, 5000), np.full(100000, 7.5e-5)]
    for x in cases:
        x = np.ascontiguousarray(x, np.float64)
        out = C.c_double(0.0)
        ua.check(lib.ursa_hip_exact_sum(x.ctypes.data_as(C.c_void_p) if x.size else None, x.size, C.byref(out)))
        tot = Fraction(0)
        for v in x:
            f = Fraction(float(v)) * (1 << 100)
            tot += Fraction(math.trunc(f) if f >= 0 else -math.trunc(-f), 1 << 100)
        assert out.value == float(tot)
