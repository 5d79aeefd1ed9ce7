"""GPU parity of the formats either side of the kernels.

* Edge weights in slot order: evaluate_weight yields one f64 per edge row and a
  weighted kernel reads weights[edge_ids[k]] per adjacency slot
  (ursa-plan/src/weight.rs:8-10, 51-106). Bit-exact against the oracle CSR's
  edge_ids; negative weights are rejected as evaluate_weight rejects them.
* Results into pinned host buffers (the Arrow egress of
  ursa-plan/src/result.rs:84-139): identical bytes to a result written into
  ordinary host memory.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as orc
from graphs import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ua():
    import ursa_amd

    ursa_amd.init(0)
    return ursa_amd


GRAPHS = cases()
IDS = [g[0] for g in GRAPHS]


@pytest.mark.parametrize("case", GRAPHS, ids=IDS)
def test_gather_edge_weights_matches_slot_order(ua, case):
    name, n, s, d = case
    t = ua.Topology.build(n, s, d)
    o = orc.Topology(n, s, d)
    w = np.random.default_rng(3).random(len(s))
    for which in ("out", "in"):
        eid = o.adjacency(which)[2]
        got = ua.gather_edge_weights(t, w, which)
        assert np.array_equal(got, w[eid]), (name, which)


def test_gather_edge_weights_rmat(ua):
    from ursa_amd import gen

    e = gen.rmat(18, 16, seed=4)
    t = e.topology()
    w = np.random.default_rng(8).random(e.m) + 0.25
    for which, adj in (("out", t.out), ("in", t.incoming)):
        eid = adj()[2]
        assert np.array_equal(ua.gather_edge_weights(t, w, which), w[eid]), which


def test_gather_edge_weights_rejects_negative_like_evaluate_weight(ua):
    t = ua.Topology.build(4, [0, 1, 2, 3], [1, 2, 3, 0])
    w = np.array([1.0, 0.0, 2.5, 3.0])
    assert np.array_equal(ua.gather_edge_weights(t, w, "in"), w[t.incoming()[2]])
    w[2] = -0.5
    with pytest.raises(ua.UrsaHipError) as e:
        ua.gather_edge_weights(t, w, "out")
    assert e.value.code == -22 and "non-negative" in str(e.value)
    # NaN is not negative: evaluate_weight's `w < 0.0` check lets it through
    w[2] = np.nan
    got = ua.gather_edge_weights(t, w, "out")
    assert np.isnan(got[t.out()[2] == 2]).all()
    with pytest.raises(ValueError):
        ua.gather_edge_weights(t, w[:3], "out")
    bad_which = ua.load().ursa_hip_gather_edge_weights(t._h, 2, w.ctypes.data_as(C.c_void_p),
                                                       np.zeros(4).ctypes.data_as(C.c_void_p))
    assert bad_which == -22
    empty = ua.Topology.build(3, [], [])
    assert ua.gather_edge_weights(empty, np.zeros(0), "in").size == 0


def test_results_into_pinned_buffers(ua):
    from ursa_amd import gen

    e = gen.rmat(16, 16, seed=1)
    t = e.topology()
    n = e.n
    ref = ua.pagerank(t)
    buf = ua.PinnedBuffer(8 * n)
    it = C.c_uint32(0)
    ua.check(ua.load().ursa_hip_pagerank(t._h, None, 0.85, 30, 1e-6, C.c_void_p(buf.ptr), C.byref(it)))
    assert np.array_equal(buf.array(np.float64, n), ref)
    deg = ua.PinnedBuffer(4 * n)
    ua.check(ua.load().ursa_hip_degree(t._h, 0, C.c_void_p(deg.ptr)))
    assert np.array_equal(deg.array(np.uint32, n), ua.degree(t, "out"))
    buf.free()
    deg.free()
    assert ua.PinnedBuffer(0).ptr is None
