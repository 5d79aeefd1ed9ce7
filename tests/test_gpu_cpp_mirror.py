"""GPU parity of the C++ host mirror (ursa_amd/host/ursa_core.hpp): every mirror
function, called from C++ through the C ABI on the device, against the oracle on the
same graphs (host/mirror_check.cpp drives them; built as host/libursa_mirror.so).

Bars as for the Python mirror: integer outputs bit-exact, pagerank / pagerank_weighted
(default PageRankParams, pagerank.rs:21-36) and exact betweenness within 1e-12 relative
per node, k_hop pairs and the shortest path identical including order, component
labels equal after canonicalisation.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as orc
from graphs import cases, rmat_dense

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MIRROR = os.path.join(ROOT, "ursa_amd", "host", "libursa_mirror.so")
RTOL = 1e-12


@pytest.fixture(scope="module")
def mirror():
    import ursa_amd

    ursa_amd.init(0)
    if not os.path.exists(MIRROR):
        raise RuntimeError(f"{MIRROR} is not built (python -m ursa_amd.build)")
    lib = C.CDLL(MIRROR)
    lib.ursa_mirror_run.restype = C.c_int
    lib.ursa_mirror_weights_len.restype = C.c_int
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _graphs():
    out = [g for g in cases() if g[0] in ("gnp100", "ba500", "diamond", "loops_dups", "multigraph")]
    n, s, d = rmat_dense(11, 7)
    out.append(("rmat11", n, s, d))
    return out


GRAPHS = _graphs()


def _rel_ok(got, ref):
    return np.all(np.abs(got - ref) <= RTOL * np.abs(ref))


@pytest.mark.parametrize("name,n,src,dst", GRAPHS, ids=[g[0] for g in GRAPHS])
def test_cpp_mirror_matches_oracle(mirror, name, n, src, dst):
    s = np.ascontiguousarray(src, np.uint32)
    d = np.ascontiguousarray(dst, np.uint32)
    m = s.shape[0]
    w = orc.gen_weights(m, 5)
    ot = orc.Topology(n, s, d)
    source = int(np.argmax(orc.degree(ot, "out")))
    lv = orc.bfs(ot, source, "out")
    goal = int(np.argmax(lv)) if lv.max() > 0 else source
    nn = max(n, 1)
    deg = [np.zeros(nn, np.uint32) for _ in range(3)]
    pr, prw, bc = (np.zeros(nn, np.float64) for _ in range(3))
    bfs = np.zeros(nn, np.int32)
    ha, hb = np.zeros(nn, np.uint32), np.zeros(nn, np.uint32)
    path = np.zeros(nn, np.uint32)
    tri = np.zeros(nn, np.uint64)
    cc = np.zeros(nn, np.uint32)
    hop_len, path_len = C.c_uint64(0), C.c_uint64(0)
    err = C.create_string_buffer(512)
    rc = mirror.ursa_mirror_run(
        C.c_uint64(n), C.c_uint64(m), _p(s), _p(d), _p(w), C.c_uint32(source), C.c_uint32(goal),
        _p(deg[0]), _p(deg[1]), _p(deg[2]), _p(pr), _p(prw), _p(bfs), _p(ha), _p(hb), C.c_uint64(nn),
        C.byref(hop_len), _p(path), C.byref(path_len), _p(bc), _p(tri), _p(cc), err, C.c_uint64(512))
    assert rc == 0, err.value.decode()
    for got, dirn in zip(deg, ("out", "in", "both")):
        assert np.array_equal(got[:n], orc.degree(ot, dirn)), dirn
    ref, _, _ = orc.pagerank(ot)
    assert _rel_ok(pr[:n], ref)
    ref, _, _ = orc.pagerank(ot, weights=w)
    assert _rel_ok(prw[:n], ref)
    assert np.array_equal(bfs[:n], lv)
    a, b = orc.k_hop(ot, [source], 2, "out")
    assert hop_len.value == a.shape[0]
    assert np.array_equal(ha[: a.shape[0]], a) and np.array_equal(hb[: b.shape[0]], b)
    want = orc.shortest_path(ot, source, goal, "out")
    assert path_len.value == (0 if want is None else want.shape[0])
    if want is not None:
        assert np.array_equal(path[: want.shape[0]], want)
    assert _rel_ok(bc[:n], orc.betweenness(ot))
    assert np.array_equal(tri[:n], orc.triangle_count(ot))
    assert np.array_equal(orc.canonicalize(cc[:n]), orc.canonicalize(orc.cc_weak(ot)))


def test_cpp_mirror_rejects_weight_length(mirror):
    s = np.array([0, 1], np.uint32)
    d = np.array([1, 2], np.uint32)
    err = C.create_string_buffer(512)
    rc = mirror.ursa_mirror_weights_len(C.c_uint64(3), C.c_uint64(2), _p(s), _p(d), C.c_uint64(3), err,
                                        C.c_uint64(512))
    assert rc != 0 and b"weights length must equal the edge count" in err.value
    assert mirror.ursa_mirror_weights_len(C.c_uint64(3), C.c_uint64(2), _p(s), _p(d), C.c_uint64(2), err,
                                          C.c_uint64(512)) == 0
