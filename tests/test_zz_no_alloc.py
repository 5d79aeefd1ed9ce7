"""The no-allocation components entry point (ursa_hip_cc_weak_noalloc) against
ursa_hip_cc_weak, through the C ABI: bit-identical labels on two calls in a row, the
second on a smaller graph (the shared device scratch keeps the first call's
contents), with and without the transpose, plus its argument checks.
"""
import ctypes as C

import numpy as np
import pytest

from graphs import rmat_dense

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ua():
    import ursa_amd

    ursa_amd.init(0)
    return ursa_amd


def noalloc_labels(ua, t, n):
    from ursa_amd import _lib

    lib = ua.load()
    p = C.c_void_p()
    _lib.check(lib.ursa_hip_device_alloc(max(4 * n, 8), C.byref(p)))
    try:
        _lib.check(lib.ursa_hip_cc_weak_noalloc(t._h, p))
        _lib.check(lib.ursa_hip_synchronize())
        out = np.zeros(n, np.uint32)
        if n:
            _lib.check(lib.ursa_hip_memcpy_to_host(out.ctypes.data_as(C.c_void_p), p, 4 * n))
        return out
    finally:
        lib.ursa_hip_device_free(p)


@pytest.mark.parametrize("transpose", [False, True], ids=["out_only", "with_transpose"])
def test_noalloc_matches_cc_weak_two_calls(ua, transpose):
    graphs = [rmat_dense(16, 5), rmat_dense(12, 6, symmetrize=True)]
    rng = np.random.default_rng(4)
    k = 3000  # many small components, no giant one
    graphs.append((3 * k + 10, np.concatenate([np.arange(k) * 3, np.arange(k) * 3 + 1]),
                   np.concatenate([np.arange(k) * 3 + 1, np.arange(k) * 3 + 2])))
    hub = 70000  # a hub whose out-list goes through the heavy queue
    graphs.append((hub + 1, np.zeros(hub, np.int64), rng.permutation(np.arange(1, hub + 1))))
    sizes = [g[0] for g in graphs]
    assert sizes[1] < sizes[0]  # the second call is on a smaller graph
    for n, s, d in graphs:
        t = ua.Topology.build(n, s, d)
        if transpose:
            t.incoming()
        want = ua.connected_components_weak(t)
        for _ in range(2):
            assert np.array_equal(noalloc_labels(ua, t, n), want), (n, transpose)


def test_noalloc_argument_checks(ua):
    from ursa_amd import _lib

    lib = ua.load()
    t = ua.Topology.build(3, [0], [1])
    host = np.zeros(3, np.uint32)
    rc = lib.ursa_hip_cc_weak_noalloc(t._h, host.ctypes.data_as(C.c_void_p))
    assert rc == _lib.EINVAL
    big = ua.Topology.build((1 << 22) + 1, [0], [1 << 22])
    p = C.c_void_p()
    _lib.check(lib.ursa_hip_device_alloc(4 * ((1 << 22) + 1), C.byref(p)))
    try:
        assert lib.ursa_hip_cc_weak_noalloc(big._h, p) == _lib.ERANGE
    finally:
        lib.ursa_hip_device_free(p)
    empty = ua.Topology.build(0, [], [])
    assert lib.ursa_hip_cc_weak_noalloc(empty._h, None) == _lib.OK
