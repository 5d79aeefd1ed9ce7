"""GPU: concurrent calls from several host threads (the C ABI allows any number of
threads on one topology; each thread has its own streams, and the lazy builds run
once per topology, topology.rs:257-260). Every thread's results must equal the
single-thread results bit for bit, also while the threads race to build a fresh
topology's caches and reuse each other's freed device blocks.
"""
import threading

import numpy as np
import pytest

from graphs import rmat_dense

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ua():
    import ursa_amd

    ursa_amd.init(0)
    return ursa_amd


def _run_threads(work, nthreads):
    errors = []

    def body(k):
        try:
            work(k)
        except Exception as e:  # reported on the main thread
            errors.append(e)

    threads = [threading.Thread(target=body, args=(k,)) for k in range(nthreads)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[0]


def test_concurrent_calls_on_one_topology(ua):
    n, s, d = rmat_dense(14, 5)
    t = ua.Topology.build(n, s, d)
    srcs = [0, 3, 7, 11]
    want = {
        "tri": ua.triangle_count(t),
        "cc": ua.connected_components_weak(t),
        "bfs": ua.bfs_multi(t, srcs),
        "pr": ua.pagerank(t),
    }

    def work(k):
        for it in range(4):
            assert np.array_equal(ua.triangle_count(t), want["tri"]), (k, it, "triangles")
            assert np.array_equal(ua.connected_components_weak(t), want["cc"]), (k, it, "components")
            assert np.array_equal(ua.bfs_multi(t, srcs), want["bfs"]), (k, it, "bfs")
            assert np.array_equal(ua.pagerank(t), want["pr"]), (k, it, "pagerank")

    _run_threads(work, 4)


def test_concurrent_lazy_builds(ua):
    # A fresh topology: the threads race to build the undirected view, the
    # triangle orientation and the transpose on first use.
    n, s, d = rmat_dense(13, 9)
    ref = ua.Topology.build(n, s, d)
    want_tri = ua.triangle_count(ref)
    want_in = ua.bfs_multi(ref, [1, 2], ua.Direction.In)
    t = ua.Topology.build(n, s, d)

    def work(k):
        if k % 2:
            assert np.array_equal(ua.triangle_count(t), want_tri), (k, "triangles")
            assert np.array_equal(ua.bfs_multi(t, [1, 2], ua.Direction.In), want_in), (k, "bfs in")
        else:
            assert np.array_equal(ua.bfs_multi(t, [1, 2], ua.Direction.In), want_in), (k, "bfs in")
            assert np.array_equal(ua.triangle_count(t), want_tri), (k, "triangles")

    _run_threads(work, 4)


def test_concurrent_pull_lists_per_direction(ua):
    # A fresh topology: six threads, two per direction, race to build the
    # multi-source pull lists, which are kept per direction on the topology.
    n, s, d = rmat_dense(14, 5)
    srcs = np.random.default_rng(7).integers(0, n, 70)
    dirs = (ua.Direction.Out, ua.Direction.In, ua.Direction.Both)
    ref = ua.Topology.build(n, s, d)
    want = [ua.bfs_multi(ref, srcs, x, mode=2) for x in dirs]
    t = ua.Topology.build(n, s, d)

    def work(k):
        for r in range(3):
            j = (k + r) % 3
            assert np.array_equal(ua.bfs_multi(t, srcs, dirs[j], mode=2), want[j]), (k, r, j)

    _run_threads(work, 6)
