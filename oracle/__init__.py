"""CPU oracle for the CSR hot path — TEST INFRASTRUCTURE ONLY.

A ctypes view of ``liboracle.so`` (built from ``oracle.cpp`` by ``oracle/Makefile``),
the C++ restatement of ursa-core's kernels. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
module, and only as the checker or the timed CPU baseline: the product path
(``ursa_amd``) never imports it.

Parity status: every function here is pinned by a check in ``tests/test_oracle.py``
(reference golden values, networkx/scipy fixtures committed under
``tests/golden/``, closed forms or brute force). None is unpinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OUT, IN, BOTH = 0, 1, 2
_DIRS = {"out": OUT, "in": IN, "both": BOTH, OUT: OUT, IN: IN, BOTH: BOTH}


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load() -> C.CDLL:
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    P = C.c_void_p
    u64, u32, i32, f64, i = C.c_uint64, C.c_uint32, C.c_int32, C.c_double, C.c_int
    sig = {
        "orc_topo_create": (P, [u64, u64, P, P]),
        "orc_topo_destroy": (None, [P]),
        "orc_topo_n": (u64, [P]),
        "orc_topo_m": (u64, [P]),
        "orc_topo_adjacency": (None, [P, i, P, P, P]),
        "orc_topo_undirected": (u64, [P, P, P]),
        "orc_intern_i64": (u64, [u64, P, P, P, P, P]),
        "orc_splitmix_stream": (None, [u64, u64, P]),
        "orc_shuffled_order": (None, [u64, u64, P]),
        "orc_sample_indices": (u64, [u64, u64, i, u64, P]),
        "orc_sample_sources": (u64, [u64, i, f64, i, u64, P]),
        "orc_default_seed": (u64, []),
        "orc_degree": (None, [P, i, P]),
        "orc_pagerank": (None, [P, P, f64, u32, f64, i, P, P, P]),
        "orc_bfs": (None, [P, u32, i, P]),
        "orc_ppr": (None, [P, P, u64, f64, u32, f64, P, P]),
        "orc_stats_sources": (u64, [u64, i, f64, P]),
        "orc_bfs_source_scan": (None, [P, P, u64, i, P, P, P]),
        "orc_avg_path_length": (f64, [P, i, f64]),
        "orc_diameter": (C.c_int64, [P, i]),
        "orc_closeness": (None, [P, P, u64, P]),
        "orc_closeness_weighted": (None, [P, P, P, u64, P]),
        "orc_clustering": (None, [P, P]),
        "orc_k_hop": (u64, [P, P, u64, u32, i, P, P, u64]),
        "orc_shortest_path": (u64, [P, u32, u32, i, P, u64]),
        "orc_betweenness_sources": (None, [P, P, u64, i, f64, i, P]),
        "orc_betweenness": (None, [P, i, f64, i, u64, i, P]),
        "orc_triangle_count": (None, [P, i, P]),
        "orc_cc_weak": (None, [P, P]),
        "orc_canonicalize_labels": (None, [u64, P, P]),
        "orc_gen_rmat": (u64, [u32, u32, f64, f64, f64, u64, i, P, P, u64]),
        "orc_gen_er": (u64, [u64, u32, u64, i, P, P, u64]),
        "orc_gen_weights": (None, [u64, u64, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class Topology:
    """ursa-core/src/topology.rs:251-404 restated: out-CSR, lazy incoming, lazy undirected."""

    def __init__(self, n: int, src, dst):
        self.src = _u32(src)
        self.dst = _u32(dst)
        if self.src.shape != self.dst.shape:
            raise ValueError("src/dst length mismatch")
        self.n = int(n)
        self.m = int(self.src.shape[0])
        h = lib().orc_topo_create(self.n, self.m, _p(self.src), _p(self.dst))
        if not h:
            raise ValueError("dense endpoint out of range or edge cap exceeded")
        self._h = C.c_void_p(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.orc_topo_destroy(h)
            self._h = None

    def adjacency(self, which: str = "out"):
        off = np.zeros(self.n + 1, np.uint64)
        tg = np.zeros(self.m, np.uint32)
        ei = np.zeros(self.m, np.uint32)
        lib().orc_topo_adjacency(self._h, 0 if which == "out" else 1, _p(off), _p(tg), _p(ei))
        return off, tg, ei

    def undirected(self):
        cnt = lib().orc_topo_undirected(self._h, None, None)
        off = np.zeros(self.n + 1, np.uint64)
        tg = np.zeros(cnt, np.uint32)
        lib().orc_topo_undirected(self._h, _p(off), _p(tg))
        return off, tg


def degree(t: Topology, direction="out") -> np.ndarray:
    out = np.zeros(t.n, np.uint32)
    lib().orc_degree(t._h, _DIRS[direction], _p(out))
    return out


def pagerank(t: Topology, damping=0.85, max_iter=30, tol=1e-6, weights=None, nthreads=1):
    """Returns (rank f64[n], iterations, last L1 delta)."""
    out = np.zeros(t.n, np.float64)
    it = C.c_uint32(0)
    dl = C.c_double(0.0)
    w = None
    if weights is not None:
        w = np.ascontiguousarray(np.asarray(weights, np.float64))
        if w.shape[0] != t.m:
            raise ValueError("weights length must equal the edge count")
    lib().orc_pagerank(t._h, _p(w), damping, max_iter, tol, nthreads, _p(out), C.byref(it), C.byref(dl))
    return out, it.value, dl.value


def ppr(t: Topology, seeds, damping=0.85, max_iter=100, tol=1e-8):
    """Personalized PageRank -> (f64[n], iterations)."""
    s = _u32(seeds)
    out = np.zeros(max(t.n, 1), np.float64)
    it = C.c_uint32(0)
    lib().orc_ppr(t._h, _p(s) if s.size else None, s.shape[0], damping, max_iter, tol, _p(out), C.byref(it))
    return out[: t.n], it.value


def stats_sources(n: int, frac=None) -> np.ndarray:
    out = np.zeros(max(n, 1), np.uint32)
    k = lib().orc_stats_sources(n, frac is not None, 0.0 if frac is None else frac, _p(out))
    return out[:k].copy()


def bfs_source_scan(t: Topology, sources, direction="out"):
    s = _u32(sources)
    sm, c, mx = C.c_double(0), C.c_uint64(0), C.c_int32(0)
    lib().orc_bfs_source_scan(t._h, _p(s) if s.size else None, s.shape[0], _DIRS[direction], C.byref(sm),
                              C.byref(c), C.byref(mx))
    return sm.value, c.value, mx.value


def avg_path_length(t: Topology, sample=None) -> float:
    return lib().orc_avg_path_length(t._h, sample is not None, 0.0 if sample is None else sample)


def diameter(t: Topology, approximate=True) -> int:
    return int(lib().orc_diameter(t._h, int(approximate)))


def closeness(t: Topology, sources=None) -> np.ndarray:
    s = _u32(np.arange(t.n) if sources is None else sources)
    out = np.zeros(max(s.shape[0], 1), np.float64)
    lib().orc_closeness(t._h, _p(s) if s.size else None, s.shape[0], _p(out))
    return out[: s.shape[0]]


def closeness_weighted(t: Topology, weights, sources=None) -> np.ndarray:
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    assert w.shape[0] == t.m, "weights length must equal the edge count"
    s = _u32(np.arange(t.n) if sources is None else sources)
    out = np.zeros(max(s.shape[0], 1), np.float64)
    lib().orc_closeness_weighted(t._h, _p(w) if w.size else None, _p(s) if s.size else None, s.shape[0], _p(out))
    return out[: s.shape[0]]


def clustering(t: Topology) -> np.ndarray:
    out = np.zeros(max(t.n, 1), np.float64)
    lib().orc_clustering(t._h, _p(out))
    return out[: t.n]


def bfs(t: Topology, source: int, direction="out") -> np.ndarray:
    out = np.zeros(t.n, np.int32)
    lib().orc_bfs(t._h, source, _DIRS[direction], _p(out))
    return out


def k_hop(t: Topology, seeds, k: int, direction="out"):
    s = _u32(seeds)
    cnt = lib().orc_k_hop(t._h, _p(s), s.shape[0], k, _DIRS[direction], None, None, 0)
    a = np.zeros(cnt, np.uint32)
    b = np.zeros(cnt, np.uint32)
    lib().orc_k_hop(t._h, _p(s), s.shape[0], k, _DIRS[direction], _p(a), _p(b), cnt)
    return a, b


def shortest_path(t: Topology, s: int, tgt: int, direction="out"):
    buf = np.zeros(max(t.n, 1), np.uint32)
    ln = lib().orc_shortest_path(t._h, s, tgt, _DIRS[direction], _p(buf), buf.shape[0])
    return None if ln == 0 else buf[:ln].copy()


def betweenness(t: Topology, sample=None, seed=None, nthreads=1) -> np.ndarray:
    out = np.zeros(t.n, np.float64)
    lib().orc_betweenness(t._h, sample is not None, 0.0 if sample is None else sample,
                          seed is not None, 0 if seed is None else seed, nthreads, _p(out))
    return out


def betweenness_sources(t: Topology, sources, scale=None, nthreads=1) -> np.ndarray:
    s = _u32(sources)
    out = np.zeros(t.n, np.float64)
    lib().orc_betweenness_sources(t._h, _p(s), s.shape[0], scale is not None,
                                  1.0 if scale is None else scale, nthreads, _p(out))
    return out


def triangle_count(t: Topology, nthreads=1) -> np.ndarray:
    out = np.zeros(t.n, np.uint64)
    lib().orc_triangle_count(t._h, nthreads, _p(out))
    return out


def cc_weak(t: Topology) -> np.ndarray:
    out = np.zeros(t.n, np.uint32)
    lib().orc_cc_weak(t._h, _p(out))
    return out


def canonicalize(labels) -> np.ndarray:
    lab = _u32(labels)
    out = np.zeros_like(lab)
    lib().orc_canonicalize_labels(lab.shape[0], _p(lab), _p(out))
    return out


def splitmix_stream(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, np.uint64)
    lib().orc_splitmix_stream(seed, count, _p(out))
    return out


def shuffled_order(n: int, seed: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_shuffled_order(n, seed, _p(out))
    return out


def sample_indices(r: int, k: int, seed=None) -> np.ndarray:
    out = np.zeros(max(min(k, r), 0), np.uint64)
    got = lib().orc_sample_indices(r, k, seed is not None, 0 if seed is None else seed, _p(out) if out.size else None)
    return out[:got]


def sample_sources(n: int, frac=None, seed=None) -> np.ndarray:
    out = np.zeros(max(n, 1), np.uint32)
    k = lib().orc_sample_sources(n, frac is not None, 0.0 if frac is None else frac,
                                 seed is not None, 0 if seed is None else seed, _p(out))
    return out[:k].copy()


def default_seed() -> int:
    return int(lib().orc_default_seed())


def intern_i64(src, dst):
    s = np.ascontiguousarray(np.asarray(src, np.int64))
    d = np.ascontiguousarray(np.asarray(dst, np.int64))
    m = s.shape[0]
    sd = np.zeros(m, np.uint32)
    dd = np.zeros(m, np.uint32)
    users = np.zeros(max(2 * m, 1), np.int64)
    n = lib().orc_intern_i64(m, _p(s), _p(d), _p(sd), _p(dd), _p(users))
    return int(n), sd, dd, users[:n].copy()


def intern_i64_batches(batches):
    """IdMap::from_edge_batches (ursa-core/src/id_map.rs:220-292): ids are global across
    the batches and rows are numbered in batch order, which is the single-array
    interning of the rows laid end to end (id_map.rs:529-550 asserts exactly that)."""
    bs = [(np.asarray(s, np.int64), np.asarray(d, np.int64)) for s, d in batches]
    if not bs:
        return 0, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.int64)
    return intern_i64(np.concatenate([s for s, _ in bs]), np.concatenate([d for _, d in bs]))


def gen_rmat(scale: int, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, symmetrize=False):
    cap = (edge_factor << scale) * (2 if symmetrize else 1)
    s = np.zeros(cap, np.int64)
    d = np.zeros(cap, np.int64)
    m = lib().orc_gen_rmat(scale, edge_factor, a, b, c, seed, int(symmetrize), _p(s), _p(d), cap)
    return s[:m].copy(), d[:m].copy()


def gen_er(n: int, edge_factor: int, seed: int, symmetrize=False):
    cap = max(edge_factor * n * (2 if symmetrize else 1), 1)
    s = np.zeros(cap, np.int64)
    d = np.zeros(cap, np.int64)
    m = lib().orc_gen_er(n, edge_factor, seed, int(symmetrize), _p(s), _p(d), cap)
    return s[:m].copy(), d[:m].copy()


def gen_weights(m: int, seed: int) -> np.ndarray:
    out = np.zeros(m, np.float64)
    lib().orc_gen_weights(m, seed, _p(out))
    return out
