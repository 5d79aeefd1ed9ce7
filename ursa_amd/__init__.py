"""ursa_amd — MI355X (gfx950) backend for ursa-core's CSR algorithm kernels.

The Python surface mirrors ``ursa_core`` (ursa-core/src/lib.rs:25-30,
ursa-core/src/algo/mod.rs:35-50): ``Topology.build``, ``Direction``,
``PageRankParams`` and one function per kernel with the reference's name,
argument meaning and edge-case behaviour. Every call goes through the C ABI of
``libursa_hip.so``; there is no CPU path.

Outputs are numpy arrays on the host. Arguments documented as "array" accept
numpy arrays (host) or integer device addresses via the ``*_ptr`` helpers used
by the benchmark.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import UrsaHipError, check, load

__all__ = [
    "Direction",
    "PageRankParams",
    "Topology",
    "UrsaHipError",
    "degree",
    "intern_i64",
    "intern_i64_batches",
    "bfs_source_scan",
    "avg_path_length",
    "diameter",
    "closeness",
    "closeness_weighted",
    "clustering_coefficient",
    "pagerank",
    "pagerank_weighted",
    "pagerank_mixed",
    "ordered_sum",
    "gather_edge_weights",
    "PinnedBuffer",
    "ppr",
    "ppr_batch",
    "ppr_multi",
    "bfs_distances",
    "bfs_multi",
    "autotune",
    "k_hop",
    "shortest_path",
    "betweenness",
    "betweenness_sources",
    "triangle_count",
    "connected_components_weak",
    "init",
    "profile",
]


class Direction(enum.IntEnum):
    """ursa-core/src/topology.rs:27-32"""

    Out = 0
    In = 1
    Both = 2


def _dir(d) -> int:
    if isinstance(d, str):
        return {"out": 0, "in": 1, "both": 2}[d.lower()]
    return int(d)


@dataclass(frozen=True)
class PageRankParams:
    """ursa-core/src/algo/pagerank.rs:21-36 (defaults :28-35)."""

    damping: float = 0.85
    max_iter: int = 30
    tol: float = 1e-6


def init(device: int = 0) -> None:
    check(load().ursa_hip_init(device))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class Topology:
    """Device-resident CSR topology (ursa-core/src/topology.rs:251-404).

    ``build`` groups the edge rows by source with original row order kept inside each
    segment (topology.rs:109-114); the transpose and the undirected view are built on
    first use and cached.
    """

    def __init__(self, handle: C.c_void_p, n: int, m: int):
        self._h = handle
        self._n = n
        self._m = m

    @classmethod
    def build(cls, n_nodes: int, src_dense, dst_dense) -> "Topology":
        """Topology::build (topology.rs:276-298)."""
        s = _u32(src_dense)
        d = _u32(dst_dense)
        if s.shape != d.shape:
            raise ValueError("src/dst length mismatch")
        return cls._from_ptrs(int(n_nodes), int(s.shape[0]), _ptr(s), _ptr(d))

    @classmethod
    def _from_ptrs(cls, n: int, m: int, src_ptr, dst_ptr) -> "Topology":
        lib = load()
        h = C.c_void_p()
        check(lib.ursa_hip_topology_create(n, m, src_ptr, dst_ptr, C.byref(h)))
        return cls(h, n, m)

    @classmethod
    def from_csr(cls, n_nodes: int, offsets, targets, edge_ids) -> "Topology":
        """A topology from an existing out-CSR (Adjacency, topology.rs:35-47): offsets
        u64[n+1], targets u32[m], edge_ids u32[m] (a permutation of the edge rows)."""
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
        tg = _u32(targets)
        ei = _u32(edge_ids)
        if off.shape[0] != int(n_nodes) + 1 or tg.shape != ei.shape:
            raise ValueError("offsets must have n+1 entries and targets/edge_ids equal length")
        h = C.c_void_p()
        check(load().ursa_hip_topology_create_from_csr(int(n_nodes), tg.shape[0], _ptr(off), _ptr(tg), _ptr(ei),
                                                       C.byref(h)))
        return cls(h, int(n_nodes), int(tg.shape[0]))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib._lib is not None:
            _lib._lib.ursa_hip_topology_destroy(h)
            self._h = None

    def n_nodes(self) -> int:
        return self._n

    def n_edges(self) -> int:
        return self._m

    def _export(self, which: int, slots: int, with_eids: bool):
        off = np.zeros(self._n + 1, np.uint64)
        tg = np.zeros(max(slots, 1), np.uint32)
        ei = np.zeros(max(slots, 1), np.uint32) if with_eids else None
        check(load().ursa_hip_topology_export(self._h, which, _ptr(off), _ptr(tg), _ptr(ei) if with_eids else None))
        return off, tg[:slots], (ei[:slots] if with_eids else None)

    def out(self):
        """(offsets u64[n+1], targets u32[m], edge_ids u32[m]) — Topology::out."""
        return self._export(0, self._m, True)

    def incoming(self):
        """Topology::incoming (topology.rs:343-348)."""
        check(load().ursa_hip_topology_incoming(self._h))
        return self._export(1, self._m, True)

    def undirected(self):
        """Topology::undirected (topology.rs:353-389): (offsets u64[n+1], targets u32[slots])."""
        slots = C.c_uint64(0)
        check(load().ursa_hip_topology_undirected(self._h, C.byref(slots)))
        off, tg, _ = self._export(2, slots.value, False)
        return off, tg


def bfs_source_scan(topo: Topology, sources, direction=Direction.Out):
    """bfs_source_scan (ursa-plan/src/stats.rs:100-140) -> (sum f64, count, max)."""
    src = _u32(sources)
    sm, c, mx = C.c_double(0), C.c_uint64(0), C.c_int32(0)
    check(load().ursa_hip_bfs_source_scan(topo._h, _ptr(src) if src.size else None, src.shape[0], _dir(direction),
                                          C.byref(sm), C.byref(c), C.byref(mx)))
    return sm.value, c.value, mx.value


def avg_path_length(topo: Topology, sample=None) -> float:
    """avg_path_length (ursa-plan/src/stats.rs:142-152)."""
    out = C.c_double(0)
    check(load().ursa_hip_avg_path_length(topo._h, sample is not None, 0.0 if sample is None else sample,
                                          C.byref(out)))
    return out.value


def diameter(topo: Topology, approximate: bool = True) -> int:
    """diameter (ursa-plan/src/stats.rs:154-168)."""
    out = C.c_int64(0)
    check(load().ursa_hip_diameter(topo._h, int(approximate), C.byref(out)))
    return out.value


def closeness(topo: Topology, sources=None) -> np.ndarray:
    """closeness (ursa-core/src/algo/closeness.rs:69-137, unweighted) for every node,
    or for the given sources -> f64."""
    if sources is None:
        out = np.zeros(max(topo.n_nodes(), 1), np.float64)
        check(load().ursa_hip_closeness(topo._h, None, 0, _ptr(out)))
        return out[: topo.n_nodes()]
    src = _u32(sources)
    out = np.zeros(max(src.shape[0], 1), np.float64)
    check(load().ursa_hip_closeness(topo._h, _ptr(src), src.shape[0], _ptr(out)))
    return out[: src.shape[0]]


def closeness_weighted(topo: Topology, weights, sources=None) -> np.ndarray:
    """closeness_weighted (ursa-core/src/algo/closeness.rs:143-222): shortest-path costs
    over out-edges with weights[edge row], for every node or the given sources -> f64."""
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    if w.shape[0] != topo.n_edges():
        raise ValueError("weights length must equal the edge count")
    wp = _ptr(w) if w.size else None
    if sources is None:
        out = np.zeros(max(topo.n_nodes(), 1), np.float64)
        check(load().ursa_hip_closeness_weighted(topo._h, wp, None, 0, _ptr(out)))
        return out[: topo.n_nodes()]
    src = _u32(sources)
    out = np.zeros(max(src.shape[0], 1), np.float64)
    check(load().ursa_hip_closeness_weighted(topo._h, wp, _ptr(src), src.shape[0], _ptr(out)))
    return out[: src.shape[0]]


def clustering_coefficient(topo: Topology) -> np.ndarray:
    """clustering_coefficient (ursa-core/src/algo/clustering.rs:21-40) -> f64[n]."""
    out = np.zeros(max(topo.n_nodes(), 1), np.float64)
    check(load().ursa_hip_clustering_coefficient(topo._h, _ptr(out)))
    return out[: topo.n_nodes()]


def intern_i64(src, dst):
    """IdMap interning on the device (ursa-core/src/id_map.rs:220-319): dense ids in
    first-seen order, src before dst per row -> (n, src_dense u32[m], dst_dense u32[m],
    user_ids i64[n])."""
    s = np.ascontiguousarray(np.asarray(src, dtype=np.int64))
    d = np.ascontiguousarray(np.asarray(dst, dtype=np.int64))
    if s.shape != d.shape:
        raise ValueError("src/dst length mismatch")
    m = s.shape[0]
    sd = np.zeros(max(m, 1), np.uint32)
    dd = np.zeros(max(m, 1), np.uint32)
    users = np.zeros(max(2 * m, 1), np.int64)
    n = C.c_uint64(0)
    check(load().ursa_hip_intern_i64(m, _ptr(s), _ptr(d), _ptr(sd), _ptr(dd), _ptr(users), C.byref(n)))
    return int(n.value), sd[:m], dd[:m], users[: n.value]


def intern_i64_batches(batches):
    """IdMap::from_edge_batches on the device (ursa-core/src/id_map.rs:207-292): interns
    a sequence of (src, dst) int64 batches in first-seen order across the batches, src
    before dst per row, without concatenating them -> (n, src_dense u32[m],
    dst_dense u32[m], user_ids i64[n]) with m the total row count."""
    keep = []
    for src, dst in batches:
        s = np.ascontiguousarray(np.asarray(src, dtype=np.int64))
        d = np.ascontiguousarray(np.asarray(dst, dtype=np.int64))
        if s.shape != d.shape or s.ndim != 1:
            raise ValueError("src/dst length mismatch")
        keep.append((s, d))
    nb = len(keep)
    m = sum(s.shape[0] for s, _ in keep)
    lens = np.array([s.shape[0] for s, _ in keep] or [0], np.uint64)
    srcs = (C.c_void_p * max(nb, 1))(*[s.ctypes.data for s, _ in keep])
    dsts = (C.c_void_p * max(nb, 1))(*[d.ctypes.data for _, d in keep])
    sd = np.zeros(max(m, 1), np.uint32)
    dd = np.zeros(max(m, 1), np.uint32)
    users = np.zeros(max(2 * m, 1), np.int64)
    n = C.c_uint64(0)
    check(load().ursa_hip_intern_i64_batches(nb, _ptr(lens), C.cast(srcs, C.c_void_p), C.cast(dsts, C.c_void_p),
                                             _ptr(sd), _ptr(dd), _ptr(users), C.byref(n)))
    return int(n.value), sd[:m], dd[:m], users[: n.value]


def degree(topo: Topology, direction=Direction.Out) -> np.ndarray:
    """degree (ursa-core/src/algo/degree.rs:12-28) -> u32[n]."""
    out = np.zeros(max(topo.n_nodes(), 1), np.uint32)
    check(load().ursa_hip_degree(topo._h, _dir(direction), _ptr(out)))
    return out[: topo.n_nodes()]


def _pagerank(topo: Topology, weights, params: PageRankParams):
    n = topo.n_nodes()
    out = np.zeros(max(n, 1), np.float64)
    it = C.c_uint32(0)
    w = None
    if weights is not None:
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
        if w.shape[0] != topo.n_edges():
            raise ValueError("weights length must equal the edge count")  # pagerank.rs:140-144
    check(load().ursa_hip_pagerank(topo._h, _ptr(w) if w is not None else None, params.damping,
                                   params.max_iter, params.tol, _ptr(out), C.byref(it)))
    return out[:n], it.value


def pagerank_mixed(topo: Topology, params: PageRankParams = PageRankParams(), weights=None):
    """PageRank with f32 compensated sweeps refined by f64 sweeps -> (f64[n], iterations, f32 iterations)."""
    n = topo.n_nodes()
    out = np.zeros(max(n, 1), np.float64)
    it, it32 = C.c_uint32(0), C.c_uint32(0)
    w = None
    if weights is not None:
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
        if w.shape[0] != topo.n_edges():
            raise ValueError("weights length must equal the edge count")
    check(load().ursa_hip_pagerank_mixed(topo._h, _ptr(w) if w is not None else None, params.damping,
                                         params.max_iter, params.tol, _ptr(out), C.byref(it), C.byref(it32)))
    return out[:n], it.value, it32.value


def ppr(topo: Topology, seeds, damping=0.85, max_iter=100, tol=1e-8):
    """Personalized PageRank (networkx personalization semantics; f32 compensated
    sweeps refined in f64) -> (f64[n], iterations)."""
    n = topo.n_nodes()
    s = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint32))
    out = np.zeros(max(n, 1), np.float64)
    it = C.c_uint32(0)
    check(load().ursa_hip_ppr(topo._h, _ptr(s) if s.size else None, s.shape[0], damping, max_iter, tol,
                              _ptr(out), C.byref(it)))
    return out[:n], it.value


def ppr_batch(topo: Topology, seed_lists, damping=0.85, max_iter=100, tol=1e-8):
    """Several personalized PageRank queries on one graph solved together (16 per
    device pass) -> (f64[k, n], iterations u32[k]); row q equals ppr(seed_lists[q])."""
    n = topo.n_nodes()
    k = len(seed_lists)
    offs = np.zeros(k + 1, np.uint64)
    for i, sl in enumerate(seed_lists):
        offs[i + 1] = offs[i] + len(sl)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint32) for x in seed_lists])
                                if k else np.zeros(0, np.uint32))
    out = np.zeros((k, max(n, 1)), np.float64)
    iters = np.zeros(max(k, 1), np.uint32)
    check(load().ursa_hip_ppr_batch(topo._h, k, _ptr(offs), _ptr(flat) if flat.size else None, damping, max_iter,
                                    tol, _ptr(out), _ptr(iters)))
    return out[:, :n], iters[:k]


def ppr_multi(topos, seed_lists, damping=0.85, max_iter=100, tol=1e-8):
    """One personalized PageRank query per graph, small graphs swept together in
    cross-graph passes -> ([f64[n_i]], iterations u32[k]); entry i equals
    ppr(topos[i], seed_lists[i])."""
    k = len(topos)
    if len(seed_lists) != k:
        raise ValueError("one seed list per topology")
    offs = np.zeros(k + 1, np.uint64)
    for i, sl in enumerate(seed_lists):
        offs[i + 1] = offs[i] + len(sl)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint32) for x in seed_lists])
                                if k else np.zeros(0, np.uint32))
    outs = [np.zeros(max(t.n_nodes(), 1), np.float64) for t in topos]
    handles = (C.c_void_p * max(k, 1))(*[t._h for t in topos])
    optrs = (C.c_void_p * max(k, 1))(*[o.ctypes.data for o in outs])
    iters = np.zeros(max(k, 1), np.uint32)
    check(load().ursa_hip_ppr_multi(k, handles, _ptr(offs), _ptr(flat) if flat.size else None, damping, max_iter,
                                    tol, optrs, _ptr(iters)))
    return [o[:t.n_nodes()] for o, t in zip(outs, topos)], iters[:k]


def pagerank(topo: Topology, params: PageRankParams = PageRankParams()) -> np.ndarray:
    """pagerank (ursa-core/src/algo/pagerank.rs:98-124) -> f64[n]."""
    return _pagerank(topo, None, params)[0]


def pagerank_weighted(topo: Topology, weights, params: PageRankParams = PageRankParams()) -> np.ndarray:
    """pagerank_weighted (ursa-core/src/algo/pagerank.rs:135-169) -> f64[n]."""
    return _pagerank(topo, weights, params)[0]


def bfs_multi(topo: Topology, sources, direction=Direction.Out, mode: int = 0) -> np.ndarray:
    """bfs_distances from each source -> i32[k, n] (ursa-core/src/algo/bfs.rs:18-42).

    mode 0 = automatic, 1 = one traversal per source, 2 = 64-lane multi-source BFS.
    """
    src = _u32(sources)
    k, n = src.shape[0], topo.n_nodes()
    out = np.zeros((max(k, 1), max(n, 1)), np.int32)
    re = C.c_uint64(0)
    if k and n:
        check(load().ursa_hip_bfs_ex(topo._h, _ptr(src), k, _dir(direction), _ptr(out), C.byref(re), mode))
    return out[:k, :n]


def autotune(topo: Topology, direction=Direction.Out, force: bool = False) -> dict:
    """Pick the traversal schedule (BFS direction switch, multi-source push rule) for
    this graph by timing candidates, or reuse the choice cached for its degree
    distribution. Results of every traversal are unchanged."""
    a, b = C.c_double(0), C.c_double(0)
    pd, src = C.c_uint32(0), C.c_int(0)
    check(load().ursa_hip_autotune(topo._h, _dir(direction), int(force), C.byref(a), C.byref(b), C.byref(pd),
                                   C.byref(src)))
    return {"alpha": a.value, "beta": b.value, "push_div": pd.value,
            "source": {0: "default", 1: "measured", 2: "cached"}[src.value]}


def gather_edge_weights(topo: Topology, weights, which="in") -> np.ndarray:
    """Per-edge-row weights (evaluate_weight's output, ursa-plan/src/weight.rs:51-106)
    put in adjacency-slot order, w[edge_ids[k]], for the out- or in-CSR; negative
    weights raise UrsaHipError (code -22) as evaluate_weight does."""
    m = topo.n_edges()
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    if w.shape[0] != m:
        raise ValueError("weights length must equal the edge count")
    out = np.zeros(max(m, 1), np.float64)
    check(load().ursa_hip_gather_edge_weights(topo._h, {"out": 0, "in": 1}[which], _ptr(w) if m else None,
                                              _ptr(out) if m else None))
    return out[:m]


class PinnedBuffer:
    """Pinned host memory from ursa_hip_host_alloc (the backing store an Arrow result
    array can take over, ursa-plan/src/result.rs:84-139). ``array(dtype, count)`` views
    it; the views must not outlive ``free()``."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load().ursa_hip_host_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    def array(self, dtype, count: int) -> np.ndarray:
        dt = np.dtype(dtype)
        if count * dt.itemsize > self.nbytes or self.ptr is None:
            raise ValueError("view larger than the buffer")
        raw = (C.c_char * (count * dt.itemsize)).from_address(self.ptr)
        return np.frombuffer(raw, dtype=dt, count=count)

    def free(self):
        if self.ptr is not None:
            check(load().ursa_hip_host_free(C.c_void_p(self.ptr)))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ordered_sum(x) -> float:
    """((0.0 + x[0]) + x[1]) + ... rounded step by step, as PageRank's dangling mass
    is summed (ursa-core/src/algo/pagerank.rs:78), computed on the device."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    out = C.c_double(0.0)
    check(load().ursa_hip_ordered_sum(_ptr(a) if a.size else None, a.shape[0], C.byref(out)))
    return out.value


def bfs_distances(topo: Topology, source: int, direction=Direction.Out) -> np.ndarray:
    """bfs_distances (ursa-core/src/algo/bfs.rs:18-42) -> i32[n]; -1 = unreachable."""
    n = topo.n_nodes()
    if n == 0:
        return np.zeros(0, np.int32)
    return bfs_multi(topo, [source], direction)[0]


def k_hop(topo: Topology, seeds, k: int, direction=Direction.Out):
    """k_hop (ursa-core/src/algo/hop.rs:52-75) -> (seeds_out u32[], reached u32[])."""
    s = _u32(seeds)
    lib = load()
    h = C.c_void_p()
    check(lib.ursa_hip_k_hop(topo._h, _ptr(s) if s.size else None, s.shape[0], k, _dir(direction), C.byref(h)))
    try:
        cnt = C.c_uint64(0)
        check(lib.ursa_hip_pairs_size(h, C.byref(cnt)))
        a = np.zeros(max(cnt.value, 1), np.uint32)
        b = np.zeros(max(cnt.value, 1), np.uint32)
        check(lib.ursa_hip_pairs_export(h, _ptr(a), _ptr(b)))
    finally:
        lib.ursa_hip_pairs_destroy(h)
    return a[: cnt.value], b[: cnt.value]


def shortest_path(topo: Topology, source: int, target: int, direction=Direction.Out):
    """shortest_path (ursa-core/src/algo/bfs.rs:48-102) -> u32[] or None."""
    cap = max(topo.n_nodes(), 1)
    buf = np.zeros(cap, np.uint32)
    ln = C.c_uint64(0)
    check(load().ursa_hip_shortest_path(topo._h, source, target, _dir(direction), _ptr(buf), cap, C.byref(ln)))
    return None if ln.value == 0 else buf[: ln.value].copy()


def betweenness(topo: Topology, sample=None, seed=None) -> np.ndarray:
    """betweenness (ursa-core/src/algo/betweenness.rs:153-177) -> f64[n]."""
    n = topo.n_nodes()
    out = np.zeros(max(n, 1), np.float64)
    check(load().ursa_hip_betweenness(topo._h, int(sample is not None), 0.0 if sample is None else float(sample),
                                      int(seed is not None), 0 if seed is None else int(seed), _ptr(out)))
    return out[:n]


def betweenness_sources(topo: Topology, sources, scale=None, batch: int = 64) -> np.ndarray:
    """Brandes over explicit sources (betweenness.rs:45-70), optional scale; batch = sources per pass."""
    n = topo.n_nodes()
    src = _u32(sources)
    out = np.zeros(max(n, 1), np.float64)
    re = C.c_uint64(0)
    check(load().ursa_hip_betweenness_sources_ex(topo._h, _ptr(src) if src.size else None, src.shape[0],
                                                 int(scale is not None), 1.0 if scale is None else float(scale),
                                                 batch, _ptr(out), C.byref(re)))
    return out[:n]


def triangle_count(topo: Topology) -> np.ndarray:
    """triangle_count (ursa-core/src/algo/triangle.rs:58-63) -> u64[n]."""
    n = topo.n_nodes()
    out = np.zeros(max(n, 1), np.uint64)
    check(load().ursa_hip_triangle_count(topo._h, _ptr(out)))
    return out[:n]


def connected_components_weak(topo: Topology) -> np.ndarray:
    """connected_components_weak (ursa-core/src/algo/components.rs:58-68) -> u32[n] (min-id labels)."""
    n = topo.n_nodes()
    out = np.zeros(max(n, 1), np.uint32)
    check(load().ursa_hip_cc_weak(topo._h, _ptr(out)))
    return out[:n]


def shuffled_order(n: int, seed: int) -> np.ndarray:
    """The reference's seeded permutation of 0..n-1 (ursa-core/src/algo/rng.rs:64-73)."""
    out = np.zeros(max(n, 1), np.uint32)
    check(load().ursa_hip_shuffled_order(n, seed, _ptr(out)))
    return out[:n]


class profile:
    """Per-kernel event timing inside the library (ursa_hip_profile_*)."""

    @staticmethod
    def enable(on: bool = True) -> None:
        check(load().ursa_hip_profile_enable(int(on)))

    @staticmethod
    def reset() -> None:
        check(load().ursa_hip_profile_reset())

    @staticmethod
    def query(kernel: str):
        ms, n, b = C.c_double(0), C.c_uint64(0), C.c_double(0)
        check(load().ursa_hip_profile_query(kernel.encode(), C.byref(ms), C.byref(n), C.byref(b)))
        return {"total_ms": ms.value, "launches": n.value, "algo_bytes": b.value}

    @staticmethod
    def names():
        lib = load()
        need = lib.ursa_hip_profile_names(None, 0)
        buf = C.create_string_buffer(int(need) + 1)
        lib.ursa_hip_profile_names(buf, need + 1)
        return [x for x in buf.value.decode().split("\n") if x]
