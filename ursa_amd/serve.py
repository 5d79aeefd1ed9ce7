"""Request server over resident device graphs (``include/ursa_hip_serve.h``).

``Server`` wraps the C++ server: graphs are made resident with ``add_graph`` and a
list of ``Query`` records is run with ``run`` (all at once, or at their arrival
times), returning one ``Result`` record per query: status, iteration count, result
length, a 64-bit digest of the result bytes and its arrival/start/done times.

``serving_graph_specs`` and ``serving_trace`` synthesise the serving mix: graphs
with node counts log-uniform in [1e3, 2e6], edge factor uniform in {4..32}, R-MAT
(0.57, 0.19, 0.19, 0.05) or Erdos-Renyi at 50/50, seeds 1000+i; Poisson arrivals
with 40% personalized PageRank (1-16 seeds, damping 0.85, tol 1e-8), 30% 3-hop
neighbourhoods, 20% shortest paths between uniform node pairs and 10% triangle
counts.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._lib import check, load
from . import Direction, Topology

PPR, KHOP, PATH, TRIANGLES = 0, 1, 2, 3
KIND_NAMES = {PPR: "ppr", KHOP: "khop", PATH: "path", TRIANGLES: "triangles"}


class Query(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("graph", C.c_uint32),
        ("nseeds", C.c_uint32),
        ("k", C.c_uint32),
        ("seeds", C.c_uint32 * 16),
        ("direction", C.c_int32),
        ("max_iter", C.c_uint32),
        ("damping", C.c_double),
        ("tol", C.c_double),
        ("arrival_s", C.c_double),
    ]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("iters", C.c_uint32),
        ("len", C.c_uint64),
        ("digest", C.c_uint64),
        ("arrival_s", C.c_double),
        ("start_s", C.c_double),
        ("done_s", C.c_double),
        ("batch", C.c_uint32),
        ("slot", C.c_uint32),
    ]


class Config(C.Structure):
    _fields_ = [
        ("workspace_bytes", C.c_uint64),
        ("graph_bytes", C.c_uint64),
        ("slots", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("serial", C.c_int32),
        ("keep_results", C.c_int32),
        ("cross_batch", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


def ppr_query(graph: int, seeds, damping=0.85, max_iter=100, tol=1e-8, arrival_s=0.0) -> Query:
    q = Query(kind=PPR, graph=graph, damping=damping, max_iter=max_iter, tol=tol, arrival_s=arrival_s)
    _set_seeds(q, seeds)
    return q


def khop_query(graph: int, seeds, k=3, direction=Direction.Out, arrival_s=0.0) -> Query:
    q = Query(kind=KHOP, graph=graph, k=k, direction=int(direction), arrival_s=arrival_s)
    _set_seeds(q, seeds)
    return q


def path_query(graph: int, source: int, target: int, direction=Direction.Out, arrival_s=0.0) -> Query:
    q = Query(kind=PATH, graph=graph, direction=int(direction), arrival_s=arrival_s)
    _set_seeds(q, [source, target])
    return q


def triangles_query(graph: int, arrival_s=0.0) -> Query:
    return Query(kind=TRIANGLES, graph=graph, arrival_s=arrival_s)


def _set_seeds(q: Query, seeds) -> None:
    seeds = [int(s) for s in seeds]
    if not 1 <= len(seeds) <= 16:
        raise ValueError("1..16 seeds per query")
    q.nseeds = len(seeds)
    for i, s in enumerate(seeds):
        q.seeds[i] = s


def digest(data) -> int:
    """ursa_hip_digest of the bytes of a numpy array (the digest results carry)."""
    a = np.ascontiguousarray(data)
    return int(load().ursa_hip_digest(a.ctypes.data_as(C.c_void_p) if a.nbytes else None, a.nbytes))


class Server:
    def __init__(self, slots=8, max_batch=16, workspace_bytes=0, graph_bytes=0, serial=False, keep_results=False,
                 cross_batch=0):
        cfg = Config(workspace_bytes=workspace_bytes, graph_bytes=graph_bytes, slots=slots, max_batch=max_batch,
                     serial=int(serial), keep_results=int(keep_results), cross_batch=cross_batch)
        h = C.c_void_p()
        check(load().ursa_hip_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._graphs: list[Topology] = []

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            check(load().ursa_hip_server_destroy(h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_graph(self, topo: Topology) -> int:
        gid = C.c_uint32(0)
        check(load().ursa_hip_server_add_graph(self._h, topo._h, C.byref(gid)))
        self._graphs.append(topo)  # the server borrows the topology
        return gid.value

    def graph_bytes(self) -> int:
        b = C.c_uint64(0)
        check(load().ursa_hip_server_graph_bytes(self._h, C.byref(b)))
        return b.value

    def run(self, queries, timed_arrivals=False):
        qs = (Query * len(queries))(*queries)
        res = (Result * len(queries))()
        check(load().ursa_hip_server_run(self._h, qs, len(queries), int(timed_arrivals), res))
        return list(res)

    def result_bytes(self, index: int) -> bytes:
        nb = C.c_uint64(0)
        check(load().ursa_hip_server_result_data(self._h, index, None, 0, C.byref(nb)))
        buf = C.create_string_buffer(max(nb.value, 1))
        check(load().ursa_hip_server_result_data(self._h, index, buf, nb.value, C.byref(nb)))
        return buf.raw[: nb.value]

    def result(self, index: int, query: Query):
        """Decoded result of query `index` of the last run (keep_results)."""
        raw = self.result_bytes(index)
        if query.kind == PPR:
            return np.frombuffer(raw, np.float64).copy()
        if query.kind == TRIANGLES:
            return np.frombuffer(raw, np.uint64).copy()
        w = np.frombuffer(raw, np.uint32).copy()
        if query.kind == KHOP:
            h = w.shape[0] // 2
            return w[:h], w[h:]
        return w if w.size else None


@dataclass
class GraphSpec:
    index: int
    kind: str  # "rmat" or "er"
    n_target: int
    edge_factor: int
    seed: int
    scale: int = 0


def serving_graph_specs(count=512, n_lo=1e3, n_hi=2e6, ef_lo=4, ef_hi=32, seed_base=1000):
    """Graph i: its own PCG64 stream seeded 1000+i draws n (log-uniform), the edge
    factor and the family. R-MAT uses scale round(log2 n) (nodes are those an edge
    names, so fewer than 2^scale); Erdos-Renyi draws edge_factor * n edges over n ids."""
    specs = []
    for i in range(count):
        rng = np.random.default_rng(seed_base + i)
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        ef = int(rng.integers(ef_lo, ef_hi + 1))
        kind = "rmat" if rng.random() < 0.5 else "er"
        sp = GraphSpec(i, kind, n, ef, seed_base + i)
        if kind == "rmat":
            sp.scale = max(1, int(round(math.log2(n))))
        specs.append(sp)
    return specs


def build_serving_graph(spec: GraphSpec) -> Topology:
    from . import gen

    if spec.kind == "rmat":
        e = gen.rmat(spec.scale, spec.edge_factor, seed=spec.seed)
    else:
        e = gen.erdos_renyi(spec.n_target, spec.edge_factor, spec.seed)
    t = e.topology()
    del e
    return t


def serving_trace(node_counts, count, qps, seed=7, mix=(0.4, 0.3, 0.2, 0.1), khop_k=3):
    """Poisson arrivals at `qps` over graphs chosen uniformly; returns a list of Query."""
    rng = np.random.default_rng(seed)
    node_counts = np.asarray(node_counts, np.int64)
    gaps = rng.exponential(1.0 / qps, size=count) if qps > 0 else np.zeros(count)
    arrivals = np.cumsum(gaps) - (gaps[0] if count else 0.0)
    kinds = rng.choice(4, size=count, p=np.asarray(mix) / np.sum(mix))
    graphs = rng.integers(0, node_counts.shape[0], size=count)
    out = []
    for i in range(count):
        g = int(graphs[i])
        n = int(node_counts[g])
        a = float(arrivals[i])
        kd = int(kinds[i])
        if kd == PPR:
            ns = int(rng.integers(1, 17))
            out.append(ppr_query(g, rng.integers(0, n, size=ns), arrival_s=a))
        elif kd == KHOP:
            out.append(khop_query(g, [int(rng.integers(0, n))], k=khop_k, arrival_s=a))
        elif kd == PATH:
            s, t = rng.integers(0, n, size=2)
            out.append(path_query(g, int(s), int(t), arrival_s=a))
        else:
            out.append(triangles_query(g, arrival_s=a))
    return out


def latency_summary(results):
    lat = np.array([r.done_s - r.arrival_s for r in results])
    span = max(r.done_s for r in results) - min(r.arrival_s for r in results) if results else 0.0
    return {
        "queries": len(results),
        "qps": (len(results) / span) if span > 0 else None,
        "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
        "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
        "mean_ms": float(lat.mean() * 1e3) if lat.size else None,
        "makespan_s": span,
        "failed": int(sum(1 for r in results if r.status != 0)),
    }
