"""Seeded synthetic graphs generated on the device (R-MAT, Erdős–Rényi).

See ``include/ursa_hip_ext.h`` for the stream definitions. The generated int64 edge
list is interned to dense ids in ursa's first-seen order, so ``Edges.topology()`` is
the topology ursa would build from the same edge frame.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, load
from . import Topology


class Edges:
    def __init__(self, handle: C.c_void_p):
        self._h = handle
        n, m = C.c_uint64(0), C.c_uint64(0)
        check(load().ursa_hip_edges_size(self._h, C.byref(n), C.byref(m)))
        self.n = n.value
        self.m = m.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().ursa_hip_edges_destroy(h)
            self._h = None

    def export(self):
        """(user_src i64[m], user_dst i64[m], src_dense u32[m], dst_dense u32[m], user_ids i64[n])."""
        us = np.zeros(max(self.m, 1), np.int64)
        ud = np.zeros(max(self.m, 1), np.int64)
        sd = np.zeros(max(self.m, 1), np.uint32)
        dd = np.zeros(max(self.m, 1), np.uint32)
        ids = np.zeros(max(self.n, 1), np.int64)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        check(load().ursa_hip_edges_export(self._h, p(us), p(ud), p(sd), p(dd), p(ids)))
        return us[: self.m], ud[: self.m], sd[: self.m], dd[: self.m], ids[: self.n]

    def topology(self) -> Topology:
        h = C.c_void_p()
        check(load().ursa_hip_edges_topology(self._h, C.byref(h)))
        return Topology(h, self.n, self.m)


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19, seed: int = 1,
         symmetrize: bool = False) -> Edges:
    h = C.c_void_p()
    check(load().ursa_hip_gen_rmat(scale, edge_factor, a, b, c, seed, int(symmetrize), C.byref(h)))
    return Edges(h)


def erdos_renyi(n: int, edge_factor: int, seed: int, symmetrize: bool = False) -> Edges:
    h = C.c_void_p()
    check(load().ursa_hip_gen_er(n, edge_factor, seed, int(symmetrize), C.byref(h)))
    return Edges(h)


def weights(m: int, seed: int) -> np.ndarray:
    out = np.zeros(max(m, 1), np.float64)
    check(load().ursa_hip_gen_weights(m, seed, out.ctypes.data_as(C.c_void_p)))
    return out[:m]
