"""ctypes binding of ``libursa_hip.so`` (the C ABI declared in ``include/ursa_hip.h``,
``include/ursa_hip_ext.h`` and ``include/ursa_hip_serve.h``).

The library is built in-tree by ``ursa_amd/build.py`` (hipcc, gfx950). Importing this
module fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libursa_hip.so")

OK = 0
EINVAL = -22
ENOMEM = -12
EDEVICE = -5
ERANGE = -34
EINTERNAL = -1


class UrsaHipError(RuntimeError):
    """A non-zero status from the C ABI, with the library's thread-local message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


P = C.c_void_p
u64, u32, i32, f64, cint = C.c_uint64, C.c_uint32, C.c_int32, C.c_double, C.c_int

# name -> (restype, argtypes); every symbol the headers declare.
SIGNATURES = {
    "ursa_hip_init": (cint, [cint]),
    "ursa_hip_shutdown": (cint, []),
    "ursa_hip_last_error": (C.c_char_p, []),
    "ursa_hip_synchronize": (cint, []),
    "ursa_hip_topology_create": (cint, [u64, u64, P, P, C.POINTER(P)]),
    "ursa_hip_topology_destroy": (cint, [P]),
    "ursa_hip_topology_size": (cint, [P, C.POINTER(u64), C.POINTER(u64)]),
    "ursa_hip_topology_incoming": (cint, [P]),
    "ursa_hip_topology_undirected": (cint, [P, C.POINTER(u64)]),
    "ursa_hip_topology_export": (cint, [P, cint, P, P, P]),
    "ursa_hip_degree": (cint, [P, cint, P]),
    "ursa_hip_pagerank": (cint, [P, P, f64, u32, f64, P, C.POINTER(u32)]),
    "ursa_hip_pagerank_mixed": (cint, [P, P, f64, u32, f64, P, C.POINTER(u32), C.POINTER(u32)]),
    "ursa_hip_bfs": (cint, [P, P, u32, cint, P]),
    "ursa_hip_bfs_ex": (cint, [P, P, u32, cint, P, C.POINTER(u64), cint]),
    "ursa_hip_k_hop": (cint, [P, P, u64, u32, cint, C.POINTER(P)]),
    "ursa_hip_pairs_size": (cint, [P, C.POINTER(u64)]),
    "ursa_hip_pairs_export": (cint, [P, P, P]),
    "ursa_hip_pairs_destroy": (cint, [P]),
    "ursa_hip_shortest_path": (cint, [P, u32, u32, cint, P, u64, C.POINTER(u64)]),
    "ursa_hip_betweenness": (cint, [P, cint, f64, cint, u64, P]),
    "ursa_hip_betweenness_sources": (cint, [P, P, u64, cint, f64, P]),
    "ursa_hip_betweenness_sources_ex": (cint, [P, P, u64, cint, f64, u32, P, C.POINTER(u64)]),
    "ursa_hip_triangle_count": (cint, [P, P]),
    "ursa_hip_shuffled_order": (cint, [u64, u64, P]),
    "ursa_hip_cc_weak": (cint, [P, P]),
    "ursa_hip_profile_enable": (cint, [cint]),
    "ursa_hip_profile_reset": (cint, []),
    "ursa_hip_profile_query": (cint, [C.c_char_p, C.POINTER(f64), C.POINTER(u64), C.POINTER(f64)]),
    "ursa_hip_profile_names": (u64, [C.c_char_p, u64]),
    "ursa_hip_copy_probe": (cint, [P, P, u64]),
    "ursa_hip_gather_probe": (cint, [P, u64, P, u64, P]),
    "ursa_hip_device_alloc": (cint, [u64, C.POINTER(P)]),
    "ursa_hip_device_free": (cint, [P]),
    "ursa_hip_memcpy_to_host": (cint, [P, P, u64]),
    "ursa_hip_memcpy_to_device": (cint, [P, P, u64]),
    "ursa_hip_gen_rmat": (cint, [u32, u32, f64, f64, f64, u64, cint, C.POINTER(P)]),
    "ursa_hip_gen_er": (cint, [u64, u32, u64, cint, C.POINTER(P)]),
    "ursa_hip_edges_size": (cint, [P, C.POINTER(u64), C.POINTER(u64)]),
    "ursa_hip_edges_export": (cint, [P, P, P, P, P, P]),
    "ursa_hip_edges_dense_ptrs": (cint, [P, C.POINTER(P), C.POINTER(P)]),
    "ursa_hip_edges_topology": (cint, [P, C.POINTER(P)]),
    "ursa_hip_edges_destroy": (cint, [P]),
    "ursa_hip_gen_weights": (cint, [u64, u64, P]),
    "ursa_hip_ppr": (cint, [P, P, u64, f64, u32, f64, P, C.POINTER(u32)]),
    "ursa_hip_graphs_enable": (cint, [cint]),
    "ursa_hip_triangle_bitmaps": (cint, [cint]),
    "ursa_hip_triangle_work": (cint, [P, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
    "ursa_hip_pagerank_schedule": (cint, [cint]),
    "ursa_hip_ppr_schedule": (cint, [cint]),
    "ursa_hip_exact_sum": (cint, [P, u64, P]),
    "ursa_hip_set_stream": (cint, [P]),
    "ursa_hip_intern_i64": (cint, [u64, P, P, P, P, P, C.POINTER(u64)]),
    "ursa_hip_bfs_source_scan": (cint, [P, P, u64, cint, C.POINTER(f64), C.POINTER(u64), C.POINTER(i32)]),
    "ursa_hip_avg_path_length": (cint, [P, cint, f64, C.POINTER(f64)]),
    "ursa_hip_diameter": (cint, [P, cint, C.POINTER(C.c_int64)]),
    "ursa_hip_closeness": (cint, [P, P, u64, P]),
    "ursa_hip_clustering_coefficient": (cint, [P, P]),
    "ursa_hip_topology_create_from_csr": (cint, [u64, u64, P, P, P, C.POINTER(P)]),
    "ursa_hip_autotune": (cint, [P, cint, cint, C.POINTER(f64), C.POINTER(f64), C.POINTER(u32), C.POINTER(cint)]),
    "ursa_hip_ordered_sum": (cint, [P, u64, C.POINTER(f64)]),
    "ursa_hip_gather_edge_weights": (cint, [P, cint, P, P]),
    "ursa_hip_closeness_weighted": (cint, [P, P, P, u64, P]),
    "ursa_hip_host_alloc": (cint, [u64, C.POINTER(P)]),
    "ursa_hip_host_free": (cint, [P]),
    "ursa_hip_ppr_batch": (cint, [P, u32, P, P, f64, u32, f64, P, P]),
    "ursa_hip_ppr_multi": (cint, [u32, P, P, P, f64, u32, f64, P, P]),
    "ursa_hip_server_create": (cint, [P, C.POINTER(P)]),
    "ursa_hip_server_add_graph": (cint, [P, P, C.POINTER(u32)]),
    "ursa_hip_server_graph_bytes": (cint, [P, C.POINTER(u64)]),
    "ursa_hip_server_run": (cint, [P, P, u64, cint, P]),
    "ursa_hip_server_result_data": (cint, [P, u64, P, u64, C.POINTER(u64)]),
    "ursa_hip_server_destroy": (cint, [P]),
    "ursa_hip_digest": (u64, [P, u64]),
    "ursa_hip_cc_weak_noalloc": (cint, [P, P]),
    "ursa_hip_intern_i64_batches": (cint, [u64, P, P, P, P, P, P, C.POINTER(u64)]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (once). Raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc, gfx950). There is no CPU fallback."
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != OK:
        msg = load().ursa_hip_last_error()
        raise UrsaHipError(status, msg.decode() if msg else "")
