"""CPU checks of the C ABI: the library loads without a GPU and exports every symbol
the public headers declare, and the Python binding covers all of them."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include"))) if h.endswith(".h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ursa_hip_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_entry_points():
    syms = declared_symbols()
    assert "ursa_hip_pagerank" in syms and "ursa_hip_topology_create" in syms


def test_library_exports_every_declared_symbol():
    from ursa_amd import _lib

    lib = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_every_declared_symbol():
    from ursa_amd import _lib

    assert declared_symbols() == set(_lib.SIGNATURES), set(_lib.SIGNATURES) ^ declared_symbols()


def test_errors_without_device_are_reported_not_fatal():
    import ursa_amd as ua

    lib = ua.load()
    h = C.c_void_p()
    # Null output handle is rejected before any device work.
    assert lib.ursa_hip_topology_create(1, 0, None, None, None) == ua._lib.EINVAL
    assert b"NULL" in lib.ursa_hip_last_error()
    # Bad direction is rejected before any device work.
    assert lib.ursa_hip_degree(None, 0, None) == ua._lib.EINVAL
    del h


def test_product_path_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "ursa_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text and "orc_" not in text, f
