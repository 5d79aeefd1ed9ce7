"""Build recipe for the in-tree HIP library (``ursa_amd/libursa_hip.so``).

``python -m ursa_amd.build`` compiles every HIP source for gfx950 with hipcc via
``ursa_amd/csrc/Makefile`` (always a full rebuild).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8) -> str:
    csrc = os.path.join(HERE, "csrc")
    subprocess.run(["make", "-B", f"-j{jobs}", "-C", csrc], check=True)
    lib = os.path.join(HERE, "libursa_hip.so")
    if not os.path.exists(lib):
        raise RuntimeError("build finished without producing libursa_hip.so")
    build_mirror_check()
    return lib


def build_mirror_check() -> str:
    """``ursa_amd/host/libursa_mirror.so``: the C++ mirror (host/ursa_core.hpp) driven by
    host/mirror_check.cpp, linked against libursa_hip.so, for the tests."""
    host = os.path.join(HERE, "host")
    out = os.path.join(host, "libursa_mirror.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-Wall", "-o", out,
                    os.path.join(host, "mirror_check.cpp"), "-L", HERE, "-lursa_hip", "-L/opt/rocm/lib",
                    "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return out


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 8))
