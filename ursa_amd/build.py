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
    return lib


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 8))
