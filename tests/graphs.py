"""Shared graph cases for the GPU parity tests: (name, n, src_dense, dst_dense)."""
import os

import numpy as np

import oracle as orc

FX = np.load(os.path.join(os.path.dirname(__file__), "golden", "nx_fixtures.npz"))


def cases():
    out = []
    for name in ("gnp100", "sparse400", "ba500", "gnp300w"):
        n, s, d, _ = orc.intern_i64(FX[f"{name}_src"], FX[f"{name}_dst"])
        out.append((name, n, s, d))
    out.append(("diamond", 3, np.array([0, 0, 1, 2]), np.array([1, 2, 2, 0])))
    out.append(("loops_dups", 3, np.array([0, 0, 1, 2, 1, 0]), np.array([1, 2, 2, 0, 1, 1])))
    rng = np.random.default_rng(17)
    out.append(("multigraph", 3000, rng.integers(0, 3000, 40000), rng.integers(0, 3000, 40000)))
    # One hub row spanning ~15 merge-path tiles (leaves -> 0, with random edges
    # among the leaves so their ranks differ), plus a long chain out of the hub.
    leaves = 30000
    la = rng.integers(1, leaves + 1, leaves)
    lb = rng.integers(1, leaves + 1, leaves)
    s = np.concatenate([np.arange(1, leaves + 1), la, [0], np.arange(leaves + 1, leaves + 200)])
    d = np.concatenate([np.zeros(leaves, np.int64), lb, [leaves + 1], np.arange(leaves + 2, leaves + 201)])
    out.append(("hub", leaves + 201, s, d))
    out.append(("isolated_only", 5, np.array([], np.int64), np.array([], np.int64)))
    return out


def rmat_dense(scale, seed, symmetrize=False, edge_factor=16):
    us, ud = orc.gen_rmat(scale, edge_factor, seed=seed, symmetrize=symmetrize)
    n, s, d, _ = orc.intern_i64(us, ud)
    return n, s, d
