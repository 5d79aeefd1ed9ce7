"""Write tests/golden/tri_rmat22.json: per-node triangle counts of the R-MAT scale-22
integer-kernel graph, as a digest the GPU's full-size output is checked against.

Run from the repository root (about 20 minutes on 8 cores):
    python tests/golden/make_tri_rmat22.py [threads]

The graph: R-MAT scale 22, edge factor 16, (a,b,c,d) = (0.57,0.19,0.19,0.05),
seed 2, self-loops and duplicates dropped, directed, int64 ids interned in
first-seen order (ursa-core/src/id_map.rs:220-292). Counts come from the oracle's
restatement of the reference's node-iterator count over the undirected view
(ursa-core/src/algo/triangle.rs:23-63, accumulated in u64), which at this size
takes the oracle tens of minutes; the fixture keeps that result so the check costs
a hash instead.

Stored: n, m, the undirected slot count, the triangle total (sum / 3), the SHA-256
of the u64 counts in dense-id order (little-endian bytes), and the exact counts of
4096 seeded nodes plus the 64 nodes of largest undirected degree.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as orc  # noqa: E402


def main() -> None:
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    t0 = time.time()
    us, ud = orc.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=2)
    n, sd, dd, _ = orc.intern_i64(us, ud)
    del us, ud
    t = orc.Topology(n, sd, dd)
    off, _ = t.undirected()
    und_slots = int(off[-1])
    print(f"graph: n={n} m={t.m} undirected slots={und_slots} ({time.time() - t0:.1f}s)", flush=True)
    t1 = time.time()
    tri = orc.triangle_count(t, nthreads=threads)
    print(f"oracle triangle_count on {threads} threads: {time.time() - t1:.1f}s", flush=True)
    total = int(tri.sum(dtype=np.uint64))
    assert total % 3 == 0
    rng = np.random.default_rng(22)
    sample = np.unique(rng.integers(0, n, 4096))
    deg = np.diff(off.astype(np.int64))
    hubs = np.argsort(-deg, kind="stable")[:64]
    nodes = np.unique(np.concatenate([sample, hubs]))
    out = {
        "graph": "R-MAT scale 22, edge factor 16, (0.57,0.19,0.19,0.05), seed 2, dedup, no self-loops",
        "n": int(n),
        "m": int(t.m),
        "undirected_slots": und_slots,
        "triangles": total // 3,
        "sum_counts": total,
        "sha256_u64_le": hashlib.sha256(tri.astype("<u8").tobytes()).hexdigest(),
        "nodes": nodes.tolist(),
        "counts": tri[nodes].astype(np.uint64).tolist(),
        "oracle_threads": threads,
        "oracle_seconds": round(time.time() - t1, 1),
    }
    with open(os.path.join(HERE, "tri_rmat22.json"), "w") as f:
        json.dump(out, f)
    print(f"triangles={total // 3} sha256={out['sha256_u64_le']}", flush=True)


if __name__ == "__main__":
    main()
