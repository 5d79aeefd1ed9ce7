#!/usr/bin/env python3
"""Benchmark of the CSR hot path on one MI355X.

Default workload: pull PageRank (f64) on R-MAT scale 24, edge factor 16,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 3, deduplicated, self-loops removed,
damping 0.85, tol 1e-10 (L1), max_iter 100. The graph is generated and interned on
the device, the topology (CSR + transpose) is built before timing, so every timed
step starts with its inputs resident in HBM and writes its output to HBM.

A step is one full PageRank solve (all iterations until delta < tol). `value` is
GTEPS = m * iterations / step time. The roofline object describes the SpMV tile
kernel (k_pr_tiles), timed with HIP events on its launch stream inside the library;
its algorithmic bytes are 12 B per edge + 32 B per node per iteration. The CPU
baseline is the oracle (a C++ restatement of ursa-core's rayon PageRank) on the
same graph for a bounded number of iterations, with all host threads.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch measured by rocprofv3 PMC runs, if committed."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def bench_pagerank(args, ua, gen):
    scale = args.scale or 24
    seed = 3
    params = ua.PageRankParams(0.85, 100, 1e-10)
    t0 = time.time()
    e = gen.rmat(scale, 16, 0.57, 0.19, 0.19, seed=seed)
    topo = e.topology()
    lib = ua.load()
    ua._lib.check(lib.ursa_hip_topology_incoming(topo._h))
    n, m = e.n, e.m
    log(f"graph R-MAT s{scale}: n={n} m={m} built in {time.time() - t0:.1f}s")

    out = C.c_void_p()
    ua._lib.check(lib.ursa_hip_device_alloc(8 * n, C.byref(out)))
    iters = C.c_uint32(0)

    def step():
        ua._lib.check(lib.ursa_hip_pagerank(topo._h, None, params.damping, params.max_iter, params.tol, out,
                                            C.byref(iters)))

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup}: iters={iters.value}")
    ua._lib.check(lib.ursa_hip_synchronize())
    t_start = time.perf_counter()
    for i in range(args.steps):
        step()
    ua._lib.check(lib.ursa_hip_synchronize())
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed / args.steps * 1e3
    gteps = m * iters.value * args.steps / elapsed / 1e9
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, iters={iters.value}, {gteps:.2f} GTEPS")

    # Roofline of the dominant kernel: HIP events around each k_pr_tiles launch.
    ua.profile.enable(True)
    ua.profile.reset()
    step()
    q = ua.profile.query("pagerank_spmv")
    ua.profile.enable(False)
    avg_ms = q["total_ms"] / max(q["launches"], 1)
    per_launch_bytes = q["algo_bytes"] / max(q["launches"], 1)
    achieved = per_launch_bytes / (avg_ms / 1e3) / 1e9
    log(f"k_pr_tiles: {q['launches']} launches, avg {avg_ms * 1e3:.1f} us, {achieved:.0f} GB/s algorithmic")
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles("k_pr_tiles", f"pagerank_rmat{scale}"),
        "kernel": "k_pr_tiles",
        "avg_launch_us": round(avg_ms * 1e3, 2),
        "algo_bytes_per_launch": per_launch_bytes,
    }

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_pagerank(e, params, args.cpu_iters)
    ua._lib.check(lib.ursa_hip_device_free(out))
    config = {
        "workload": f"pagerank_f64 rmat{scale}",
        "graph": f"R-MAT scale {scale}, edge factor 16, (0.57,0.19,0.19,0.05), seed {seed}, dedup, no self-loops",
        "n": n,
        "m": m,
        "damping": params.damping,
        "tol": params.tol,
        "max_iter": params.max_iter,
        "iterations": iters.value,
    }
    return {
        "metric": "GTEPS (pull PageRank SpMV, m*iterations/s)",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "ms_per_step": round(ms_per_step, 3),
        "dtype": "f64",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def cpu_baseline_pagerank(e, params, cpu_iters):
    import oracle as orc

    threads = os.cpu_count() or 1
    t0 = time.time()
    _, _, sd, dd, _ = e.export()
    t = orc.Topology(e.n, sd, dd)
    orc.degree(t, "in")  # builds the transpose outside the timed region
    log(f"cpu baseline: oracle topology built in {time.time() - t0:.1f}s")
    t1 = time.perf_counter()
    _, it, _ = orc.pagerank(t, params.damping, cpu_iters, params.tol, nthreads=threads)
    el = time.perf_counter() - t1
    gteps = e.m * it / el / 1e9
    log(f"cpu baseline: {it} iterations in {el:.2f}s on {threads} threads: {gteps:.3f} GTEPS")
    return {
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"same graph, {it} PageRank iterations (max_iter={cpu_iters}) of the oracle's "
                  f"rayon-structured pull loop, {threads} threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="pagerank", choices=["pagerank"])
    ap.add_argument("--scale", type=int, default=None)
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md, not implemented)"}))
        sys.exit(2)

    import ursa_amd as ua
    from ursa_amd import gen

    ua.init(0)
    res = bench_pagerank(args, ua, gen)
    line = {
        "metric": res["metric"],
        "value": res["value"],
        "unit": res["unit"],
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": res["dtype"],
        "data": "synthetic (seeded R-MAT generated on the device)",
        "config": res["config"],
        "roofline": res["roofline"],
        "cpu_baseline": res["cpu_baseline"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
