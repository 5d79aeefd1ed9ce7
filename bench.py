#!/usr/bin/env python3
"""Benchmark of ursa-core's CSR hot path on one MI355X. Prints one JSON line.

Default workload (the integer-kernel configuration): R-MAT scale 22, edge factor 16,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 2, self-loops and duplicates dropped,
directed, u32 CSR plus transpose. One step runs, on that graph:
  * degree, out and in                       -> u32[n] each
  * BFS from 64 sources (Out), i32 levels     -> i32[64, n]; the sources are the
    first 64 of the reference's seeded shuffle (seed 2) of the nodes with
    out-degree > 0, in dense-id order
  * weakly connected components               -> u32[n] (min-id labels)
  * triangle_count on the undirected view     -> u64[n]
The graph is generated and interned on the device and its CSR, transpose,
undirected view and triangle orientation are built before timing (ursa caches
them on the topology too), so a timed step starts with everything in HBM and
writes its outputs to HBM.

value = aggregate GTEPS of the step: (sum of BFS E_reach over the 64 sources
+ m for components + undirected m for triangles) / step time, the TEPS
numerators of each kernel. `kernels` gives per-kernel GTEPS and %HBM from HIP
events on the kernels' own stream. `roofline` describes the dominant kernel of
the step. `cpu_baseline` is the oracle's serial BFS (the reference's
bfs_distances loop) on a sample of the same sources, one core.

The `pagerank` section (R-MAT scale 24, seed 3, damping 0.85, tol 1e-10) times
the plain f64 solve and the f32+refinement solve on the same graph and reports
their error against the f64 oracle. Skip it with --no-pagerank.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(workload: str, kernel: str):
    """HBM bytes per launch measured by rocprofv3 PMC runs, if committed."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def traffic_source(workload: str):
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get(workload, {}).get("source")
    except (OSError, ValueError):
        return None


def traffic_total(t):
    """HBM bytes per launch (FETCH + WRITE, calibrated) from a traffic.json entry, or None."""
    return None if not t else t.get("total_bytes")


class Dev:
    """A device buffer owned by the library's allocator."""

    def __init__(self, lib, nbytes):
        self.lib = lib
        self.p = C.c_void_p()
        check(lib.ursa_hip_device_alloc(max(int(nbytes), 8), C.byref(self.p)))

    def free(self):
        if self.p:
            self.lib.ursa_hip_device_free(self.p)
            self.p = C.c_void_p()


def check(rc):
    from ursa_amd import _lib

    _lib.check(rc)


def kernel_stats(ua, names):
    tot_ms, launches, nbytes = 0.0, 0, 0.0
    for k in names:
        q = ua.profile.query(k)
        tot_ms += q["total_ms"]
        launches += q["launches"]
        nbytes += q["algo_bytes"]
    return tot_ms, launches, nbytes


# --------------------------------------------------------------------------- C2
def bench_c2(args, ua, gen):
    lib = ua.load()
    scale = args.scale or 22
    t0 = time.time()
    e = gen.rmat(scale, 16, 0.57, 0.19, 0.19, seed=2)
    topo = e.topology()
    n, m = e.n, e.m
    outdeg = ua.degree(topo, "out")
    cand = np.nonzero(outdeg > 0)[0].astype(np.uint32)
    perm = ua.shuffled_order(len(cand), 2)
    sources = np.ascontiguousarray(cand[perm[:64]])
    # Lazily built topology structures (cached on the topology, as in ursa).
    tb = time.time()
    check(lib.ursa_hip_topology_incoming(topo._h))
    t_in = time.time() - tb
    und_slots = C.c_uint64(0)
    tb = time.time()
    check(lib.ursa_hip_topology_undirected(topo._h, C.byref(und_slots)))
    t_und = time.time() - tb
    m_und = und_slots.value // 2
    log(f"graph R-MAT s{scale}: n={n} m={m} undirected m={m_und}; generated in {time.time() - t0:.1f}s "
        f"(transpose {t_in * 1e3:.0f} ms, undirected view {t_und * 1e3:.0f} ms)")
    # Traversal schedule (BFS switch thresholds, multi-source push rule) from the
    # autotuner, once per graph before any timing; results do not depend on it.
    tb = time.time()
    tuning = ua.autotune(topo)
    tuning["tuning_s"] = round(time.time() - tb, 3)
    log(f"autotuned traversal schedule: {tuning}")

    lv = Dev(lib, 4 * 64 * n)
    d_out, d_in, d_cc, d_tri = Dev(lib, 4 * n), Dev(lib, 4 * n), Dev(lib, 4 * n), Dev(lib, 8 * n)
    src_ptr = sources.ctypes.data_as(C.c_void_p)
    reached = C.c_uint64(0)
    walls = {}

    def step(record=False):
        def run(name, fn):
            if record:
                check(lib.ursa_hip_synchronize())
                w0 = time.perf_counter()
            fn()
            if record:
                check(lib.ursa_hip_synchronize())
                walls[name] = walls.get(name, 0.0) + time.perf_counter() - w0

        run("degree", lambda: (check(lib.ursa_hip_degree(topo._h, 0, d_out.p)),
                               check(lib.ursa_hip_degree(topo._h, 1, d_in.p))))
        run("bfs", lambda: check(lib.ursa_hip_bfs_ex(topo._h, src_ptr, 64, 0, lv.p, C.byref(reached), 0)))
        run("cc", lambda: check(lib.ursa_hip_cc_weak(topo._h, d_cc.p)))
        run("triangle", lambda: check(lib.ursa_hip_triangle_count(topo._h, d_tri.p)))

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup}")
    check(lib.ursa_hip_synchronize())
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
    check(lib.ursa_hip_synchronize())
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed / args.steps * 1e3
    e_reach = reached.value
    numer = e_reach + m + m_und
    gteps = numer * args.steps / elapsed / 1e9
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step; BFS E_reach={e_reach}; aggregate {gteps:.2f} GTEPS")

    # Per-kernel event timing (one profiled step) and wall time per call (one more).
    ua.profile.enable(True)
    ua.profile.reset()
    step()
    groups = {
        "bfs": ["bfs_top_down", "bfs_bottom_up", "msbfs_expand"],
        "cc": ["cc_hook_round", "cc_compress", "cc_hook_out", "cc_hook_in"],
        "triangle": ["triangle_count"],
        "degree": ["degree"],
    }
    stats = {g: kernel_stats(ua, ks) for g, ks in groups.items()}
    ua.profile.enable(False)
    step(record=True)
    tri_q = ua.profile.query("triangle_count")
    algo = {
        "bfs": 4.0 * e_reach + 12.0 * n * 64,
        "cc": stats["cc"][2],
        "triangle": stats["triangle"][2],
        "degree": stats["degree"][2],
    }
    teps = {"bfs": e_reach, "cc": m, "triangle": m_und, "degree": None}
    kernels = {}
    for g, (ms, launches, _) in stats.items():
        gbs = algo[g] / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        moved = traffic_total(traffic_from_profiles(f"c2_rmat{scale}", g))
        kernels[g] = {
            "kernel_ms": round(ms, 3),
            "call_ms": round(walls.get(g, 0.0) * 1e3, 3),
            "launches": launches,
            "algo_GB": round(algo[g] / 1e9, 4),
            "GBps": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "GTEPS": round(teps[g] / (ms / 1e3) / 1e9, 3) if teps[g] and ms > 0 else None,
        }
        if moved and ms > 0:
            # HBM bytes one call moves (rocprof counters, calibrated) over this run's kernel time
            kernels[g]["moved_GB"] = round(moved / 1e9, 4)
            kernels[g]["moved_GBps"] = round(moved / (ms / 1e3) / 1e9, 1)
            kernels[g]["moved_frac"] = round(moved / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    kernels["bfs"]["algo_model"] = ("4 E_reach + 12 n summed over the 64 sources; the 64 lanes share every edge "
                                    "scan, so moved_GB is the bus figure")
    # Triangle work: the walk model (what the one-sided probe reads) is the
    # algorithmic figure; the sorted-merge model (both lists of every pair) is
    # reported beside it.
    e_or, merge, walk = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
    check(lib.ursa_hip_triangle_work(topo._h, C.byref(e_or), C.byref(merge), C.byref(walk)))
    tri_ms = stats["triangle"][0]
    merge_bytes = 4.0 * merge.value + 12.0 * n
    kernels["triangle"]["algo_model"] = ("one-sided walk: 4 B per walked list entry + 12 B per oriented edge "
                                         "(partner id, two offsets) + 12 B per node")
    kernels["triangle"]["walk_steps"] = walk.value
    kernels["triangle"]["oriented_edges"] = e_or.value
    kernels["triangle"]["merge_model_GB"] = round(merge_bytes / 1e9, 4)
    if tri_ms > 0:
        kernels["triangle"]["merge_model_frac"] = round(merge_bytes / (tri_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    log("kernels: " + json.dumps(kernels))
    dom = max(("bfs", "cc", "triangle"), key=lambda g: stats[g][0])
    if dom == "bfs":
        ms = stats["bfs"][0]
        per_launch = algo["bfs"] / 64
        avg_us = ms / 64 * 1e3
        kname = "k_ms_push+k_ms_pull (64-lane multi-source BFS; per-traversal average)"
        launches = 64
    else:
        ms, launches, b = stats[dom]
        per_launch = b / max(launches, 1)
        avg_us = ms / max(launches, 1) * 1e3
        kname = {"cc": "k_cc_* (Afforest)",
                 "triangle": "k_tc_count + k_tc_count_big, out- and in-list kinds on two streams; "
                             "avg_launch_us is one call's span"}[dom]
    achieved = per_launch / (avg_us / 1e6) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_total(traffic_from_profiles(f"c2_rmat{scale}", dom)),
        "traffic_source": traffic_source(f"c2_rmat{scale}"),
        "kernel": kname,
        "avg_launch_us": round(avg_us, 2),
        "algo_bytes_per_launch": per_launch,
        "launches": launches,
    }
    host_step = c2_host_step(ua, lib, topo, n, src_ptr)
    cpu, cpu_kernels, parity = None, None, None
    if not args.no_cpu_baseline:
        cpu, cpu_kernels, parity = c2_cpu_leg(args, lib, e, sources, lv, d_out, d_in, d_cc, d_tri, scale)
    ingest = bench_ingest(args, lib, e)
    for b in (lv, d_out, d_in, d_cc, d_tri):
        b.free()
    config = {
        "workload": f"integer kernels on R-MAT scale {scale}: degree in/out, BFS x64 (Out), weak CC, triangle_count",
        "graph": f"R-MAT scale {scale}, edge factor 16, (0.57,0.19,0.19,0.05), seed 2, dedup, no self-loops",
        "n": n,
        "m": m,
        "m_undirected": m_und,
        "bfs_sources": 64,
        "bfs_E_reach": e_reach,
        "build_ms": {"transpose": round(t_in * 1e3, 1), "undirected": round(t_und * 1e3, 1)},
        "traversal_tuning": tuning,
    }
    return {
        "metric": "GTEPS (aggregate over the step's kernels: BFS E_reach + CC m + triangle m_und per second)",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "ms_per_step": round(ms_per_step, 3),
        "dtype": "u32/i32/u64 (integer kernels)",
        "config": config,
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "cpu_kernels": cpu_kernels,
        "parity": parity,
        "step_host_outputs": host_step,
        "tri_launch": tri_q,
        "ingest": ingest,
    }


def c2_host_step(ua, lib, topo, n, src_ptr, reps=3):
    """The same step through the C ABI with every output written to pinned host
    memory (ursa_hip_host_alloc, the buffers an Arrow array can own; result.rs:84-139):
    64 x n i32 levels, 2 x n u32 degrees, n u32 labels, n u64 triangle counts."""
    bufs = {"levels": ua.PinnedBuffer(4 * 64 * n), "out": ua.PinnedBuffer(4 * n), "in": ua.PinnedBuffer(4 * n),
            "cc": ua.PinnedBuffer(4 * n), "tri": ua.PinnedBuffer(8 * n)}
    reached = C.c_uint64(0)
    p = {k: C.c_void_p(b.ptr) for k, b in bufs.items()}

    def step():
        check(lib.ursa_hip_degree(topo._h, 0, p["out"]))
        check(lib.ursa_hip_degree(topo._h, 1, p["in"]))
        check(lib.ursa_hip_bfs_ex(topo._h, src_ptr, 64, 0, p["levels"], C.byref(reached), 0))
        check(lib.ursa_hip_cc_weak(topo._h, p["cc"]))
        check(lib.ursa_hip_triangle_count(topo._h, p["tri"]))

    step()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    ms = (time.perf_counter() - t0) / reps * 1e3
    out_mb = (4 * 64 + 4 * 3 + 8) * n / 1e6
    for b in bufs.values():
        b.free()
    log(f"step with pinned host outputs: {ms:.2f} ms ({out_mb:.0f} MB of results to host)")
    return {"ms_per_step": round(ms, 3), "host_output_MB": round(out_mb, 1)}


def bench_ingest(args, lib, e):
    """IdMap interning (id_map.rs:220-319) of the step graph's int64 edge list on the
    device, inputs and outputs in HBM; bytes = 16 per row read + 8 per row written +
    8 per distinct id written. CPU baseline: the oracle's serial hash-map interning."""
    us, ud, _, _, _ = e.export()
    m = int(us.shape[0])
    bufs = [Dev(lib, 8 * m), Dev(lib, 8 * m), Dev(lib, 4 * m), Dev(lib, 4 * m), Dev(lib, 16 * m), Dev(lib, 8)]
    check(lib.ursa_hip_memcpy_to_device(bufs[0].p, us.ctypes.data_as(C.c_void_p), 8 * m))
    check(lib.ursa_hip_memcpy_to_device(bufs[1].p, ud.ctypes.data_as(C.c_void_p), 8 * m))
    nn = C.c_uint64(0)
    times = []
    for _ in range(4):
        t0 = time.perf_counter()
        check(lib.ursa_hip_intern_i64(m, bufs[0].p, bufs[1].p, bufs[2].p, bufs[3].p, bufs[4].p, C.byref(nn)))
        times.append(time.perf_counter() - t0)
    el = min(times[1:])
    gb = (24.0 * m + 8.0 * nn.value) / 1e9
    res = {"rows": m, "ids": int(nn.value), "ms": round(el * 1e3, 3), "GBps": round(gb / el, 1),
           "hbm_frac": round(gb / el / HBM_PEAK_GBS, 4)}
    # The same rows handed over as 16 chunk pairs (IdMap::from_edge_batches,
    # id_map.rs:220-292), each a device pointer into the edge list.
    sd1, dd1 = host_copy(lib, bufs[2], np.uint32, m), host_copy(lib, bufs[3], np.uint32, m)
    nb = 16
    cut = [m * i // nb for i in range(nb + 1)]
    lens = np.array([cut[i + 1] - cut[i] for i in range(nb)], np.uint64)
    srcs = (C.c_void_p * nb)(*[bufs[0].p.value + 8 * cut[i] for i in range(nb)])
    dsts = (C.c_void_p * nb)(*[bufs[1].p.value + 8 * cut[i] for i in range(nb)])
    nb_n = C.c_uint64(0)
    times = []
    for _ in range(4):
        t0 = time.perf_counter()
        check(lib.ursa_hip_intern_i64_batches(nb, lens.ctypes.data_as(C.c_void_p), C.cast(srcs, C.c_void_p),
                                              C.cast(dsts, C.c_void_p), bufs[2].p, bufs[3].p, bufs[4].p,
                                              C.byref(nb_n)))
        times.append(time.perf_counter() - t0)
    res["batches16_ms"] = round(min(times[1:]) * 1e3, 3)
    res["batches16_equal"] = bool(nb_n.value == nn.value
                                  and np.array_equal(host_copy(lib, bufs[2], np.uint32, m), sd1)
                                  and np.array_equal(host_copy(lib, bufs[3], np.uint32, m), dd1))
    if not args.no_cpu_baseline:
        import oracle as orc

        t0 = time.perf_counter()
        rn, _, _, _ = orc.intern_i64(us, ud)
        res["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        res["cpu_ids_equal"] = rn == nn.value
    for b in bufs:
        b.free()
    log(f"ingest: {m} rows -> {nn.value} ids in {el * 1e3:.2f} ms ({gb / el:.0f} GB/s); "
        f"as {nb} batches {res['batches16_ms']:.2f} ms, equal {res['batches16_equal']}")
    return res


def host_copy(lib, dev, dtype, count):
    out = np.zeros(count, dtype)
    if count:
        check(lib.ursa_hip_memcpy_to_host(out.ctypes.data_as(C.c_void_p), dev.p, out.nbytes))
    return out


def oracle_bfs_rows(o, sources, threads, direction="out"):
    """The oracle's serial BFS (bfs.rs:18-42) for each source, `threads` sources at a
    time (ctypes releases the GIL during each call)."""
    import oracle as orc
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda x: orc.bfs(o, int(x), direction), sources))


def cpu_threads():
    """Host cores this process may use (the box's CPU share), capped at 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def c2_cpu_leg(args, lib, e, sources, lv, d_out, d_in, d_cc, d_tri, scale):
    """CPU baseline of the step on the oracle (the reference's algorithms restated in
    C++), timed on this host, and the parity check of every GPU output of the step
    against it at full size: 64 BFS level rows, canonical component labels and in/out
    degrees bit-exact; the u64 triangle counts against the committed oracle digest
    (tests/golden/tri_rmat22.json: the node-iterator count takes the oracle ~20 min
    at this size, so the digest was computed once in the build container)."""
    import hashlib

    import oracle as orc

    threads = cpu_threads()
    n = e.n
    t0 = time.perf_counter()
    _, _, sd, dd, _ = e.export()
    o = orc.Topology(n, sd, dd)
    build_s = time.perf_counter() - t0
    parity = {}
    per = {}
    # degrees
    t1 = time.perf_counter()
    r_out, r_in = orc.degree(o, "out"), orc.degree(o, "in")
    per["degree_in_out"] = {"s": round(time.perf_counter() - t1, 4), "threads": 1}
    parity["degree_out_equal"] = bool(np.array_equal(host_copy(lib, d_out, np.uint32, n), r_out))
    parity["degree_in_equal"] = bool(np.array_equal(host_copy(lib, d_in, np.uint32, n), r_in))
    # BFS: one thread (a sample, cold + warm) and all threads (every source)
    outdeg = r_out.astype(np.int64)
    k1 = max(1, min(args.cpu_sources, len(sources)))
    times1 = []
    for rep in range(3):
        t1 = time.perf_counter()
        oracle_bfs_rows(o, sources[:k1], 1)
        times1.append(time.perf_counter() - t1)
    t1 = time.perf_counter()
    rows = oracle_bfs_rows(o, sources, threads)
    tall = time.perf_counter() - t1
    reach_all = sum(int(outdeg[r >= 0].sum()) for r in rows)
    reach_1 = sum(int(outdeg[r >= 0].sum()) for r in rows[:k1])
    got = host_copy(lib, lv, np.int32, 64 * n).reshape(64, n)
    parity["bfs_rows_equal"] = int(sum(np.array_equal(got[i], rows[i]) for i in range(len(sources))))
    parity["bfs_rows"] = len(sources)
    del got
    per["bfs_1thread"] = {"sources": k1, "cold_s": round(times1[0], 4), "warm_median_s": round(float(np.median(times1[1:])), 4),
                          "warm_min_s": round(min(times1[1:]), 4),
                          "GTEPS": round(reach_1 / float(np.median(times1[1:])) / 1e9, 5)}
    per["bfs_all_threads"] = {"sources": len(sources), "threads": threads, "s": round(tall, 4),
                              "GTEPS": round(reach_all / tall / 1e9, 5)}
    # weak components: serial union-find, as the reference runs it
    times = []
    for rep in range(3):
        t1 = time.perf_counter()
        lab = orc.cc_weak(o)
        times.append(time.perf_counter() - t1)
    parity["cc_labels_equal"] = bool(np.array_equal(host_copy(lib, d_cc, np.uint32, n), orc.canonicalize(lab)))
    per["cc"] = {"threads": 1, "cold_s": round(times[0], 4), "warm_median_s": round(float(np.median(times[1:])), 4),
                 "GTEPS": round(e.m / float(np.median(times[1:])) / 1e9, 5)}
    # triangles: digest of the full per-node u64 counts
    tri = host_copy(lib, d_tri, np.uint64, n)
    fx_path = os.path.join(ROOT, "tests", "golden", f"tri_rmat{scale}.json")
    if os.path.exists(fx_path):
        with open(fx_path) as f:
            fx = json.load(f)
        parity["triangles_sha256_equal"] = hashlib.sha256(tri.astype("<u8").tobytes()).hexdigest() == fx["sha256_u64_le"]
        parity["triangles_total"] = int(tri.sum(dtype=np.uint64)) // 3
        parity["triangles_total_oracle"] = fx["triangles"]
        per["triangle_oracle_fixture"] = {"s": fx["oracle_seconds"], "threads": fx["oracle_threads"],
                                          "where": "build container (node-iterator count, not rerun here)"}
    else:
        parity["triangles_sha256_equal"] = None
    parity["all_equal"] = bool(parity["degree_out_equal"] and parity["degree_in_equal"]
                               and parity["bfs_rows_equal"] == len(sources) and parity["cc_labels_equal"]
                               and parity["triangles_sha256_equal"] is not False)
    log(f"c2 parity vs oracle: {json.dumps(parity)}")
    cpu = {
        "value": per["bfs_all_threads"]["GTEPS"],
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle bfs_distances (the reference's serial frontier loop) for all {len(sources)} benchmark "
                  f"sources of the step, one source per thread on {threads} threads; per-kernel CPU times "
                  f"(1 thread, cold + warm) in cpu_kernels",
        "oracle_topology_build_s": round(build_s, 3),
    }
    return cpu, per, parity


# --------------------------------------------------------------------- PageRank
def bench_bfs_large(ua, lib, topo, n, m):
    """BFS on the PageRank graph (R-MAT s24): 64 sources as one multi-source batch and
    4 of them as single-source direction-optimizing traversals; kernel time from HIP
    events, algorithmic bytes 4 E_reach + 12 n per source."""
    outdeg = ua.degree(topo, "out")
    cand = np.nonzero(outdeg > 0)[0].astype(np.uint32)
    sources = np.ascontiguousarray(cand[ua.shuffled_order(len(cand), 3)[:64]])
    lv = Dev(lib, 4 * 64 * n)
    reached = C.c_uint64(0)
    tb = time.time()
    tuning = ua.autotune(topo)  # once per graph, before timing
    tuning["tuning_s"] = round(time.time() - tb, 3)
    out = {"traversal_tuning": tuning}
    for name, k, mode, kern in (("msbfs_x64", 64, 2, ["msbfs_expand"]),
                                ("single_x4", 4, 1, ["bfs_top_down", "bfs_bottom_up"])):
        check(lib.ursa_hip_bfs_ex(topo._h, sources.ctypes.data_as(C.c_void_p), k, 0, lv.p, C.byref(reached), mode))
        ua.profile.enable(True)
        ua.profile.reset()
        t0 = time.perf_counter()
        check(lib.ursa_hip_bfs_ex(topo._h, sources.ctypes.data_as(C.c_void_p), k, 0, lv.p, C.byref(reached), mode))
        wall = time.perf_counter() - t0
        ms = sum(ua.profile.query(x)["total_ms"] for x in kern)
        ua.profile.enable(False)
        algo = 4.0 * reached.value + 12.0 * n * k
        gbs = algo / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        out[name] = {"sources": k, "E_reach": reached.value, "kernel_ms": round(ms, 3), "call_ms": round(wall * 1e3, 3),
                     "GTEPS": round(reached.value / (ms / 1e3) / 1e9, 2)}
        if k == 1 or mode == 1:
            out[name].update({"GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)})
            log(f"bfs s24 {name}: {ms:.2f} ms kernels, {gbs:.0f} GB/s ({gbs / HBM_PEAK_GBS:.1%} of HBM peak)")
        else:
            # 4 E_reach + 12 n summed over sources that share every edge scan: a
            # per-source equivalent, not bus traffic (which the counters give).
            out[name]["per_source_model_GBps"] = round(gbs, 1)
            log(f"bfs s24 {name}: {ms:.2f} ms kernels, {reached.value / (ms / 1e3) / 1e9:.0f} GTEPS")
    lv.free()
    return out


def bench_egress(ua, lib, topo, n, m, out):
    """The formats either side of the kernels at R-MAT s24: the f64 result (8 n bytes)
    copied to pageable host memory and to a pinned buffer from ursa_hip_host_alloc
    (the Arrow egress), and per-row edge weights gathered into in-CSR slot order on
    the device (algorithmic bytes 20 m: edge id 4 + weight 8 + slot write 8)."""
    res = {}
    nbytes = 8 * n
    pageable = np.empty(n, np.float64)
    pinned = ua.PinnedBuffer(nbytes)
    for name, dst in (("pageable", pageable.ctypes.data_as(C.c_void_p)), ("pinned", C.c_void_p(pinned.ptr))):
        check(lib.ursa_hip_memcpy_to_host(dst, out.p, nbytes))
        t0 = time.perf_counter()
        for _ in range(3):
            check(lib.ursa_hip_memcpy_to_host(dst, out.p, nbytes))
        ms = (time.perf_counter() - t0) / 3 * 1e3
        res[f"result_to_{name}"] = {"MB": round(nbytes / 1e6, 1), "ms": round(ms, 3), "GBps": round(nbytes / ms / 1e6, 1)}
        log(f"egress: f64 result {nbytes / 1e6:.0f} MB to {name} host memory in {ms:.2f} ms ({nbytes / ms / 1e6:.1f} GB/s)")
    pinned.free()
    w_host = np.random.default_rng(11).random(m)
    w_dev, slots = Dev(lib, 8 * m), Dev(lib, 8 * m)
    check(lib.ursa_hip_memcpy_to_device(w_dev.p, w_host.ctypes.data_as(C.c_void_p), 8 * m))
    del w_host
    check(lib.ursa_hip_gather_edge_weights(topo._h, 1, w_dev.p, slots.p))
    ua.profile.enable(True)
    ua.profile.reset()
    check(lib.ursa_hip_gather_edge_weights(topo._h, 1, w_dev.p, slots.p))
    q = ua.profile.query("gather_edge_weights")
    ua.profile.enable(False)
    w_dev.free()
    slots.free()
    gbs = q["algo_bytes"] / (q["total_ms"] / 1e3) / 1e9 if q["total_ms"] > 0 else 0.0
    res["weight_gather_in_csr"] = {"m": m, "kernel_ms": round(q["total_ms"], 3), "GBps": round(gbs, 1),
                                   "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    log(f"egress: weight gather into in-CSR order, {m} slots in {q['total_ms']:.2f} ms ({gbs:.0f} GB/s)")
    return res


def bench_c1(args, ua, gen):
    """R-MAT scale 16, seed 1, damping 0.85, tol 1e-6, max_iter 100 (the reference's
    CPU-runnable case): the f64 GPU solve through the C ABI with its inputs and output
    in HBM, and the oracle's solve on this host at one thread and at all threads
    (cold run + median/min of warm runs), with the GPU result checked against it."""
    lib = ua.load()
    e = gen.rmat(16, 16, 0.57, 0.19, 0.19, seed=1)
    topo = e.topology()
    n, m = e.n, e.m
    out = Dev(lib, 8 * n)
    it = C.c_uint32(0)
    run = lambda p: check(lib.ursa_hip_pagerank(topo._h, None, 0.85, 100, 1e-6, p, C.byref(it)))
    run(out.p)
    check(lib.ursa_hip_synchronize())
    t0 = time.perf_counter()
    for _ in range(20):
        run(out.p)
    check(lib.ursa_hip_synchronize())
    ms = (time.perf_counter() - t0) / 20 * 1e3
    res = {"graph": {"scale": 16, "seed": 1, "n": n, "m": m, "tol": 1e-6, "max_iter": 100},
           "gpu": {"ms": round(ms, 3), "iters": it.value, "GTEPS": round(m * it.value / (ms / 1e3) / 1e9, 3)}}
    if not args.no_cpu_baseline:
        import oracle as orc

        got = np.zeros(n, np.float64)
        run(got.ctypes.data_as(C.c_void_p))
        _, _, sd, dd, _ = e.export()
        o = orc.Topology(n, sd, dd)
        for th in sorted({1, cpu_threads()}):
            ts = []
            for _ in range(6):
                t1 = time.perf_counter()
                ref, rit, _ = orc.pagerank(o, 0.85, 100, 1e-6, nthreads=th)
                ts.append(time.perf_counter() - t1)
            res[f"cpu_{th}thread"] = {"cold_ms": round(ts[0] * 1e3, 3),
                                      "warm_median_ms": round(float(np.median(ts[1:])) * 1e3, 3),
                                      "warm_min_ms": round(min(ts[1:]) * 1e3, 3), "iters": rit}
        res["parity"] = {"iters_equal": it.value == rit,
                         "max_rel": float((np.abs(got - ref) / np.abs(ref)).max())}
    out.free()
    log(f"c1: {json.dumps(res)}")
    return res


def bench_pagerank(args, ua, gen):
    lib = ua.load()
    scale = args.pr_scale
    params = ua.PageRankParams(0.85, 100, 1e-10)
    t0 = time.time()
    e = gen.rmat(scale, 16, 0.57, 0.19, 0.19, seed=3)
    topo = e.topology()
    check(lib.ursa_hip_topology_incoming(topo._h))
    n, m = e.n, e.m
    log(f"pagerank graph R-MAT s{scale}: n={n} m={m} built in {time.time() - t0:.1f}s")
    out = Dev(lib, 8 * n)
    it, it32 = C.c_uint32(0), C.c_uint32(0)

    def plain():
        check(lib.ursa_hip_pagerank(topo._h, None, params.damping, params.max_iter, params.tol, out.p, C.byref(it)))

    def mixed():
        check(lib.ursa_hip_pagerank_mixed(topo._h, None, params.damping, params.max_iter, params.tol, out.p,
                                          C.byref(it), C.byref(it32)))

    def timed_solve(fn):
        fn()
        check(lib.ursa_hip_synchronize())
        t1 = time.perf_counter()
        for _ in range(args.pr_steps):
            fn()
        check(lib.ursa_hip_synchronize())
        return (time.perf_counter() - t1) / args.pr_steps * 1e3

    def sweep_roofline(qq, traffic_key):
        if not qq["launches"]:
            return None
        avg = qq["total_ms"] / qq["launches"]
        bpl = qq["algo_bytes"] / qq["launches"]
        gbs = bpl / (avg / 1e3) / 1e9
        return {"avg_launch_us": round(avg * 1e3, 2), "algo_bytes_per_launch": bpl, "GBps": round(gbs, 1),
                "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(f"pagerank_rmat{scale}", traffic_key)}

    # Two sweep schedules (ursa_hip_pagerank_schedule): binned (streamed terms,
    # exact fixed-point row sums; the automatic choice at this size) and pull.
    res = {"pull": {}}
    for sched, sname in ((2, "binned"), (1, "pull")):
        check(lib.ursa_hip_pagerank_schedule(sched))
        dst = res if sched == 2 else res["pull"]
        for name, fn in (("f64", plain), ("mixed", mixed)):
            ms = timed_solve(fn)
            dst[name] = {"ms": round(ms, 3), "iters": it.value, "GTEPS": round(m * it.value / (ms / 1e3) / 1e9, 3),
                         "schedule": sname}
            if name == "mixed":
                dst[name]["f32_iters"] = it32.value
            log(f"pagerank {sname} {name}: {ms:.2f} ms, {it.value} iterations")
        # per-sweep kernel time (the binned sweep's two phases count as one sweep)
        ua.profile.enable(True)
        ua.profile.reset()
        plain()
        q = ua.profile.query("pagerank_spmv")
        mixed()
        q32 = ua.profile.query("pagerank_spmv_f32")
        ua.profile.enable(False)
        pre = "k_bn_sweep" if sched == 2 else "k_pr_sweep"
        dst[pre + "_f64"] = sweep_roofline(q, pre + "_f64")
        dst[pre + "_f32"] = sweep_roofline(q32, pre + "_f32")
        dst["mixed_speedup"] = round(dst["f64"]["ms"] / dst["mixed"]["ms"], 3)
    check(lib.ursa_hip_pagerank_schedule(0))
    res["mixed_speedup_vs_pull_f64"] = round(res["pull"]["f64"]["ms"] / res["mixed"]["ms"], 3)
    res["binned_f64_speedup_vs_pull"] = round(res["pull"]["f64"]["ms"] / res["f64"]["ms"], 3)
    res["c1"] = bench_c1(args, ua, gen)
    res["bfs"] = bench_bfs_large(ua, lib, topo, n, m)
    res["egress"] = bench_egress(ua, lib, topo, n, m, out)
    # Weighted variant: f32-exact weights U(0,1] per edge row (pagerank.rs:135-169).
    w_dev = Dev(lib, 8 * m)
    check(lib.ursa_hip_gen_weights(m, 3, C.c_void_p(w_dev.p.value)))

    def plain_w():
        check(lib.ursa_hip_pagerank(topo._h, w_dev.p, params.damping, params.max_iter, params.tol, out.p, C.byref(it)))

    def mixed_w():
        check(lib.ursa_hip_pagerank_mixed(topo._h, w_dev.p, params.damping, params.max_iter, params.tol, out.p,
                                          C.byref(it), C.byref(it32)))

    for name, fn in (("weighted_f64", plain_w), ("weighted_mixed", mixed_w)):
        ms = timed_solve(fn)
        res[name] = {"ms": round(ms, 3), "iters": it.value, "GTEPS": round(m * it.value / (ms / 1e3) / 1e9, 3)}
        if name.endswith("mixed"):
            res[name]["f32_iters"] = it32.value
        log(f"pagerank {name}: {ms:.2f} ms, {it.value} iterations")
    check(lib.ursa_hip_pagerank_schedule(1))
    ms = timed_solve(plain_w)
    res["pull"]["weighted_f64"] = {"ms": round(ms, 3), "iters": it.value, "schedule": "pull"}
    log(f"pagerank pull weighted_f64: {ms:.2f} ms, {it.value} iterations")
    check(lib.ursa_hip_pagerank_schedule(0))
    res["weighted_mixed_speedup"] = round(res["weighted_f64"]["ms"] / res["weighted_mixed"]["ms"], 3)
    # CPU leg: the oracle's f64 solves (the reference's pull loop restated) timed on
    # this host, and the accuracy of every GPU solve against them.
    if not args.no_cpu_baseline:
        import oracle as orc

        threads = cpu_threads()
        got = {}
        for name, wp, fn, sched in (("f64", None, lib.ursa_hip_pagerank, 0),
                                    ("mixed", None, lib.ursa_hip_pagerank_mixed, 0),
                                    ("weighted_f64", w_dev.p, lib.ursa_hip_pagerank, 0),
                                    ("weighted_mixed", w_dev.p, lib.ursa_hip_pagerank_mixed, 0),
                                    ("pull_f64", None, lib.ursa_hip_pagerank, 1),
                                    ("pull_mixed", None, lib.ursa_hip_pagerank_mixed, 1)):
            check(lib.ursa_hip_pagerank_schedule(sched))
            g = np.zeros(n, np.float64)
            extra = (C.byref(it32),) if name.endswith("mixed") else ()
            check(fn(topo._h, wp, params.damping, params.max_iter, params.tol, g.ctypes.data_as(C.c_void_p),
                     C.byref(it), *extra))
            got[name] = (g, it.value)
        check(lib.ursa_hip_pagerank_schedule(0))
        w_host = np.zeros(m, np.float64)
        check(lib.ursa_hip_memcpy_to_host(w_host.ctypes.data_as(C.c_void_p), w_dev.p, 8 * m))
        t2 = time.time()
        _, _, sd, dd, _ = e.export()
        ot = orc.Topology(n, sd, dd)
        del sd, dd
        indeg = orc.degree(ot, "in")
        refs, cpu = {}, {}
        for kind, wts in (("", None), ("weighted_", w_host)):
            t3 = time.perf_counter()
            ref, rit, _ = orc.pagerank(ot, params.damping, params.max_iter, params.tol, weights=wts, nthreads=threads)
            cpu_s = time.perf_counter() - t3
            refs[kind] = (ref, rit)
            cpu[kind + "f64"] = {"s": round(cpu_s, 3), "iters": rit, "threads": threads,
                                 "GTEPS": round(m * rit / cpu_s / 1e9, 4)}
            log(f"pagerank oracle {kind or 'unweighted '}: {rit} iterations in {cpu_s:.1f}s on {threads} threads")
        log(f"pagerank oracle setup {time.time() - t2:.0f}s")
        for name, (g, git) in got.items():
            ref, rit = refs["weighted_" if name.startswith("weighted") else ""]
            rel = np.abs(g - ref) / np.abs(ref)
            r = res["pull"][name[5:]] if name.startswith("pull_") else res[name]
            r["iters_checked"] = git
            r["L1_vs_oracle"] = float(np.abs(g - ref).sum())
            r["maxabs_vs_oracle"] = float(np.abs(g - ref).max())
            r["max_rel_vs_oracle"] = float(rel.max())
            r["oracle_iters"] = rit
            w = int(np.argmax(rel))
            r["worst_node"] = {"in_degree": int(indeg[w]), "rank": float(ref[w]),
                               "nodes_above_1e-12": int((rel > 1e-12).sum())}
            top_r = set(np.argsort(-ref, kind="stable")[:1000].tolist())
            r["top1000_overlap"] = len(top_r & set(np.argsort(-g, kind="stable")[:1000].tolist()))
        res["cpu_oracle"] = cpu
        res["parity"] = {
            "f64_max_rel": res["f64"]["max_rel_vs_oracle"],
            "f64_iters_equal": got["f64"][1] == refs[""][1],
            "weighted_f64_max_rel": res["weighted_f64"]["max_rel_vs_oracle"],
            "weighted_f64_iters_equal": got["weighted_f64"][1] == refs["weighted_"][1],
            "mixed_L1": res["mixed"]["L1_vs_oracle"],
            "weighted_mixed_L1": res["weighted_mixed"]["L1_vs_oracle"],
            "pull_f64_max_rel": res["pull"]["f64"]["max_rel_vs_oracle"],
            "pull_f64_iters_equal": got["pull_f64"][1] == refs[""][1],
            "pull_mixed_L1": res["pull"]["mixed"]["L1_vs_oracle"],
            "within_bars": bool(res["f64"]["max_rel_vs_oracle"] <= 1e-12
                                and res["weighted_f64"]["max_rel_vs_oracle"] <= 1e-12
                                and res["pull"]["f64"]["max_rel_vs_oracle"] <= 1e-12
                                and res["mixed"]["L1_vs_oracle"] <= 1e-9
                                and res["pull"]["mixed"]["L1_vs_oracle"] <= 1e-9
                                and res["weighted_mixed"]["L1_vs_oracle"] <= 1e-9
                                and got["f64"][1] == refs[""][1] and got["pull_f64"][1] == refs[""][1]
                                and got["weighted_f64"][1] == refs["weighted_"][1]),
        }
        log(f"pagerank parity: {json.dumps(res['parity'])}")
        del w_host, ot
    w_dev.free()
    out.free()
    res["graph"] = {"scale": scale, "seed": 3, "n": n, "m": m, "damping": 0.85, "tol": 1e-10, "max_iter": 100}
    return res


# ------------------------------------------------------------------ batched (C4)
def bench_batched(args, ua, gen):
    """R-MAT scale 20, seed 4, symmetrised: Brandes from k=256 seeded sources and
    BFS from 4096 seeded sources, batched 64 per pass vs one traversal per launch
    sequence, on the same GPU with identical results."""
    lib = ua.load()
    scale = args.batch_scale
    e = gen.rmat(scale, 16, 0.57, 0.19, 0.19, seed=4, symmetrize=True)
    topo = e.topology()
    check(lib.ursa_hip_topology_incoming(topo._h))
    n, m = e.n, e.m
    res = {"graph": {"scale": scale, "seed": 4, "symmetrised": True, "n": n, "m": m}}
    tb = time.time()
    res["traversal_tuning"] = ua.autotune(topo)  # once per graph, before timing
    res["traversal_tuning"]["tuning_s"] = round(time.time() - tb, 3)
    # Brandes, k = 256 sources = shuffled_order(n, 4)[:256], scaled by n/256.
    src = np.ascontiguousarray(ua.shuffled_order(n, 4)[:256])
    out_b = np.zeros(n, np.float64)
    out_1 = np.zeros(n, np.float64)
    re = C.c_uint64(0)
    legs = (("batched64", 64, out_b),) + (() if args.no_compare else (("one_per_launch", 1, out_1),))
    for name, batch, out in legs:
        check(lib.ursa_hip_betweenness_sources_ex(topo._h, src.ctypes.data_as(C.c_void_p), 256, 1, n / 256.0,
                                                  batch, out.ctypes.data_as(C.c_void_p), C.byref(re)))
        t0 = time.perf_counter()
        check(lib.ursa_hip_betweenness_sources_ex(topo._h, src.ctypes.data_as(C.c_void_p), 256, 1, n / 256.0,
                                                  batch, out.ctypes.data_as(C.c_void_p), C.byref(re)))
        el = time.perf_counter() - t0
        res[f"brandes_{name}"] = {"s": round(el, 4), "sources_per_s": round(256 / el, 1),
                                  "GTEPS": round(re.value / el / 1e9, 3)}
        log(f"brandes k=256 {name}: {el * 1e3:.1f} ms")
    if not args.no_compare:
        rel = np.abs(out_b - out_1) / np.maximum(np.abs(out_1), 1e-300)
        res["brandes_speedup"] = round(res["brandes_one_per_launch"]["s"] / res["brandes_batched64"]["s"], 2)
        res["brandes_max_rel_diff_batched_vs_single"] = float(rel[out_1 != 0].max()) if np.any(out_1 != 0) else 0.0
    o = None
    if not args.no_cpu_baseline:
        # CPU leg: the oracle's Brandes over the same 256 sources (betweenness.rs:45-70,
        # 64-source chunks in parallel), timed here, and the GPU result checked against it.
        import oracle as orc

        threads = cpu_threads()
        _, _, sd, dd, _ = e.export()
        o = orc.Topology(n, sd, dd)
        del sd, dd
        t0 = time.perf_counter()
        ref = orc.betweenness_sources(o, src, scale=n / 256.0, nthreads=threads)
        cpu_s = time.perf_counter() - t0
        nzr = ref != 0
        relb = np.abs(out_b - ref)[nzr] / np.abs(ref[nzr])
        res["brandes_cpu_oracle"] = {"s": round(cpu_s, 3), "threads": threads,
                                     "sources_per_s": round(256 / cpu_s, 2)}
        res["brandes_parity"] = {"max_rel_vs_oracle": float(relb.max()) if relb.size else 0.0,
                                 "zeros_equal": bool(np.array_equal(out_b[~nzr], ref[~nzr])),
                                 "within_1e-12": bool((relb.size == 0 or relb.max() <= 1e-12)
                                                      and np.array_equal(out_b[~nzr], ref[~nzr]))}
        log(f"brandes k=256 oracle on {threads} threads: {cpu_s:.2f} s; parity {json.dumps(res['brandes_parity'])}")
    # Multi-source BFS from 4096 sources = shuffled_order(n, 4 ^ 1)[:4096].
    nsrc = args.msbfs_sources
    bsrc = np.ascontiguousarray(ua.shuffled_order(n, 4 ^ 1)[:nsrc])
    lv = Dev(lib, 4 * nsrc * n)
    for name, mode in (("batched64", 2),) + (() if args.no_compare else (("one_per_launch", 1),)):
        check(lib.ursa_hip_bfs_ex(topo._h, bsrc.ctypes.data_as(C.c_void_p), 64, 0, lv.p, C.byref(re), mode))
        t0 = time.perf_counter()
        check(lib.ursa_hip_bfs_ex(topo._h, bsrc.ctypes.data_as(C.c_void_p), nsrc, 0, lv.p, C.byref(re), mode))
        check(lib.ursa_hip_synchronize())
        el = time.perf_counter() - t0
        res[f"msbfs_{name}"] = {"s": round(el, 4), "sources_per_s": round(nsrc / el, 1),
                                "GTEPS": round(re.value / el / 1e9, 3)}
        log(f"bfs x{nsrc} {name}: {el * 1e3:.1f} ms, {re.value / el / 1e9:.1f} GTEPS")
        if name == "batched64":
            first = np.zeros((64, n), np.int32)
            check(lib.ursa_hip_memcpy_to_host(first.ctypes.data_as(C.c_void_p), lv.p, first.nbytes))
    if not args.no_compare:
        single = np.zeros((64, n), np.int32)
        check(lib.ursa_hip_memcpy_to_host(single.ctypes.data_as(C.c_void_p), lv.p, single.nbytes))
        res["msbfs_first64_identical"] = bool(np.array_equal(first, single))
    if o is not None:
        threads = cpu_threads()
        t0 = time.perf_counter()
        rows = oracle_bfs_rows(o, bsrc[:64], threads)
        cpu_s = time.perf_counter() - t0
        res["msbfs_parity"] = {"rows_checked": 64,
                               "rows_equal": int(sum(np.array_equal(first[i], rows[i]) for i in range(64)))}
        res["msbfs_cpu_oracle"] = {"sources": 64, "threads": threads, "s": round(cpu_s, 3)}
        log(f"msbfs parity vs oracle: {json.dumps(res['msbfs_parity'])}")
    if not args.no_compare:
        res["msbfs_speedup"] = round(res["msbfs_one_per_launch"]["s"] / res["msbfs_batched64"]["s"], 2)
    lv.free()
    return res


def bench_ppr(args, ua, gen):
    """Personalized PageRank on the R-MAT scale 22 / seed 2 graph (tol 1e-8, damping
    0.85): one query per solve and 16 queries per solve (one lane each), device
    output; sweep time and bytes from the library's per-kernel events."""
    lib = ua.load()
    e = gen.rmat(args.scale or 22, 16, 0.57, 0.19, 0.19, seed=2)
    topo = e.topology()
    n, m = e.n, e.m
    rng = np.random.default_rng(12)
    seed_lists = [rng.integers(0, n, int(rng.integers(1, 17))).astype(np.uint32) for _ in range(16)]
    out = Dev(lib, 8 * n * 16)
    res = {"graph": {"scale": args.scale or 22, "seed": 2, "n": n, "m": m}}
    offs = np.zeros(17, np.uint64)
    for i, sl in enumerate(seed_lists):
        offs[i + 1] = offs[i] + len(sl)
    flat = np.ascontiguousarray(np.concatenate(seed_lists))
    iters = np.zeros(16, np.uint32)
    for name, k in (("one_query", 1), ("sixteen_queries", 16)):
        call = lambda: check(lib.ursa_hip_ppr_batch(topo._h, k, offs.ctypes.data_as(C.c_void_p),
                                                    flat.ctypes.data_as(C.c_void_p), 0.85, 100, 1e-8, out.p,
                                                    iters.ctypes.data_as(C.c_void_p)))
        call()
        check(lib.ursa_hip_synchronize())
        t0 = time.perf_counter()
        call()
        check(lib.ursa_hip_synchronize())
        ms = (time.perf_counter() - t0) * 1e3
        ua.profile.enable(True)
        ua.profile.reset()
        call()
        q = ua.profile.query("ppr_rows")
        ua.profile.enable(False)
        res[name] = {"ms": round(ms, 3), "iters": iters[:k].tolist(), "sweeps": q["launches"],
                     "sweep_us": round(q["total_ms"] / max(q["launches"], 1) * 1e3, 2),
                     "GBps": round(q["algo_bytes"] / (q["total_ms"] / 1e3) / 1e9, 1) if q["total_ms"] else None}
        log(f"ppr {name}: {json.dumps(res[name])}")
    out.free()
    return res


def bench_serving(args, ua):
    """Serving mix: resident graphs (node counts log-uniform in [1e3, 2e6], edge
    factor U{4..32}, R-MAT or Erdos-Renyi 50/50, seeds 1000+i) and a query trace of
    40% PPR / 30% 3-hop / 20% shortest path / 10% triangle counts over uniformly
    chosen graphs. The plain serial loop (one query at a time through the plain
    entry points) and the continuous-batching server run the same trace: first all
    queries at once (throughput), then Poisson arrivals at a fixed rate (latency).
    Every query's result digest must match between the two."""
    from ursa_amd import serve

    lib = ua.load()
    specs = serve.serving_graph_specs(args.serve_graphs)
    t0 = time.perf_counter()
    topos, ns, ms = [], [], []
    for i, sp in enumerate(specs):
        t = serve.build_serving_graph(sp)
        topos.append(t)
        ns.append(t.n_nodes())
        ms.append(t.n_edges())
        if (i + 1) % 64 == 0:
            log(f"serving graphs built: {i + 1}/{len(specs)} ({time.perf_counter() - t0:.1f} s)")
    ser = serve.Server(serial=True)
    srv = serve.Server(slots=args.serve_slots, max_batch=16)
    t1 = time.perf_counter()
    for t in topos:
        ser.add_graph(t)
        srv.add_graph(t)
    check(lib.ursa_hip_synchronize())
    log(f"serving pool resident: {len(topos)} graphs, {sum(ns)} nodes, {sum(ms)} edges, "
        f"{srv.graph_bytes() / 2**30:.1f} GiB, caches {time.perf_counter() - t1:.1f} s")
    res = {"graphs": len(topos), "nodes": int(sum(ns)), "edges": int(sum(ms)),
           "resident_GiB": round(srv.graph_bytes() / 2**30, 2),
           "build_s": round(t1 - t0, 1), "slots": args.serve_slots, "max_batch": 16}
    warm = serve.serving_trace(ns, 64, qps=0, seed=99)
    ser.run(warm)
    srv.run(warm)

    def summary(rs, trace):
        s = serve.latency_summary(rs)
        per = {}
        for k, name in serve.KIND_NAMES.items():
            lat = [(r.done_s - r.arrival_s) * 1e3 for r, q in zip(rs, trace) if q.kind == k]
            if lat:
                per[name] = {"n": len(lat), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                             "p99_ms": round(float(np.percentile(lat, 99)), 3)}
        out = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()}
        out["by_kind"] = per
        out["mean_batch"] = round(float(np.mean([r.batch for r in rs])), 2)
        return out

    # Throughput: every query waiting at t = 0.
    sat = serve.serving_trace(ns, args.serve_queries, qps=0, seed=7)
    rs = ser.run(sat)
    rm = srv.run(sat)
    mism = sum(1 for a, b in zip(rs, rm) if a.digest != b.digest or a.status != b.status or a.status != 0)
    res["saturation"] = {"serial": summary(rs, sat), "served": summary(rm, sat), "digest_mismatches": mism}
    qs_serial = res["saturation"]["serial"]["qps"]
    qs_served = res["saturation"]["served"]["qps"]
    res["throughput_speedup"] = round(qs_served / qs_serial, 2) if qs_serial else None
    log(f"serving saturation: serial {qs_serial:.1f} q/s, served {qs_served:.1f} q/s, mismatches {mism}")
    # Latency at a fixed Poisson rate both can sustain (default 0.7 x serial capacity).
    qps = args.serve_qps or round(0.7 * qs_serial, 1)
    tr = serve.serving_trace(ns, args.serve_queries, qps=qps, seed=8)
    ls = ser.run(tr, timed_arrivals=True)
    lm = srv.run(tr, timed_arrivals=True)
    mism2 = sum(1 for a, b in zip(ls, lm) if a.digest != b.digest or a.status != b.status or a.status != 0)
    res["poisson"] = {"qps": qps, "serial": summary(ls, tr), "served": summary(lm, tr), "digest_mismatches": mism2}
    log(f"serving poisson {qps} q/s: serial p50 {res['poisson']['serial']['p50_ms']} ms p99 "
        f"{res['poisson']['serial']['p99_ms']} ms; served p50 {res['poisson']['served']['p50_ms']} ms p99 "
        f"{res['poisson']['served']['p99_ms']} ms")
    ser.close()
    srv.close()
    del topos
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=None, help="R-MAT scale of the integer-kernel graph (22)")
    ap.add_argument("--pr-scale", type=int, default=24)
    ap.add_argument("--pr-steps", type=int, default=3)
    ap.add_argument("--no-pagerank", action="store_true")
    ap.add_argument("--cpu-sources", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-scale", type=int, default=20)
    ap.add_argument("--msbfs-sources", type=int, default=4096)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--serve-graphs", type=int, default=512)
    ap.add_argument("--serve-queries", type=int, default=1000)
    ap.add_argument("--serve-qps", type=float, default=0.0, help="Poisson rate (default 0.7 x serial capacity)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--only", choices=["c2", "pagerank", "batched", "serving", "ppr"], default=None,
                    help="run one section only (for profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run the calibration copy kernel over this many MiB")
    ap.add_argument("--gather-probe", type=int, default=0, metavar="MIB",
                    help="also run the calibration gather kernel: 2^24 random 8-byte reads, each on its own "
                         "128-byte line of a buffer of this many MiB")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the one-query-per-launch comparison legs (for profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (how the path would shard over "
                                   "several GPUs is described in DESIGN.md, not implemented)"}))
        sys.exit(2)

    import ursa_amd as ua
    from ursa_amd import gen

    ua.init(0)
    if args.copy_probe:
        lib = ua.load()
        nbytes = args.copy_probe << 20
        a, b = Dev(lib, nbytes), Dev(lib, nbytes)
        for _ in range(3):
            check(lib.ursa_hip_copy_probe(a.p, b.p, nbytes))
        a.free()
        b.free()
        log(f"copy probe: 3 x {args.copy_probe} MiB")
    if args.gather_probe:
        lib = ua.load()
        lines = (args.gather_probe << 20) // 128
        count = 1 << 24
        idx = (np.random.default_rng(5).permutation(lines)[:count] * 16).astype(np.uint32)
        count = int(idx.shape[0])
        src, di, out = Dev(lib, lines * 128), Dev(lib, 4 * count), Dev(lib, 8 * count)
        check(lib.ursa_hip_memcpy_to_device(di.p, idx.ctypes.data_as(C.c_void_p), 4 * count))
        for _ in range(3):
            check(lib.ursa_hip_gather_probe(src.p, lines * 16, di.p, count, out.p))
        for b in (src, di, out):
            b.free()
        log(f"gather probe: 3 x {count} gathers over {args.gather_probe} MiB")
    if args.only == "pagerank":
        print(json.dumps({"pagerank": bench_pagerank(args, ua, gen)}), flush=True)
        return
    if args.only == "batched":
        print(json.dumps({"batched": bench_batched(args, ua, gen)}), flush=True)
        return
    if args.only == "serving":
        print(json.dumps({"serving": bench_serving(args, ua)}), flush=True)
        return
    if args.only == "ppr":
        print(json.dumps({"ppr": bench_ppr(args, ua, gen)}), flush=True)
        return
    res = bench_c2(args, ua, gen)
    if args.only == "c2":
        args.no_pagerank = True
        args.no_batched = True
        args.no_serving = True
    pr = None if args.no_pagerank else bench_pagerank(args, ua, gen)
    batched = None if args.no_batched else bench_batched(args, ua, gen)
    serving = None if args.no_serving else bench_serving(args, ua)
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
        "cpu_kernels": res["cpu_kernels"],
        "parity": res["parity"],
        "kernels": res["kernels"],
        "ingest": res["ingest"],
        "pagerank": pr,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
