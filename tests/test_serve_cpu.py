"""CPU checks of the serving layer: the result digest (host code in the library)
against a Python restatement of its definition, the serving-mix synthesis, and
latency bookkeeping. No device calls."""
import numpy as np

from ursa_amd import serve

M64 = (1 << 64) - 1


def py_digest(b: bytes) -> int:
    def sm(z):
        z = (z + 0x9E3779B97F4A7C15) & M64
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    h = sm(len(b))
    pad = b + bytes((-len(b)) % 8)
    for i in range(len(pad) // 8):
        w = int.from_bytes(pad[8 * i: 8 * i + 8], "little")
        h = (h + sm((w + i * 0xD1B54A32D192ED03) & M64)) & M64
    return h


def test_digest_matches_definition():
    rng = np.random.default_rng(0)
    for nb in (0, 1, 7, 8, 9, 64, 1001):
        a = rng.integers(0, 256, nb, dtype=np.uint8)
        assert serve.digest(a) == py_digest(a.tobytes()), nb


def test_digest_is_position_and_length_sensitive():
    a = np.arange(16, dtype=np.uint32)
    b = a.copy()
    b[[3, 4]] = b[[4, 3]]
    assert serve.digest(a) != serve.digest(b)
    assert serve.digest(a[:15]) != serve.digest(a)
    assert serve.digest(np.zeros(2, np.uint32)) != serve.digest(np.zeros(1, np.uint64)[:0])


def test_graph_specs_follow_the_mix():
    specs = serve.serving_graph_specs(512)
    assert [s.seed for s in specs] == list(range(1000, 1512))
    ns = np.array([s.n_target for s in specs])
    assert ns.min() >= 1000 and ns.max() <= 2_000_000
    # log-uniform: about half below the geometric mean
    assert 0.4 < np.mean(ns < np.sqrt(1e3 * 2e6)) < 0.6
    assert {s.edge_factor for s in specs} <= set(range(4, 33))
    assert 0.4 < np.mean([s.kind == "rmat" for s in specs]) < 0.6
    assert specs == serve.serving_graph_specs(512)


def test_trace_mix_and_arrivals():
    ns = [1000, 5000, 70000]
    tr = serve.serving_trace(ns, 4000, qps=200.0, seed=1)
    kinds = np.array([q.kind for q in tr])
    frac = [np.mean(kinds == k) for k in range(4)]
    assert abs(frac[0] - 0.4) < 0.03 and abs(frac[1] - 0.3) < 0.03
    assert abs(frac[2] - 0.2) < 0.03 and abs(frac[3] - 0.1) < 0.03
    arr = np.array([q.arrival_s for q in tr])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    assert abs(arr[-1] / (len(tr) - 1) - 1 / 200.0) < 0.1 / 200.0
    for q in tr:
        assert q.graph < 3
        seeds = list(q.seeds)[: q.nseeds]
        assert all(s < ns[q.graph] for s in seeds)
        if q.kind == serve.PPR:
            assert 1 <= q.nseeds <= 16 and q.damping == 0.85 and q.tol == 1e-8
        if q.kind == serve.KHOP:
            assert q.k == 3 and q.nseeds == 1
        if q.kind == serve.PATH:
            assert q.nseeds == 2


def test_latency_summary():
    rs = []
    for i in range(100):
        r = serve.Result()
        r.arrival_s = i * 0.01
        r.done_s = r.arrival_s + (0.001 if i < 99 else 0.5)
        rs.append(r)
    s = serve.latency_summary(rs)
    assert abs(s["p50_ms"] - 1.0) < 1e-6 and s["p99_ms"] > 1.0 and s["failed"] == 0
