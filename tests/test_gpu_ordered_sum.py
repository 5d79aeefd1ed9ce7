"""GPU parity of the ordered f64 sum behind PageRank's dangling mass.

The reference sums the dangling ranks left to right over the ascending node
list (ursa-core/src/algo/pagerank.rs:76-79). The device reproduces that
loop's rounding in parallel (binade-wise integer sums of rounded quotients,
with single steps where the running sum changes binade or meets a halfway
case), so the bar is bit-exact against a plain Python loop, on inputs chosen
to hit every single-step case: many binade changes, exact halfway cases,
zeros, subnormals, negative values, infinities and NaN.
"""
import math

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ua():
    import ursa_amd

    ursa_amd.init(0)
    return ursa_amd


def seq_sum(x):
    s = 0.0
    for v in x.tolist():
        s += v
    return s


def same(a, b):
    return (math.isnan(a) and math.isnan(b)) or np.float64(a).tobytes() == np.float64(b).tobytes()


def cases():
    rng = np.random.default_rng(11)
    n = 1_300_000
    yield "uniform_small", rng.random(n) * 1e-7
    yield "identical_1_over_n", np.full(3_000_017, 1.0 / 3_000_017)
    yield "powers_of_two", 2.0 ** -rng.integers(0, 60, 400_000).astype(np.float64)
    yield "few_mantissa_bits", rng.integers(1, 1024, 500_000) / 1024.0
    yield "pareto", (rng.pareto(1.2, 700_000) + 1.0) * 1e-6
    pick = rng.integers(0, 6, 300_000)
    vals = np.array([0.0, 5e-324, 1e-300, 0.5, 2.0 ** -30, 3.0])
    yield "mixed_special", vals[pick] * (1.0 + (pick == 3) * rng.random(300_000))
    yield "with_negatives", rng.random(200_000) - 0.3
    yield "ascending", np.sort(rng.random(300_000))
    yield "descending", np.sort(rng.random(300_000))[::-1].copy()
    x = rng.random(100_000)
    x[77_777] = np.inf
    yield "with_inf", x
    x = rng.random(100_000)
    x[5] = np.nan
    yield "with_nan", x
    for k in (1, 2, 63, 64, 65, 4095, 4096, 4097, 262_144, 262_145):
        yield f"n{k}", rng.random(k) * 1e-3


@pytest.mark.parametrize("case", list(cases()), ids=lambda c: c[0])
def test_ordered_sum_is_bit_exact(ua, case):
    name, x = case
    got = ua.ordered_sum(x)
    want = seq_sum(x)
    assert same(got, want), (name, got, want, float(np.sum(x)))


def test_ordered_sum_empty(ua):
    assert ua.ordered_sum(np.zeros(0)) == 0.0


def test_pagerank_dangling_mass_in_reference_order(ua):
    """A graph whose ranks are set by the dangling mass: 2^21 nodes, 90 % of them
    dangling, in-degree at most 1, so every row sum is exact and the only rounding
    that depends on summation order is the dangling sum's. Summed pairwise instead
    of in order, the ranks differ from the reference's by up to 3.3e-11 relative on
    this graph; summed in order they are bit-identical."""
    n = 1 << 21
    rng = np.random.default_rng(5)
    src = rng.choice(n, n // 10, replace=False).astype(np.uint32)
    dst = ((src.astype(np.uint64) * 7 + 1) % n).astype(np.uint32)
    t = ua.Topology.build(n, src, dst)
    o = orc.Topology(n, src, dst)
    got, it = ua._pagerank(t, None, ua.PageRankParams(0.85, 100, 1e-12))
    ref, rit, _ = orc.pagerank(o, 0.85, 100, 1e-12)
    assert it == rit
    assert np.array_equal(got, ref), float(np.max(np.abs(got - ref) / np.abs(ref)))
