"""Host-side closed form of PageRank's first dangling mass.

The first iteration's dangling mass is 1/n added D times, left to right
(ursa-core/src/algo/pagerank.rs:72, 78). ursa_amd/csrc/repeated_sum.hpp
computes it per binade instead of per term; tests/cpp/repeated_sum_test.cpp
compares it bit for bit with the plain loop, including exact halfway cases.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_repeated_sum_matches_the_sequential_loop(tmp_path):
    exe = tmp_path / "repeated_sum_test"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "cpp", "repeated_sum_test.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("0 of 3063 differ"), r.stdout


def _fx_exact(xs):
    """Exact sum of the terms, each truncated toward zero to a multiple of
    2^-100, rounded to the nearest double (ties to even: float(Fraction))."""
    from fractions import Fraction
    import math

    tot = Fraction(0)
    for x in xs:
        f = Fraction(x) * (1 << 100)
        tot += Fraction(math.trunc(f), 1 << 100) if f >= 0 else -Fraction(math.trunc(-f), 1 << 100)
    return float(tot)


@pytest.fixture(scope="module")
def fx_tool(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("fx") / "fixed_sum_tool"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(HERE, "cpp", "fixed_sum_tool.cpp")],
                   check=True)
    return str(exe)


def _run_fx(exe, cases, *args):
    text = "".join(f"{len(c)} " + " ".join(float(x).hex() for x in c) + "\n" for c in cases)
    r = subprocess.run([exe, *args], input=text, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout.split("\n")


def test_fixed_point_sum_is_the_correctly_rounded_exact_sum(fx_tool):
    """The binned PageRank sweep's row accumulator (fixed_sum.hpp): integer parts
    added in any order, then rounded once. Checked against Fraction arithmetic on
    mixed magnitudes, signs and cancellation, terms below 2^-100 (dropped),
    subnormals, exact ties between two doubles, and sums whose low parts carry."""
    import numpy as np

    rng = np.random.default_rng(5)
    cases = [[], [0.0], [1.0], [-1.0], [3.999999], [2.0 ** -100], [2.0 ** -101], [5e-324, 1e-310],
             [1.0, 2.0 ** -53], [1.0, 2.0 ** -53, 2.0 ** -100],  # a tie, then just above it
             [1.0, 3 * 2.0 ** -53], [1.0, -2.0 ** -54], [0.1] * 10, [1e-16] * 1000 + [1.0]]
    for _ in range(300):
        k = int(rng.integers(1, 400))
        mag = 10.0 ** rng.uniform(-35, 0, k)
        sgn = np.where(rng.random(k) < 0.3, -1.0, 1.0)
        cases.append(list(mag * sgn))
    for _ in range(50):  # cancellation down to the last bits
        k = int(rng.integers(2, 100))
        a = list(rng.uniform(-1, 1, k))
        cases.append(a + [-x for x in a[:-1]])
    cases.append(list(np.full(200000, np.nextafter(2.0 ** -60, 1.0))))  # low parts near 2^40 each
    out = _run_fx(fx_tool, cases)
    for c, line in zip(cases, out):
        assert float.fromhex(line) == _fx_exact(c), (c[:4], line)


def test_fixed_point_split_parts(fx_tool):
    """Each term's parts satisfy lo + hi * 2^40 = trunc(x * 2^100), 0 <= lo < 2^40."""
    from fractions import Fraction
    import math

    xs = [1.0, -1.0, 0.5, -0.75, 2.0 ** -60, -(2.0 ** -60), 1e-13, -3.3e-20, 2.0 ** -100, -(2.0 ** -99), 0.0, -0.0]
    out = _run_fx(fx_tool, [xs], "split")
    for x, line in zip(xs, out):
        lo, hi = (int(v) for v in line.split())
        f = Fraction(x) * (1 << 100)
        want = math.trunc(f) if f >= 0 else -math.trunc(-f)
        assert 0 <= lo < (1 << 40) and lo + hi * (1 << 40) == want, (x, lo, hi)
