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
