"""The C++ host mirror (ursa_amd/host/ursa_core.hpp) compiles and links against the
C ABI library on the CPU (no device calls are made)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include "ursa_amd/host/ursa_core.hpp"
#include <cstdio>
int main(int argc, char**) {
    if (argc > 5) {  // never taken: proves every wrapper links
        auto t = ursa_core::Topology::build(3, {0, 1}, {1, 2});
        using D = ursa_core::Direction;
        auto a = ursa_core::algo::degree(t, D::Out);
        auto b = ursa_core::algo::pagerank(t);
        auto c = ursa_core::algo::bfs_distances(t, 0, D::Both);
        auto d = ursa_core::algo::k_hop(t, {0}, 2, D::Out);
        auto e = ursa_core::algo::shortest_path(t, 0, 2, D::Out);
        auto f = ursa_core::algo::betweenness(t, std::nullopt, std::nullopt);
        auto g = ursa_core::algo::triangle_count(t);
        auto h = ursa_core::algo::connected_components_weak(t);
        auto i = ursa_core::algo::pagerank_weighted(t, {1.0, 1.0});
        std::printf("%zu\n", a.size() + b.size() + c.size() + d.first.size() + f.size() + g.size() + h.size() + i.size() + (e ? 1 : 0));
    }
    std::puts("ok");
    return 0;
}
'''


def test_cpp_mirror_compiles_and_links(tmp_path):
    src = tmp_path / "mirror.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "mirror"
    lib_dir = os.path.join(ROOT, "ursa_amd")
    cmd = ["g++", "-std=c++17", "-I", ROOT, str(src), "-o", str(exe), "-L", lib_dir, "-lursa_hip",
           "-L/opt/rocm/lib", "-Wl,-rpath," + lib_dir, "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stderr
