// Checks ursa::repeated_sum (the closed form of a sequential f64 sum of equal
// terms) against the plain loop, bit for bit; built and run by
// tests/test_host_sums.py.
#include "../../ursa_amd/csrc/repeated_sum.hpp"
#include <cmath>
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(7);
    int bad = 0, n = 0;
    double xs[] = {1.0 / 8870160, 1.0 / 3, 0.1, 1.0 / 1024, 3.0, 1e-300, 5e-324, 1.0 / 65536 * 3, 0.75};
    for (double x : xs) for (uint64_t c : {0ull, 1ull, 2ull, 3ull, 1000ull, 1489159ull, 4000000ull}) {
        double a = 0.0; for (uint64_t i = 0; i < c; ++i) a += x;
        double b = ursa::repeated_sum(x, c); ++n;
        if (a != b) { ++bad; printf("x=%a c=%llu loop=%a fast=%a\n", x, (unsigned long long)c, a, b); }
    }
    for (int t = 0; t < 3000; ++t) {
        double x = std::ldexp((double)(g() >> 11) / 9007199254740992.0 + 0.5, -(int)(g() % 40));
        if (t % 3 == 0) x = std::ldexp((double)(g() % 4096 + 1), -(int)(g() % 30));  // few mantissa bits: halfway cases
        uint64_t c = g() % 200000;
        double a = 0.0; for (uint64_t i = 0; i < c; ++i) a += x;
        double b = ursa::repeated_sum(x, c); ++n;
        if (a != b) { ++bad; if (bad < 10) printf("x=%a c=%llu loop=%a fast=%a\n", x, (unsigned long long)c, a, b); }
    }
    printf("%d of %d differ\n", bad, n);
    return bad != 0;
}
