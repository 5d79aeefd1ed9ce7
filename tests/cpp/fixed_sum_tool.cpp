// Reads lines "k x_1 ... x_k" (hexadecimal doubles) and prints, per line, the
// fixed-point exact sum of ursa_amd/csrc/fixed_sum.hpp (fx_split of every term,
// the parts added as integers, fx_value), in hexadecimal; with the argument
// "split", prints each term's (lo, hi) parts instead. Built and driven by
// tests/test_host_sums.py, which checks it against exact rational arithmetic.
#include "../../ursa_amd/csrc/fixed_sum.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
int main(int argc, char** argv) {
    const bool split = argc > 1 && std::strcmp(argv[1], "split") == 0;
    unsigned long long k;
    while (std::scanf("%llu", &k) == 1) {
        unsigned long long lo_sum = 0, lo_rev = 0;
        long long hi_sum = 0, hi_rev = 0;
        double* xs = static_cast<double*>(std::malloc((k ? k : 1) * sizeof(double)));
        for (unsigned long long i = 0; i < k; ++i) {
            char buf[64];
            if (std::scanf("%63s", buf) != 1) return 2;
            xs[i] = std::strtod(buf, nullptr);
        }
        for (unsigned long long i = 0; i < k; ++i) {
            unsigned long long lo;
            long long hi;
            ursa::fx_split(xs[i], lo, hi);
            if (split) std::printf("%llu %lld\n", lo, hi);
            lo_sum += lo;
            hi_sum += hi;
        }
        for (unsigned long long i = k; i-- > 0;) {  // the other order: the same integers
            unsigned long long lo;
            long long hi;
            ursa::fx_split(xs[i], lo, hi);
            lo_rev += lo;
            hi_rev += hi;
        }
        std::free(xs);
        if (lo_sum != lo_rev || hi_sum != hi_rev) return 3;
        if (!split) std::printf("%a\n", ursa::fx_value(lo_sum, hi_sum));
    }
    return 0;
}
