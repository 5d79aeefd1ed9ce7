// repeated_sum.hpp — the value of ((0 + x) + x) + ... + x (count terms) as a
// sequential f64 loop rounds it, in O(binades) steps on the host.
//
// PageRank's first dangling mass is such a sum: every rank starts at 1/n and the
// reference adds the dangling ones left to right (pagerank.rs:72, 78). While the
// running sum s stays in one binade [2^e, 2^(e+1)), with u = 2^(e-52) and
// significand S = s/u, each step adds round(x/u) to S — a constant — except in
// the halfway case, where round-half-even makes S even after the first such
// step and the increment constant after that. So each binade is a division
// plus one ordinary add at the binade change.
#pragma once
#include <cstdint>
#include <cstring>

namespace ursa {

inline double repeated_sum(double x, uint64_t count) {
    if (count == 0) return 0.0;
    double s = 0.0 + x;
    uint64_t i = 1;
    const uint64_t lim = 1ull << 53;
    const uint64_t mant_mask = (1ull << 52) - 1;
    while (i < count) {
        uint64_t sb, xb;
        std::memcpy(&sb, &s, 8);
        std::memcpy(&xb, &x, 8);
        const int se = (int)(sb >> 52), xe = (int)(xb >> 52);  // sign bits land above 0x7FF
        // Outside positive normal running sums and summands, step plainly.
        if (se < 23 || se >= 0x7FF || xe < 1 || xe >= 0x7FF) {
            s = s + x;
            ++i;
            continue;
        }
        uint64_t sig = (sb & mant_mask) | (1ull << 52);
        const uint64_t mx = (xb & mant_mask) | (1ull << 52);
        const int k = se - xe;  // x / u = mx / 2^k
        if (k > 63) return s;  // x < u / 2^10: no step moves the sum any more
        if (k < 1) {
            s = s + x;
            ++i;
            continue;
        }
        const uint64_t q = mx >> k, rem = mx & ((1ull << k) - 1), half = 1ull << (k - 1);
        uint64_t add;
        if (rem < half) {
            add = q;
        } else if (rem > half) {
            add = q + 1;
        } else {
            if (sig & 1u) {  // halfway from an odd significand: take one step, which leaves it even
                s = s + x;
                ++i;
                continue;
            }
            add = q + (q & 1u);  // halfway from an even significand: the result is even again
        }
        if (add == 0) return s;  // the sum no longer moves
        uint64_t steps = (lim - 1 - sig) / add;
        if (steps > count - i) steps = count - i;
        sig += steps * add;
        i += steps;
        sb = ((uint64_t)se << 52) | (sig & mant_mask);
        std::memcpy(&s, &sb, 8);
        if (i < count) {  // the step that leaves the binade
            s = s + x;
            ++i;
        }
    }
    return s;
}

}  // namespace ursa
