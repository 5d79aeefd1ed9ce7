// rng.hpp — the reference's seeded source sampler (ursa-core/src/algo/rng.rs:
// 9-73, betweenness.rs:349-362): splitmix64, a Fisher-Yates permutation from
// i = n-1 down with j = next() % (i + 1), and ceil(frac * n) sources clamped to
// [1, n]. Host code; the same seed selects the same sources as ursa.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace ursa {

constexpr uint64_t kDefaultSeed = 0x5EED5EED5EED5EEDull;

struct SplitMix64 {
    uint64_t state;
    uint64_t next() {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

inline std::vector<uint32_t> shuffled_order(uint64_t n, uint64_t seed) {
    std::vector<uint32_t> order(n);
    for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
    SplitMix64 r{seed ^ 0xD1B54A32D192ED03ull};
    for (uint64_t i = n; i-- > 1;) std::swap(order[i], order[r.next() % (i + 1)]);
    return order;
}

// None (has_frac false) -> every node in order.
inline std::vector<uint32_t> sample_sources(uint64_t n, bool has_frac, double frac, bool has_seed, uint64_t seed) {
    if (!has_frac) {
        std::vector<uint32_t> all(n);
        for (uint64_t i = 0; i < n; ++i) all[i] = (uint32_t)i;
        return all;
    }
    const double f = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
    const double kf = std::ceil((double)n * f);
    uint64_t k = (kf != kf || kf <= 0.0) ? 0 : (kf >= 1.8e19 ? UINT64_MAX : (uint64_t)kf);
    k = std::min<uint64_t>(std::max<uint64_t>(k, 1), n);
    std::vector<uint32_t> order = shuffled_order(n, has_seed ? seed : kDefaultSeed);
    order.resize(k);
    return order;
}

}  // namespace ursa
