// gen.hip — seeded synthetic graphs generated on the device: R-MAT and
// Erdős–Rényi edge lists, deduplicated keeping the first occurrence, optionally
// symmetrised, then interned to dense ids in first-seen order (src then dst
// per row, the numbering of ursa-core/src/id_map.rs:220-292).
//
// Streams are splitmix64 in counter form: the t-th output (1-based) of the
// stream seeded with `seed` is mix(seed + t * 0x9E3779B97F4A7C15), so every
// draw is computed independently. R-MAT edge e takes draws e*scale+1 ..
// e*scale+scale, most significant bit first; each draw's top 53 bits form a
// uniform double r that picks the quadrant by the thresholds a, a+b, a+b+c.
// Erdős–Rényi edge e takes draws 2e+1 (src) and 2e+2 (dst), each mod n.
#include "../../include/ursa_hip_ext.h"
#include "prim.hpp"
#include "topology.hpp"

struct ursa_hip_edges {
    uint64_t n = 0, m = 0;
    ursa::DevBuf<uint32_t> usrc, udst;  // kept edges (user ids), row order
    ursa::DevBuf<uint32_t> src, dst;    // dense ids
    ursa::DevBuf<int64_t> users;        // dense -> user id
};

namespace ursa {
namespace {

constexpr unsigned B = 256;
constexpr uint64_t GAMMA = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_gen_rmat(uint64_t draws, uint32_t scale, double a, double ab, double abc, uint64_t seed,
                           uint32_t* us, uint32_t* vs) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < draws;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t u = 0, v = 0;
        for (uint32_t j = 0; j < scale; ++j) {
            const uint64_t x = mix64(seed + (e * scale + j + 1) * GAMMA);
            const double r = (double)(x >> 11) * 0x1.0p-53;
            uint32_t bu = 0, bv = 0;
            if (r < a) {
            } else if (r < ab) {
                bv = 1;
            } else if (r < abc) {
                bu = 1;
            } else {
                bu = 1;
                bv = 1;
            }
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        us[e] = u;
        vs[e] = v;
    }
}

__global__ void k_gen_er(uint64_t draws, uint64_t n, uint64_t seed, uint32_t* us, uint32_t* vs) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < draws;
         e += (uint64_t)gridDim.x * blockDim.x) {
        us[e] = (uint32_t)(mix64(seed + (2 * e + 1) * GAMMA) % n);
        vs[e] = (uint32_t)(mix64(seed + (2 * e + 2) * GAMMA) % n);
    }
}

__global__ void k_pair_keys(const uint32_t* us, const uint32_t* vs, uint64_t m, int bits, uint64_t* keys,
                            uint32_t* idx) {
    const uint64_t marker = 1ull << (2 * bits);
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t u = us[e], v = vs[e];
        keys[e] = (u == v) ? marker : ((u << bits) | v);
        idx[e] = (uint32_t)e;
    }
}

// keep[row] = 1 for the first row of each distinct key (sorted stably, so the
// first in a run has the smallest row), never for self-loop markers.
__global__ void k_first_flags(const uint64_t* keys, const uint32_t* idx, uint64_t m, int bits, uint8_t* keep) {
    const uint64_t marker = 1ull << (2 * bits);
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[k];
        const bool first = (k == 0 || keys[k - 1] != key) && key != marker;
        keep[idx[k]] = first ? 1 : 0;
    }
}

__global__ void k_first_seen(const uint32_t* us, const uint32_t* vs, uint64_t m, uint32_t* fp) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < m;
         r += (uint64_t)gridDim.x * blockDim.x) {
        atomicMin(&fp[us[r]], (uint32_t)(2 * r));
        atomicMin(&fp[vs[r]], (uint32_t)(2 * r + 1));
    }
}

__global__ void k_present(const uint32_t* fp, uint64_t range, uint8_t* flag, uint32_t* iota) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < range;
         i += (uint64_t)gridDim.x * blockDim.x) {
        flag[i] = fp[i] != 0xFFFFFFFFu;
        iota[i] = (uint32_t)i;
    }
}

__global__ void k_gather_u32(const uint32_t* src, const uint32_t* idx, uint64_t n, uint32_t* out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

__global__ void k_assign_dense(const uint32_t* ids_by_rank, uint64_t n, uint32_t* dense_of, int64_t* users) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t id = ids_by_rank[r];
        dense_of[id] = (uint32_t)r;
        users[r] = (int64_t)id;
    }
}

__global__ void k_map_dense(const uint32_t* us, const uint32_t* vs, uint64_t m, const uint32_t* dense_of,
                            uint32_t* s, uint32_t* d) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < m;
         r += (uint64_t)gridDim.x * blockDim.x) {
        s[r] = dense_of[us[r]];
        d[r] = dense_of[vs[r]];
    }
}

__global__ void k_concat_rev(const uint32_t* us, const uint32_t* vs, uint64_t m, uint32_t* u2, uint32_t* v2) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < m;
         r += (uint64_t)gridDim.x * blockDim.x) {
        u2[r] = us[r];
        v2[r] = vs[r];
        u2[m + r] = vs[r];
        v2[m + r] = us[r];
    }
}

__global__ void k_widen_i64(const uint32_t* in, uint64_t n, int64_t* out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (int64_t)in[i];
}

__global__ void k_weights(uint64_t m, uint64_t seed, double* out) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = mix64(seed + (e + 1) * GAMMA);
        out[e] = (double)((x >> 40) + 1) * 0x1.0p-24;
    }
}

// Drop self-loops and later duplicates of (us, vs) in place of the buffers;
// returns the kept count. Row order is preserved.
uint64_t dedup_first(DevBuf<uint32_t>& us, DevBuf<uint32_t>& vs, uint64_t m, int bits, hipStream_t s) {
    if (m == 0) return 0;
    DevBuf<uint64_t> keys(m), keys_sorted(m);
    DevBuf<uint32_t> idx(m), idx_sorted(m);
    k_pair_keys<<<blocks_for(m, B), B, 0, s>>>(us.get(), vs.get(), m, bits, keys.get(), idx.get());
    check_launch();
    prim::sort_pairs_u64_u32(keys.get(), keys_sorted.get(), idx.get(), idx_sorted.get(), m, 2 * bits + 1, s);
    keys.reset();
    idx.reset();
    DevBuf<uint8_t> keep(m);
    k_first_flags<<<blocks_for(m, B), B, 0, s>>>(keys_sorted.get(), idx_sorted.get(), m, bits, keep.get());
    check_launch();
    keys_sorted.reset();
    idx_sorted.reset();
    DevBuf<uint32_t> u2(m), v2(m);
    DevBuf<uint64_t> cnt(1);
    prim::select_flagged_u32(us.get(), keep.get(), u2.get(), cnt.get(), m, s);
    prim::select_flagged_u32(vs.get(), keep.get(), v2.get(), cnt.get(), m, s);
    uint64_t h = 0;
    HIP_OK(hipMemcpyAsync(&h, cnt.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    us = std::move(u2);
    vs = std::move(v2);
    return h;
}

void finish(ursa_hip_edges* g, DevBuf<uint32_t>& us, DevBuf<uint32_t>& vs, uint64_t draws, uint64_t range,
            int symmetrize, hipStream_t s) {
    const int bits = bits_for(range);
    URSA_REQUIRE(bits <= 31, URSA_HIP_ERANGE, "id range too large for the generator");
    uint64_t m = dedup_first(us, vs, draws, bits, s);
    if (symmetrize && m) {
        DevBuf<uint32_t> u2(2 * m), v2(2 * m);
        k_concat_rev<<<blocks_for(m, B), B, 0, s>>>(us.get(), vs.get(), m, u2.get(), v2.get());
        check_launch();
        m = dedup_first(u2, v2, 2 * m, bits, s);
        us = std::move(u2);
        vs = std::move(v2);
    }
    URSA_REQUIRE(2 * m < 0xFFFFFFFFull, URSA_HIP_ERANGE, "too many edges for first-seen interning");
    g->m = m;
    // First-seen interning: fp[id] = smallest endpoint position 2*row(+1).
    DevBuf<uint32_t> fp(range);
    HIP_OK(hipMemsetAsync(fp.get(), 0xFF, range * sizeof(uint32_t), s));
    if (m) {
        k_first_seen<<<blocks_for(m, B), B, 0, s>>>(us.get(), vs.get(), m, fp.get());
        check_launch();
    }
    DevBuf<uint8_t> flag(range);
    DevBuf<uint32_t> iota(range), present(range);
    DevBuf<uint64_t> cnt(1);
    k_present<<<blocks_for(range, B), B, 0, s>>>(fp.get(), range, flag.get(), iota.get());
    check_launch();
    prim::select_flagged_u32(iota.get(), flag.get(), present.get(), cnt.get(), range, s);
    uint64_t n = 0;
    HIP_OK(hipMemcpyAsync(&n, cnt.get(), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    g->n = n;
    g->users.alloc(n ? n : 1);
    g->src.alloc(m ? m : 1);
    g->dst.alloc(m ? m : 1);
    if (n) {
        DevBuf<uint32_t> key(n), key_sorted(n), ids_sorted(n);
        k_gather_u32<<<blocks_for(n, B), B, 0, s>>>(fp.get(), present.get(), n, key.get());
        check_launch();
        prim::sort_pairs_u32(key.get(), key_sorted.get(), present.get(), ids_sorted.get(), n, 32, s);
        DevBuf<uint32_t> dense_of(range);
        k_assign_dense<<<blocks_for(n, B), B, 0, s>>>(ids_sorted.get(), n, dense_of.get(), g->users.get());
        check_launch();
        if (m) {
            k_map_dense<<<blocks_for(m, B), B, 0, s>>>(us.get(), vs.get(), m, dense_of.get(), g->src.get(),
                                                       g->dst.get());
            check_launch();
        }
        HIP_OK(hipStreamSynchronize(s));
    }
    g->usrc = std::move(us);
    g->udst = std::move(vs);
}

}  // namespace
}  // namespace ursa

using namespace ursa;

extern "C" {

int ursa_hip_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                      int symmetrize, ursa_hip_edges** out) {
    return guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        *out = nullptr;
        URSA_REQUIRE(scale >= 1 && scale <= 30, URSA_HIP_EINVAL, "scale must be in [1, 30]");
        const uint64_t draws = (uint64_t)edge_factor << scale;
        URSA_REQUIRE(draws < (1ull << 32), URSA_HIP_ERANGE, "too many edge draws");
        hipStream_t s = stream();
        auto* g = new ursa_hip_edges;
        try {
            DevBuf<uint32_t> us(draws ? draws : 1), vs(draws ? draws : 1);
            const double ab = a + b, abc = a + b + c;
            if (draws) {
                k_gen_rmat<<<blocks_for(draws, B), B, 0, s>>>(draws, scale, a, ab, abc, seed, us.get(), vs.get());
                check_launch();
            }
            finish(g, us, vs, draws, 1ull << scale, symmetrize, s);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int ursa_hip_gen_er(uint64_t n, uint32_t edge_factor, uint64_t seed, int symmetrize, ursa_hip_edges** out) {
    return guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        *out = nullptr;
        URSA_REQUIRE(n >= 1 && n < (1ull << 31), URSA_HIP_EINVAL, "n must be in [1, 2^31)");
        const uint64_t draws = (uint64_t)edge_factor * n;
        URSA_REQUIRE(draws < (1ull << 32), URSA_HIP_ERANGE, "too many edge draws");
        hipStream_t s = stream();
        auto* g = new ursa_hip_edges;
        try {
            DevBuf<uint32_t> us(draws ? draws : 1), vs(draws ? draws : 1);
            if (draws) {
                k_gen_er<<<blocks_for(draws, B), B, 0, s>>>(draws, n, seed, us.get(), vs.get());
                check_launch();
            }
            finish(g, us, vs, draws, n, symmetrize, s);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int ursa_hip_edges_size(const ursa_hip_edges* g, uint64_t* n, uint64_t* m) {
    return guard([&] {
        URSA_REQUIRE(g != nullptr, URSA_HIP_EINVAL, "edges handle is NULL");
        if (n) *n = g->n;
        if (m) *m = g->m;
    });
}

int ursa_hip_edges_export(ursa_hip_edges* g, int64_t* user_src, int64_t* user_dst, uint32_t* src_dense,
                          uint32_t* dst_dense, int64_t* user_ids) {
    return guard([&] {
        URSA_REQUIRE(g != nullptr, URSA_HIP_EINVAL, "edges handle is NULL");
        hipStream_t s = stream();
        const uint64_t m = g->m, n = g->n;
        auto widen = [&](int64_t* user, const DevBuf<uint32_t>& from) {
            if (!user || !m) return;
            DevOut<int64_t> o(user, m);
            k_widen_i64<<<blocks_for(m, B), B, 0, s>>>(from.get(), m, o.p);
            check_launch();
            o.finish(s);
        };
        widen(user_src, g->usrc);
        widen(user_dst, g->udst);
        auto copy = [&](uint32_t* user, const DevBuf<uint32_t>& from) {
            if (!user || !m) return;
            DevOut<uint32_t> o(user, m);
            HIP_OK(hipMemcpyAsync(o.p, from.get(), m * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
            o.finish(s);
        };
        copy(src_dense, g->src);
        copy(dst_dense, g->dst);
        if (user_ids && n) {
            DevOut<int64_t> o(user_ids, n);
            HIP_OK(hipMemcpyAsync(o.p, g->users.get(), n * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
            o.finish(s);
        }
    });
}

int ursa_hip_edges_dense_ptrs(ursa_hip_edges* g, const uint32_t** src_dense, const uint32_t** dst_dense) {
    return guard([&] {
        URSA_REQUIRE(g != nullptr, URSA_HIP_EINVAL, "edges handle is NULL");
        if (src_dense) *src_dense = g->src.get();
        if (dst_dense) *dst_dense = g->dst.get();
    });
}

int ursa_hip_edges_topology(ursa_hip_edges* g, ursa_hip_topo** out) {
    if (!g) {
        set_last_error("edges handle is NULL");
        return URSA_HIP_EINVAL;
    }
    return ursa_hip_topology_create(g->n, g->m, g->src.get(), g->dst.get(), out);
}

int ursa_hip_edges_destroy(ursa_hip_edges* g) {
    return guard([&] {
        if (!g) return;
        ensure_device();
        delete g;
    });
}

int ursa_hip_gen_weights(uint64_t m, uint64_t seed, double* out) {
    return guard([&] {
        if (m == 0) return;
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        hipStream_t s = stream();
        DevOut<double> o(out, m);
        k_weights<<<blocks_for(m, B), B, 0, s>>>(m, seed ^ 0x5745494748540000ull, o.p);
        check_launch();
        o.finish(s);
    });
}

}  // extern "C"
