// cc.hip — weakly connected components (connected_components_weak,
// ursa-core/src/algo/components.rs:58-68) by sampled hook-and-compress
// (Afforest, Sutton et al. 2018).
//
//  1. comp[v] = v.
//  2. Two neighbour rounds: every vertex hooks to its r-th out-neighbour
//     (r = 0, 1), each round followed by a full compress.
//  3. The most frequent root among 1024 sampled vertices is the giant
//     component c.
//  4. Every vertex not in c hooks to its remaining out-neighbours (from slot 2)
//     and, when the transpose already exists, to all its in-neighbours; vertices
//     in c are skipped (their edges into other components are seen from the
//     other side through the in-edges). Without a transpose every remaining
//     out-edge is processed.
//  5. Final compress.
// Hooking only ever points a root at a smaller id, so comp[v] <= v always and
// each component's root is its minimum dense id: the output is the canonical
// labelling (the reference's union-by-rank root is an arbitrary member;
// relabelling either output by component minimum gives identical arrays).
//
// Hooks run lock-free with CAS on roots (the Link of Afforest / GAP cc.cc);
// loads of comp[] are agent-scope atomics so concurrent hooks are seen.
// Heavy vertices (degree > 64) are gathered into a queue and hooked one wave
// per vertex so no thread walks a hub's list alone.
//
// Algorithmic bytes: 8 B per edge processed + 8 B per node per sweep.
#include <algorithm>
#include <unordered_map>
#include <vector>

#include "frontier.hpp"

namespace ursa {

namespace {

constexpr int CB = 256;
constexpr uint32_t HEAVY = 64;

__device__ __forceinline__ uint32_t ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void hook(uint32_t u, uint32_t v, uint32_t* comp) {
    uint32_t p1 = ld(&comp[u]);
    uint32_t p2 = ld(&comp[v]);
    while (p1 != p2) {
        const uint32_t high = p1 > p2 ? p1 : p2;
        const uint32_t low = p1 + p2 - high;
        const uint32_t p_high = ld(&comp[high]);
        if (p_high == low) break;
        if (p_high == high && atomicCAS(&comp[high], high, low) == high) break;
        p1 = ld(&comp[ld(&comp[high])]);
        p2 = ld(&comp[low]);
    }
}

__global__ void k_cc_init(uint32_t* comp, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        comp[i] = (uint32_t)i;
}

// Round 0, first half: u's first neighbour v = tgt[off[u]], when smaller than
// u, becomes u's parent by a plain store (each thread writes only comp[u], and
// every comp[] is still its own id). The pointers, all to smaller ids, form a
// forest whose trees join only adjacent vertices, as hooking would: this
// replaces most of round 0's compare-and-swap hooks, which all ran at once on
// a fresh forest and retried along growing parent chains.
__global__ void k_cc_first(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt, uint64_t n,
                           uint32_t* comp) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = off[u], b = off[u + 1];
        if (a < b) {
            const uint32_t v = tgt[a];
            if (v < u) comp[u] = v;
        }
    }
}

// Round 0, second half: hooks the first neighbours larger than the vertex,
// which k_cc_first left out.
__global__ void k_cc_round0(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt, uint64_t n,
                            uint32_t* comp) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = off[u], b = off[u + 1];
        if (a < b) {
            const uint32_t v = tgt[a];
            if (v > u) hook((uint32_t)u, v, comp);
        }
    }
}

__device__ __forceinline__ uint32_t root_of(uint32_t x, const uint32_t* comp) {
    uint32_t c = ld(&comp[x]);
    while (c != x) {
        x = c;
        c = ld(&comp[x]);
    }
    return x;
}

// Round 1: the roots of u and of its second neighbour are linked, larger to
// smaller, by a plain store while the larger is still a root. After round 0
// whole trees share one root, and compare-and-swap hooks from all their
// vertices on that root's word serialised; a store that races with another
// may drop a link, never adds a wrong one (both roots are connected through
// the edge, and every pointer still goes to a smaller id). Slot 1 is swept
// again with exact hooks afterwards, so a dropped link is made there.
__global__ void k_cc_round1(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt, uint64_t n,
                            uint32_t* comp) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = off[u], b = off[u + 1];
        if (a + 1 < b) {
            const uint32_t p1 = root_of((uint32_t)u, comp), p2 = root_of(tgt[a + 1], comp);
            if (p1 != p2) {
                const uint32_t high = max(p1, p2), low = min(p1, p2);
                if (ld(&comp[high]) == high) __hip_atomic_store(&comp[high], low, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

__global__ void k_cc_compress(uint32_t* comp, uint64_t n) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t c = ld(&comp[u]);
        uint32_t cc = ld(&comp[c]);
        while (c != cc) {
            c = cc;
            cc = ld(&comp[c]);
        }
        comp[u] = c;
    }
}

// Light vertices (not in the giant component) hook their edges from slot
// `from` on; heavy ones are queued for k_cc_heavy.
__global__ void k_cc_light(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt, uint64_t n,
                           uint32_t from, uint32_t giant, uint32_t* comp, uint32_t* heavy, uint32_t* heavy_len) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = off[u] + from, b = off[u + 1];
        bool qh = false;
        if (a < b && ld(&comp[u]) != giant) {
            if (b - a > HEAVY)
                qh = true;
            else
                for (uint32_t k = a; k < b; ++k) hook((uint32_t)u, tgt[k], comp);
        }
        if (qh) heavy[atomicAdd(heavy_len, 1u)] = (uint32_t)u;
    }
}

__global__ void k_cc_heavy(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt,
                           const uint32_t* __restrict__ heavy, const uint32_t* heavy_len, uint32_t from,
                           uint32_t* comp) {
    const uint64_t gw = blockIdx.x * (uint64_t)(CB / 64) + (threadIdx.x >> 6);
    const uint64_t nw = (uint64_t)gridDim.x * (CB / 64);
    const uint32_t cnt = *heavy_len;
    for (uint64_t i = gw; i < cnt; i += nw) {
        const uint32_t u = heavy[i];
        const uint32_t a = off[u] + from, b = off[u + 1];
        for (uint32_t k = a + lane_id(); k < b; k += 64) hook(u, tgt[k], comp);
    }
}

__global__ void k_cc_sample(const uint32_t* __restrict__ comp, const uint32_t* __restrict__ idx, uint32_t k,
                            uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) out[i] = comp[idx[i]];
}

}  // namespace

void cc_weak(ursa_hip_topo* t, uint32_t* out_user, uint32_t* sweeps_out) {
    const uint64_t n = t->n;
    if (n == 0) return;
    hipStream_t s = stream();
    DevOut<uint32_t> out(out_user, n);
    uint32_t* comp = out.p;
    const DevCsr& o = t->out;
    const bool have_in = t->in.built;
    const unsigned g = blocks_for(n, CB, 1u << 16);
    const double nb = 8.0 * (double)n;
    k_cc_init<<<g, CB, 0, s>>>(comp, n);
    check_launch();
    const uint32_t rounds = 2;
    for (uint32_t r = 0; r < rounds; ++r) {
        timed("cc_hook_round", nb + 8.0 * (double)n, s, [&] {
            if (r == 0) {
                k_cc_first<<<g, CB, 0, s>>>(o.off.get(), o.tgt.get(), n, comp);
                k_cc_round0<<<g, CB, 0, s>>>(o.off.get(), o.tgt.get(), n, comp);
            } else {
                k_cc_round1<<<g, CB, 0, s>>>(o.off.get(), o.tgt.get(), n, comp);
            }
        });
        timed("cc_compress", nb, s, [&] { k_cc_compress<<<g, CB, 0, s>>>(comp, n); });
        check_launch();
    }
    // Giant component from 1024 sampled vertices (fixed sample).
    const uint32_t ks = (uint32_t)std::min<uint64_t>(1024, n);
    std::vector<uint32_t> idx(ks), samp(ks);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (uint32_t i = 0; i < ks; ++i) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        idx[i] = (uint32_t)(x % n);
    }
    DevBuf<uint32_t> didx(ks), dsamp(ks);
    HIP_OK(hipMemcpyAsync(didx.get(), idx.data(), ks * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    k_cc_sample<<<blocks_for(ks, 256), 256, 0, s>>>(comp, didx.get(), ks, dsamp.get());
    check_launch();
    HIP_OK(hipMemcpyAsync(samp.data(), dsamp.get(), ks * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    std::unordered_map<uint32_t, uint32_t> freq;
    uint32_t giant = 0, best = 0;
    for (uint32_t v : samp) {
        uint32_t f = ++freq[v];
        if (f > best || (f == best && v < giant)) {
            best = f;
            giant = v;
        }
    }
    // Without the transpose nothing may be skipped: use an id no root has.
    const uint32_t skip = have_in ? giant : 0xFFFFFFFFu;
    DevBuf<uint32_t> heavy(n), hlen(1);
    auto sweep = [&](const DevCsr& c, uint32_t from, const char* name) {
        HIP_OK(hipMemsetAsync(hlen.get(), 0, sizeof(uint32_t), s));
        timed(name, 8.0 * (double)c.slots + nb, s, [&] {
            k_cc_light<<<g, CB, 0, s>>>(c.off.get(), c.tgt.get(), n, from, skip, comp, heavy.get(), hlen.get());
            k_cc_heavy<<<blocks_for(n, CB, 1u << 14), CB, 0, s>>>(c.off.get(), c.tgt.get(), heavy.get(), hlen.get(),
                                                                  from, comp);
        });
        check_launch();
    };
    sweep(o, 1, "cc_hook_out");  // from slot 1: round 1's links may have been dropped
    if (have_in) sweep(t->in, 0, "cc_hook_in");
    timed("cc_compress", nb, s, [&] { k_cc_compress<<<g, CB, 0, s>>>(comp, n); });
    check_launch();
    out.finish(s);
    if (sweeps_out) *sweeps_out = have_in ? 2 : 1;
}

}  // namespace ursa

extern "C" int ursa_hip_cc_weak(ursa_hip_topo* t, uint32_t* labels) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (t->n == 0) return;
        URSA_REQUIRE(labels != nullptr, URSA_HIP_EINVAL, "labels is NULL");
        ursa::cc_weak(t, labels, nullptr);
    });
}
