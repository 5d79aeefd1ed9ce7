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
#include <mutex>
#include <unordered_map>
#include <vector>

#include "frontier.hpp"

namespace ursa {

namespace {

constexpr int CB = 256;
constexpr uint32_t HEAVY = 64;
constexpr uint32_t CC_SAMPLES = 1024;
constexpr uint64_t CC_SAMPLE_SEED = 0x9E3779B97F4A7C15ull;

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
// The skipped component is *giant_p when giant_p is set, else giant_v.
__global__ void k_cc_light(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt, uint64_t n,
                           uint32_t from, const uint32_t* giant_p, uint32_t giant_v, uint32_t* comp, uint32_t* heavy,
                           uint32_t* heavy_len) {
    const uint32_t giant = giant_p ? *giant_p : giant_v;
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

// The giant-component vote on the device (one workgroup of CC_SAMPLES threads):
// sample i is the root of vertex x_{i+1} % n of the xorshift sequence the host
// path draws, and the winner is the most frequent root, the smallest on ties —
// the host path's choice. Also clears the heavy-queue counter.
__global__ __launch_bounds__(CC_SAMPLES) void k_cc_giant(const uint32_t* __restrict__ comp, uint64_t n,
                                                        uint32_t* __restrict__ giant, uint32_t* __restrict__ hlen) {
    __shared__ uint32_t samp[CC_SAMPLES];
    __shared__ unsigned long long best[CC_SAMPLES / 64];
    const uint32_t i = threadIdx.x;
    const uint32_t ks = n < CC_SAMPLES ? (uint32_t)n : CC_SAMPLES;
    if (i < ks) {
        uint64_t x = CC_SAMPLE_SEED;
        for (uint32_t j = 0; j <= i; ++j) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
        }
        samp[i] = __hip_atomic_load(&comp[x % n], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    unsigned long long key = 0;  // (count, ~root): the maximum is the most frequent, smallest root
    if (i < ks) {
        const uint32_t v = samp[i];
        uint32_t cnt = 0;
        for (uint32_t j = 0; j < ks; ++j) cnt += samp[j] == v;
        key = (unsigned long long)cnt << 32 | (uint32_t)~v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(key, off, 64);
        key = o > key ? o : key;
    }
    if ((i & 63) == 0) best[i >> 6] = key;
    __syncthreads();
    if (i == 0) {
        unsigned long long b = 0;
        for (uint32_t w = 0; w < CC_SAMPLES / 64; ++w) b = best[w] > b ? best[w] : b;
        *giant = ~(uint32_t)b;
        *hlen = 0;
    }
}

__global__ void k_cc_zero(uint32_t* p) { *p = 0; }

// Scratch of ursa_hip_cc_weak_noalloc: the heavy-vertex queue, its counter and
// the giant root, for graphs of at most CC_NOALLOC_MAX_N vertices. Each device
// holds its own copy (16 MiB + 1 KiB) for the life of the process.
constexpr uint64_t CC_NOALLOC_MAX_N = 1ull << 22;
constexpr uint64_t CC_SCRATCH_HLEN = CC_NOALLOC_MAX_N, CC_SCRATCH_GIANT = CC_NOALLOC_MAX_N + 1;
constexpr uint64_t CC_SCRATCH_WORDS = CC_NOALLOC_MAX_N + 256;
static __device__ __align__(256) uint32_t scratch[CC_SCRATCH_WORDS];

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
            k_cc_light<<<g, CB, 0, s>>>(c.off.get(), c.tgt.get(), n, from, nullptr, skip, comp, heavy.get(),
                                        hlen.get());
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

// Device address of `scratch` per device, taken once per device with that
// device current (ursa_hip_init, else the first no-allocation call there). The
// array's name is a host-side stand-in; only these addresses reach kernels.
namespace {
std::mutex g_scratch_mu;
uint32_t* g_scratch[64] = {};
}  // namespace

void cc_noalloc_setup() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    URSA_REQUIRE(dev >= 0 && dev < 64, URSA_HIP_EDEVICE, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return;
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    URSA_REQUIRE(p != nullptr, URSA_HIP_EDEVICE, "no device address for the components scratch");
    g_scratch[dev] = static_cast<uint32_t*>(p);
}

// cc_weak without any allocation: the caller's device labels are the parent
// array, the heavy queue and counters live in this device's `scratch`, the
// giant component is voted on the device, and nothing is copied to the host or
// waited for: the work is queued on the calling thread's stream.
void cc_weak_noalloc(ursa_hip_topo* t, uint32_t* comp) {
    const uint64_t n = t->n;
    if (n == 0) return;
    URSA_REQUIRE(n <= CC_NOALLOC_MAX_N, URSA_HIP_ERANGE,
                 "ursa_hip_cc_weak_noalloc takes at most 4194304 vertices (its scratch size)");
    URSA_REQUIRE(is_device_ptr(comp), URSA_HIP_EINVAL, "labels must be device memory for the no-allocation entry point");
    int dev = ensure_device();
    HIP_OK(hipGetDevice(&dev));
    uint32_t* sp = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        sp = (dev >= 0 && dev < 64) ? g_scratch[dev] : nullptr;
    }
    if (!sp) {
        cc_noalloc_setup();
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        sp = g_scratch[dev];
    }
    uint32_t* heavy = sp;
    uint32_t* hlen = sp + CC_SCRATCH_HLEN;
    uint32_t* giant = sp + CC_SCRATCH_GIANT;
    hipStream_t s = stream();
    const DevCsr& o = t->out;
    const bool have_in = t->in.built;
    const unsigned g = blocks_for(n, CB, 1u << 16);
    k_cc_init<<<g, CB, 0, s>>>(comp, n);
    k_cc_first<<<g, CB, 0, s>>>(o.off.get(), o.tgt.get(), n, comp);
    k_cc_round0<<<g, CB, 0, s>>>(o.off.get(), o.tgt.get(), n, comp);
    k_cc_compress<<<g, CB, 0, s>>>(comp, n);
    k_cc_round1<<<g, CB, 0, s>>>(o.off.get(), o.tgt.get(), n, comp);
    k_cc_compress<<<g, CB, 0, s>>>(comp, n);
    k_cc_giant<<<1, CC_SAMPLES, 0, s>>>(comp, n, giant, hlen);
    check_launch();
    // Without the transpose nothing may be skipped: an id no root has.
    k_cc_light<<<g, CB, 0, s>>>(o.off.get(), o.tgt.get(), n, 1, have_in ? giant : nullptr, 0xFFFFFFFFu, comp, heavy,
                                hlen);
    k_cc_heavy<<<blocks_for(n, CB, 1u << 14), CB, 0, s>>>(o.off.get(), o.tgt.get(), heavy, hlen, 1, comp);
    check_launch();
    if (have_in) {
        k_cc_zero<<<1, 1, 0, s>>>(hlen);
        k_cc_light<<<g, CB, 0, s>>>(t->in.off.get(), t->in.tgt.get(), n, 0, giant, 0u, comp, heavy, hlen);
        k_cc_heavy<<<blocks_for(n, CB, 1u << 14), CB, 0, s>>>(t->in.off.get(), t->in.tgt.get(), heavy, hlen, 0, comp);
        check_launch();
    }
    k_cc_compress<<<g, CB, 0, s>>>(comp, n);
    check_launch();
}

}  // namespace ursa

extern "C" int ursa_hip_cc_weak_noalloc(ursa_hip_topo* t, uint32_t* labels) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (t->n == 0) return;
        URSA_REQUIRE(labels != nullptr, URSA_HIP_EINVAL, "labels is NULL");
        ursa::cc_weak_noalloc(t, labels);
    });
}

extern "C" int ursa_hip_cc_weak(ursa_hip_topo* t, uint32_t* labels) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (t->n == 0) return;
        URSA_REQUIRE(labels != nullptr, URSA_HIP_EINVAL, "labels is NULL");
        ursa::cc_weak(t, labels, nullptr);
    });
}
