// closeness_w.hip — weighted closeness (closeness.rs:143-222): shortest path
// costs by edge weight from each source over out-edges, then
// reachable / (sum of the finite costs of the other nodes, in node order).
//
// 64 sources per batch, vertex-major (the cost of vertex v from source b at
// v*64 + b). Costs are relaxed label-correcting style: each round takes the
// vertices whose cost improved for some source and relaxes their out-edges
// for those sources with an atomic minimum. Non-negative f64 values order like
// their bit patterns, so the minimum is a u64 atomicMin. The result equals
// Dijkstra's bit for bit: both produce, for every vertex, the minimum over
// paths of the path cost added left to right (dist[u] + w, rounded), because
// rounded addition of a non-negative weight is monotone in dist[u] and never
// decreases it. The final sum runs over v = 0..n in order on one lane per
// source, the reference's summation order.
#include <cmath>
#include <vector>

#include "../../include/ursa_hip.h"
#include "frontier.hpp"
#include "topology.hpp"

namespace ursa {

void gather_edge_weights(ursa_hip_topo* t, int which, const double* w_rows, double* w_slots);

namespace {

constexpr int WB = 256;
constexpr int WW = WB / 64;
constexpr unsigned long long INF_BITS = 0x7FF0000000000000ull;

__global__ void k_wc_init(unsigned long long* __restrict__ dist, unsigned long long* __restrict__ cur, uint64_t n,
                          const uint32_t* __restrict__ src, uint32_t nb) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n * 64; i += (uint64_t)gridDim.x * blockDim.x) {
        dist[i] = INF_BITS;
        if ((i & 63) == 0) cur[i >> 6] = 0;
    }
}

__global__ void k_wc_seed(unsigned long long* __restrict__ dist, unsigned long long* __restrict__ cur,
                          const uint32_t* __restrict__ src, uint32_t nb) {
    for (uint32_t b = 0; b < nb; ++b) {
        dist[(uint64_t)src[b] * 64 + b] = 0;  // +0.0
        cur[src[b]] |= 1ull << b;
    }
}

// Queue of the vertices improved in the last round; their masks move to `mask`
// and `cur` is cleared for the next round.
__global__ __launch_bounds__(WB) void k_wc_compact(unsigned long long* __restrict__ cur, uint64_t n,
                                                   uint32_t* __restrict__ q,
                                                   unsigned long long* __restrict__ mask, uint32_t* len) {
    __shared__ uint32_t s_cnt[WW + 1];
    for (uint64_t b0 = blockIdx.x * (uint64_t)WB; b0 < n; b0 += (uint64_t)gridDim.x * WB) {
        const uint64_t v = b0 + threadIdx.x;
        unsigned long long m = 0;
        if (v < n) {
            m = cur[v];
            if (m) cur[v] = 0;
        }
        const uint32_t slot = block_append_slot<WB>(m != 0, len, s_cnt);
        if (m) {
            q[slot] = (uint32_t)v;
            mask[slot] = m;
        }
    }
}

// One wave per queued vertex u, one lane per source: relax u's out-edges for
// the sources whose cost at u improved.
__global__ __launch_bounds__(WB) void k_wc_relax(const uint32_t* __restrict__ q,
                                                 const unsigned long long* __restrict__ mask, uint32_t qlen,
                                                 const uint32_t* __restrict__ off, const uint32_t* __restrict__ tgt,
                                                 const double* __restrict__ wsl,
                                                 unsigned long long* __restrict__ dist,
                                                 unsigned long long* __restrict__ nxt) {
    const int b = lane_id();
    const uint64_t gw = blockIdx.x * (uint64_t)WW + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * WW;
    for (uint64_t i = gw; i < qlen; i += nwaves) {
        const uint32_t u = q[i];
        const bool act = (mask[i] >> b) & 1ull;
        const double du = __longlong_as_double((long long)dist[(uint64_t)u * 64 + b]);
        const uint32_t ee = off[u + 1];
        for (uint32_t e = off[u]; e < ee; e += 4) {  // four edges' atomics in flight at once
            uint32_t xs[4];
            bool better[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                xs[j] = e + j < ee ? tgt[e + j] : 0u;
                better[j] = false;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (act && e + j < ee) {
                    const double nd = du + wsl[e + j];  // rounded as the reference's `du + wt`
                    const unsigned long long nb = (unsigned long long)__double_as_longlong(nd);
                    better[j] = nb < atomicMin(&dist[(uint64_t)xs[j] * 64 + b], nb);
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t bm = __ballot(better[j]);
                if (bm && b == __ffsll((unsigned long long)bm) - 1) atomicOr(&nxt[xs[j]], (unsigned long long)bm);
            }
        }
    }
}

// Lane b: node-order sum of source b's finite costs (closeness.rs:199-215).
__global__ void k_wc_fold(const unsigned long long* __restrict__ dist, uint64_t n, const uint32_t* __restrict__ src,
                          uint32_t nb, double* __restrict__ out) {
    const uint32_t b = threadIdx.x;
    if (b >= nb) return;
    const uint32_t u = src[b];
    double sum = 0.0;
    long long reach = 0;
    for (uint64_t v = 0; v < n; ++v) {
        const double d = __longlong_as_double((long long)dist[v * 64 + b]);
        if (v != u && d < INFINITY) {
            sum = __dadd_rn(sum, d);
            ++reach;
        }
    }
    out[b] = sum > 0.0 ? __ddiv_rn((double)reach, sum) : 0.0;
}

}  // namespace

void closeness_weighted_sources(ursa_hip_topo* t, const double* w_rows, const std::vector<uint32_t>& src,
                                double* out_user) {
    const uint64_t n = t->n, m = t->m, k = src.size();
    std::vector<double> res(k, 0.0);
    std::vector<uint32_t> valid;
    std::vector<uint64_t> where;
    for (uint64_t i = 0; i < k; ++i)
        if (src[i] < n) {
            valid.push_back(src[i]);
            where.push_back(i);
        }
    if (!valid.empty()) {
        hipStream_t s = stream();
        DevBuf<double> wsl(m ? m : 1);
        if (m) gather_edge_weights(t, 0, w_rows, wsl.get());
        DevBuf<unsigned long long> dist(n * 64), cur(n), nxt(n), mask(n);
        DevBuf<uint32_t> q(n), dsrc(64), len(1);
        DevBuf<double> dout(64);
        const unsigned g = blocks_for(n, WB, GRID_CAP);
        for (uint64_t b0 = 0; b0 < valid.size(); b0 += 64) {
            const uint32_t nb = (uint32_t)std::min<uint64_t>(64, valid.size() - b0);
            HIP_OK(hipMemcpyAsync(dsrc.get(), valid.data() + b0, nb * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            k_wc_init<<<blocks_for(n * 64, 256, 1u << 14), 256, 0, s>>>(dist.get(), cur.get(), n, dsrc.get(), nb);
            HIP_OK(hipMemsetAsync(nxt.get(), 0, n * sizeof(unsigned long long), s));
            k_wc_seed<<<1, 1, 0, s>>>(dist.get(), cur.get(), dsrc.get(), nb);
            check_launch();
            unsigned long long* c = cur.get();
            unsigned long long* x = nxt.get();
            for (;;) {
                HIP_OK(hipMemsetAsync(len.get(), 0, sizeof(uint32_t), s));
                k_wc_compact<<<g, WB, 0, s>>>(c, n, q.get(), mask.get(), len.get());
                check_launch();
                uint32_t qlen = 0;
                HIP_OK(hipMemcpyAsync(&qlen, len.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                HIP_OK(hipStreamSynchronize(s));
                if (qlen == 0) break;
                timed("closeness_weighted_relax", 0.0, s, [&] {
                    k_wc_relax<<<blocks_for(qlen, WW, GRID_CAP), WB, 0, s>>>(q.get(), mask.get(), qlen,
                                                                              t->out.off.get(), t->out.tgt.get(),
                                                                              wsl.get(), dist.get(), x);
                });
                check_launch();
                std::swap(c, x);  // x (the old cur) was cleared by the compaction
            }
            k_wc_fold<<<1, 64, 0, s>>>(dist.get(), n, dsrc.get(), nb, dout.get());
            check_launch();
            double h[64];
            HIP_OK(hipMemcpyAsync(h, dout.get(), nb * sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            for (uint32_t b = 0; b < nb; ++b) res[where[b0 + b]] = h[b];
        }
    }
    if (k) {
        if (is_device_ptr(out_user))
            HIP_OK(hipMemcpy(out_user, res.data(), k * sizeof(double), hipMemcpyHostToDevice));
        else
            std::copy(res.begin(), res.end(), out_user);
    }
}

}  // namespace ursa

extern "C" int ursa_hip_closeness_weighted(ursa_hip_topo* t, const double* weights, const uint32_t* sources,
                                           uint64_t k, double* out) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(t->m == 0 || weights != nullptr, URSA_HIP_EINVAL, "weights is NULL");
        std::vector<uint32_t> src;
        if (sources == nullptr) {
            src.resize(t->n);
            for (uint64_t i = 0; i < t->n; ++i) src[i] = (uint32_t)i;
        } else {
            src.resize(k);
            if (k && ursa::is_device_ptr(sources))
                HIP_OK(hipMemcpy(src.data(), sources, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
            else if (k)
                std::copy(sources, sources + k, src.begin());
        }
        URSA_REQUIRE(src.empty() || out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        ursa::closeness_weighted_sources(t, weights, src, out);
    });
}
