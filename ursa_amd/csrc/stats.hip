// stats.hip — BFS-based graph statistics of ursa-plan/src/stats.rs:100-168:
// bfs_source_scan (the (sum, count, max) fold over finite distances >= 1 from
// each source), avg_path_length and diameter.
//
// Sources run 64 at a time through the multi-source BFS; each batch's levels
// are folded on the device. Distances are integers, so the sum is accumulated
// exactly in u64 and converted once: the reference's f64 fold of integer
// distances is exact below 2^53, so both give the same bits.
#include <functional>
#include <vector>

#include "frontier.hpp"

namespace ursa {

TuneParams effective_tune(ursa_hip_topo* t, int dir);
void msbfs_levels_each(ursa_hip_topo* t, const std::vector<uint32_t>& src, int dir, int32_t* lv64, uint32_t push_div,
                       const std::function<void(uint64_t, uint32_t)>& each);

namespace {

struct ScanAcc {
    unsigned long long sum;
    unsigned long long count;
    int max;
    int pad;
};

__global__ void k_scan_fold(const int32_t* __restrict__ lv, uint64_t cnt, ScanAcc* acc) {
    unsigned long long sum = 0, c = 0;
    int mx = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) {
        const int d = lv[i];
        if (d >= 1) {
            sum += (unsigned long long)d;
            ++c;
            mx = max(mx, d);
        }
    }
    sum = wave_sum_u64(sum);
    c = wave_sum_u64(c);
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_down(mx, off, 64));
    if (lane_id() == 0) {
        if (sum) atomicAdd(&acc->sum, sum);
        if (c) atomicAdd(&acc->count, c);
        if (mx) atomicMax(&acc->max, mx);
    }
}

// Per lane b of a batch: sum and count of the finite distances >= 1 in
// lv[b*n .. (b+1)*n) (blockIdx.y = lane).
__global__ void k_lane_fold(const int32_t* __restrict__ lv, uint64_t n, unsigned long long* __restrict__ acc) {
    const uint64_t b = blockIdx.y;
    const int32_t* row = lv + b * n;
    unsigned long long sum = 0, c = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int d = row[i];
        if (d >= 1) {
            sum += (unsigned long long)d;
            ++c;
        }
    }
    sum = wave_sum_u64(sum);
    c = wave_sum_u64(c);
    if (lane_id() == 0) {
        if (sum) atomicAdd(&acc[2 * b], sum);
        if (c) atomicAdd(&acc[2 * b + 1], c);
    }
}

// triangle.rs-derived local clustering (clustering.rs:21-40): 2*T / (k*(k-1)),
// 0 for undirected degree < 2; same operation order as the reference.
__global__ void k_clustering(const uint32_t* __restrict__ uoff, const unsigned long long* __restrict__ tri, uint64_t n,
                             double* __restrict__ out) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        const double k = (double)(uoff[u + 1] - uoff[u]);
        out[u] = k < 2.0 ? 0.0 : __ddiv_rn(__dmul_rn(2.0, (double)tri[u]), __dmul_rn(k, __dsub_rn(k, 1.0)));
    }
}

// stats.rs:80-98: None -> every node; Some(frac) -> round(frac*n) (at least
// one, at most n) evenly strided ids, floor(i * max(n/k, 1)), capped at n-1.
std::vector<uint32_t> stats_sources(uint64_t n, bool has_frac, double frac) {
    std::vector<uint32_t> out;
    if (!has_frac) {
        out.resize(n);
        for (uint64_t i = 0; i < n; ++i) out[i] = (uint32_t)i;
        return out;
    }
    const double f = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
    double kk = std::round((double)n * f);
    if (kk < 1.0) kk = 1.0;
    uint64_t k = std::min<uint64_t>((uint64_t)kk, n);
    if (k == 0) return out;
    const double step = std::max((double)n / (double)k, 1.0);
    for (uint64_t i = 0; i < k; ++i)
        out.push_back((uint32_t)std::min<uint64_t>((uint64_t)std::floor((double)i * step), n - 1));
    return out;
}

}  // namespace

void bfs_source_scan(ursa_hip_topo* t, const std::vector<uint32_t>& sources_in, int dir, double* sum, uint64_t* count,
                     int32_t* mx) {
    const uint64_t n = t->n;
    *sum = 0.0;
    *count = 0;
    *mx = 0;
    std::vector<uint32_t> src;
    for (uint32_t v : sources_in)
        if (v < n) src.push_back(v);  // bfs.rs:21-23: an unknown source reaches nothing
    if (src.empty()) return;
    hipStream_t s = stream();
    DevBuf<int32_t> lv(64 * n);
    DevBuf<ScanAcc> acc(1);
    HIP_OK(hipMemsetAsync(acc.get(), 0, sizeof(ScanAcc), s));
    const TuneParams tp = effective_tune(t, dir);
    msbfs_levels_each(t, src, dir, lv.get(), tp.ms_push_div, [&](uint64_t, uint32_t nb) {
        k_scan_fold<<<blocks_for(nb * n, 256, 4096), 256, 0, s>>>(lv.get(), (uint64_t)nb * n, acc.get());
        check_launch();
    });
    ScanAcc h{};
    HIP_OK(hipMemcpyAsync(&h, acc.get(), sizeof(ScanAcc), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    *sum = (double)h.sum;
    *count = h.count;
    *mx = h.max;
}

void triangle_count(ursa_hip_topo* t, uint64_t* out_user);

// closeness.rs:69-137 (unweighted): for each source u, reachable / sum of the
// finite distances >= 1, or 0 when it reaches nothing. Both are exact integers,
// so one f64 division reproduces the reference's value.
void closeness_sources(ursa_hip_topo* t, const std::vector<uint32_t>& src, double* out_user) {
    const uint64_t n = t->n, k = src.size();
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
        DevBuf<int32_t> lv(64 * n);
        DevBuf<unsigned long long> acc(128);
        std::vector<unsigned long long> h(128);
        const TuneParams tp = effective_tune(t, URSA_HIP_OUT);
        msbfs_levels_each(t, valid, URSA_HIP_OUT, lv.get(), tp.ms_push_div, [&](uint64_t b0, uint32_t nb) {
            HIP_OK(hipMemsetAsync(acc.get(), 0, 128 * sizeof(unsigned long long), s));
            dim3 grid(blocks_for(n, 256, 256), nb);
            k_lane_fold<<<grid, 256, 0, s>>>(lv.get(), n, acc.get());
            check_launch();
            HIP_OK(hipMemcpyAsync(h.data(), acc.get(), 128 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            for (uint32_t b = 0; b < nb; ++b) {
                const unsigned long long sum = h[2 * b], cnt = h[2 * b + 1];
                res[where[b0 + b]] = sum > 0 ? (double)cnt / (double)sum : 0.0;
            }
        });
    }
    if (k) {
        if (is_device_ptr(out_user))
            HIP_OK(hipMemcpy(out_user, res.data(), k * sizeof(double), hipMemcpyHostToDevice));
        else
            std::copy(res.begin(), res.end(), out_user);
    }
}

void clustering(ursa_hip_topo* t, double* out_user) {
    const uint64_t n = t->n;
    if (n == 0) return;
    hipStream_t s = stream();
    DevBuf<uint64_t> tri(n);
    triangle_count(t, tri.get());
    const DevCsr& U = t->undirected(s);
    DevOut<double> out(out_user, n);
    k_clustering<<<blocks_for(n, 256, 4096), 256, 0, s>>>(U.off.get(), reinterpret_cast<unsigned long long*>(tri.get()),
                                                          n, out.p);
    check_launch();
    out.finish(s);
}

}  // namespace ursa

using namespace ursa;

extern "C" {

int ursa_hip_bfs_source_scan(ursa_hip_topo* t, const uint32_t* sources, uint64_t k, int direction, double* sum,
                             uint64_t* count, int32_t* max_dist) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr && sum && count && max_dist, URSA_HIP_EINVAL, "NULL argument");
        URSA_REQUIRE(direction >= 0 && direction <= 2, URSA_HIP_EINVAL, "bad direction");
        URSA_REQUIRE(k == 0 || sources != nullptr, URSA_HIP_EINVAL, "sources is NULL");
        std::vector<uint32_t> src(k);
        if (k && is_device_ptr(sources))
            HIP_OK(hipMemcpy(src.data(), sources, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
        else if (k)
            std::copy(sources, sources + k, src.begin());
        bfs_source_scan(t, src, direction, sum, count, max_dist);
    });
}

int ursa_hip_avg_path_length(ursa_hip_topo* t, int has_sample, double sample, double* out) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr && out != nullptr, URSA_HIP_EINVAL, "NULL argument");
        double sum = 0.0;
        uint64_t count = 0;
        int32_t mx = 0;
        bfs_source_scan(t, stats_sources(t->n, has_sample != 0, sample), URSA_HIP_OUT, &sum, &count, &mx);
        *out = count == 0 ? 0.0 : sum / (double)count;  // stats.rs:150-151
    });
}

int ursa_hip_closeness(ursa_hip_topo* t, const uint32_t* sources, uint64_t k, double* out) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        std::vector<uint32_t> src;
        if (sources == nullptr) {
            src.resize(t->n);
            for (uint64_t i = 0; i < t->n; ++i) src[i] = (uint32_t)i;
        } else {
            src.resize(k);
            if (k && is_device_ptr(sources))
                HIP_OK(hipMemcpy(src.data(), sources, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
            else if (k)
                std::copy(sources, sources + k, src.begin());
        }
        URSA_REQUIRE(src.empty() || out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        closeness_sources(t, src, out);
    });
}

int ursa_hip_clustering_coefficient(ursa_hip_topo* t, double* out) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (t->n == 0) return;
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        clustering(t, out);
    });
}

int ursa_hip_diameter(ursa_hip_topo* t, int approximate, int64_t* out) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr && out != nullptr, URSA_HIP_EINVAL, "NULL argument");
        const uint64_t n = t->n;
        std::vector<uint32_t> src;
        if (approximate && n > 0)  // stats.rs:161-163: a bounded strided sample of ~128 sources
            src = stats_sources(n, true, std::min(128.0 / (double)n, 1.0));
        else
            src = stats_sources(n, false, 0.0);
        double sum = 0.0;
        uint64_t count = 0;
        int32_t mx = 0;
        bfs_source_scan(t, src, URSA_HIP_OUT, &sum, &count, &mx);
        *out = mx;
    });
}

}  // extern "C"
