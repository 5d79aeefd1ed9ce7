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
