// egress.hip — the data formats on either side of the kernels:
//   * edge weights in slot order: evaluate_weight yields one f64 per edge row
//     (ursa-plan/src/weight.rs:51-106) and a weighted kernel reads
//     weights[edge_ids[k]] per adjacency slot (weight.rs:8-10,
//     ursa-core/src/algo/pagerank.rs:159-161); the gather runs here on the
//     device, with evaluate_weight's non-negativity check;
//   * result buffers in pinned host memory: algo_array moves each kernel's Vec
//     into an Arrow array without a copy (ursa-plan/src/result.rs:84-139); a
//     buffer from ursa_hip_host_alloc can back that array, and a device result
//     lands in it by DMA at full link speed instead of through the runtime's
//     bounce buffer for pageable memory.
#include "../../include/ursa_hip_ext.h"
#include "topology.hpp"

namespace ursa {
namespace {

// w_slots[k] = w_rows[eid[k]]; any negative weight raises *neg (the gather
// touches every row once, since eid is a permutation of the rows).
__global__ void k_gather_weights(const uint32_t* __restrict__ eid, const double* __restrict__ w, uint64_t m,
                                 double* __restrict__ out, uint32_t* __restrict__ neg) {
    bool bad = false;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
        const double x = w[eid[k]];
        out[k] = x;
        bad |= x < 0.0;
    }
    if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(neg, 1u);
}

}  // namespace

void gather_edge_weights(ursa_hip_topo* t, int which, const double* w_rows, double* w_slots) {
    hipStream_t s = stream();
    const uint64_t m = t->m;
    if (m == 0) return;
    const DevCsr* csr = nullptr;
    if (which == 0)
        csr = &t->out;
    else if (which == 1)
        csr = &t->incoming(s);
    else
        throw Error(URSA_HIP_EINVAL, "which must be 0 (out-CSR) or 1 (in-CSR)");
    DevIn<double> w(w_rows, m, s);
    DevOut<double> out(w_slots, m);
    DevBuf<uint32_t> neg(1);
    HIP_OK(hipMemsetAsync(neg.get(), 0, sizeof(uint32_t), s));
    timed("gather_edge_weights", 20.0 * (double)m, s, [&] {
        k_gather_weights<<<blocks_for(m, 256, 1u << 14), 256, 0, s>>>(csr->eid.get(), w.p, m, out.p, neg.get());
    });
    check_launch();
    uint32_t h = 0;
    HIP_OK(hipMemcpyAsync(&h, neg.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    out.finish(s);
    URSA_REQUIRE(h == 0, URSA_HIP_EINVAL,
                 "weight expression produced a negative value; weights must be non-negative");
}

}  // namespace ursa

extern "C" {

int ursa_hip_gather_edge_weights(ursa_hip_topo* t, int which, const double* w_rows, double* w_slots) {
    return ursa::guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(t->m == 0 || (w_rows && w_slots), URSA_HIP_EINVAL, "NULL weight buffer");
        ursa::gather_edge_weights(t, which, w_rows, w_slots);
    });
}

int ursa_hip_host_alloc(uint64_t bytes, void** out) {
    return ursa::guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        *out = nullptr;
        ursa::ensure_device();
        if (bytes == 0) return;
        void* p = nullptr;
        const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw ursa::Error(URSA_HIP_ENOMEM, "hipHostMalloc of pinned result memory failed");
        }
        *out = p;
    });
}

int ursa_hip_host_free(void* p) {
    return ursa::guard([&] {
        if (p) HIP_OK(hipHostFree(p));
    });
}

}  // extern "C"
