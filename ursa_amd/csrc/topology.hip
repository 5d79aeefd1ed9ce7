// topology.hip — device CSR construction (Topology::build, the lazy transpose
// and the undirected view) plus degree and CSR export.
//
// Construction is a stable LSD radix sort of row indices keyed by the grouping
// endpoint (rocPRIM), which reproduces ursa-core's counting sort slot for slot
// (topology.rs:79-107: within a segment, rows stay in original order). The
// transpose rebuilds (src, dst) by original row from the out-CSR and groups
// by dst (topology.rs:307-321, :343-348). The undirected view sorts the
// packed (u, v) / (v, u) endpoint pairs and drops duplicates and self-loops
// (topology.rs:360-389).
#include "prim.hpp"
#include "topology.hpp"

namespace ursa {

namespace {

__global__ void k_iota(uint32_t* v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}

__global__ void k_gather(uint32_t* out, const uint32_t* src, const uint32_t* idx, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

// off[u] = first slot whose key >= u, for u in [0, n]; keys sorted ascending.
__global__ void k_offsets_from_sorted(const uint32_t* keys, uint64_t m, uint64_t n, uint32_t* off) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k <= m;
         k += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = (k == 0) ? 0 : (uint64_t)keys[k - 1] + 1;
        uint64_t hi = (k == m) ? n : (uint64_t)keys[k];
        for (uint64_t u = lo; u <= hi; ++u) off[u] = (uint32_t)k;
    }
}

__global__ void k_head_flags(const uint32_t* off, uint64_t n, uint32_t* flags) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t a = off[u], b = off[u + 1];
        if (a < b) flags[a] = (uint32_t)u;
    }
}

__global__ void k_reconstruct(const uint32_t* owner, const uint32_t* tgt, const uint32_t* eid, uint64_t m,
                              uint32_t* src_by_row, uint32_t* dst_by_row) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m;
         k += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t row = eid[k];
        src_by_row[row] = owner[k];
        dst_by_row[row] = tgt[k];
    }
}

// Packed (u << bits | v) and (v << bits | u) per out-slot; a self-loop becomes
// the marker n << bits, which sorts after every real pair.
__global__ void k_und_pairs(const uint32_t* owner, const uint32_t* tgt, uint64_t m, uint64_t n, int bits,
                            uint64_t* keys) {
    const uint64_t marker = n << bits;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m;
         k += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t u = owner[k], v = tgt[k];
        bool loop = (u == v);
        keys[2 * k] = loop ? marker : ((u << bits) | v);
        keys[2 * k + 1] = loop ? marker : ((v << bits) | u);
    }
}

__global__ void k_split_keys(const uint64_t* keys, uint64_t cnt, int bits, uint32_t* hi, uint32_t* lo) {
    const uint64_t mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < cnt;
         k += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t x = keys[k];
        hi[k] = (uint32_t)(x >> bits);
        lo[k] = (uint32_t)(x & mask);
    }
}

__global__ void k_degree_of(const uint32_t* off, uint64_t n, uint32_t* out) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x)
        out[u] = off[u + 1] - off[u];
}

__global__ void k_degree_both(const uint32_t* a, const uint32_t* b, uint64_t n, uint32_t* out) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x)
        out[u] = (a[u + 1] - a[u]) + (b[u + 1] - b[u]);
}

__global__ void k_widen(const uint32_t* in, uint64_t n, uint64_t* out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

__global__ void k_check_range(const uint32_t* a, const uint32_t* b, uint64_t m, uint64_t n, uint32_t* bad) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (a[i] >= n || b[i] >= n) atomicOr(bad, 1u);
}

constexpr unsigned B = 256;

// Plain streaming copy, 16 B per thread per step: the calibration kernel for
// hardware byte counters (its traffic is exactly `bytes` read + `bytes` written).
__global__ void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// out[i] = src[idx[i]]: the calibration kernel for random gathers (with every
// index on its own cache line of a buffer much larger than the caches, each
// gather costs one line fetch past L2, whatever the counters call it).
__global__ void k_gather_probe(const double* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t count,
                               double* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

}  // namespace

void build_grouped(uint64_t n, uint64_t m, const uint32_t* keys, const uint32_t* other, DevCsr& out,
                   hipStream_t s) {
    out.off.alloc(n + 1);
    out.tgt.alloc(m);
    out.eid.alloc(m);
    out.slots = m;
    if (m == 0) {
        HIP_OK(hipMemsetAsync(out.off.get(), 0, (n + 1) * sizeof(uint32_t), s));
        out.built = true;
        return;
    }
    DevBuf<uint32_t> rows(m), keys_sorted(m);
    k_iota<<<blocks_for(m, B), B, 0, s>>>(rows.get(), m);
    check_launch();
    prim::sort_pairs_u32(keys, keys_sorted.get(), rows.get(), out.eid.get(), m, bits_for(n), s);
    k_gather<<<blocks_for(m, B), B, 0, s>>>(out.tgt.get(), other, out.eid.get(), m);
    k_offsets_from_sorted<<<blocks_for(m + 1, B), B, 0, s>>>(keys_sorted.get(), m, n, out.off.get());
    check_launch();
    HIP_OK(hipStreamSynchronize(s));
    out.built = true;
}

void offsets_from_sorted(const uint32_t* keys, uint64_t m, uint64_t n, uint32_t* off, hipStream_t s) {
    k_offsets_from_sorted<<<blocks_for(m + 1, B), B, 0, s>>>(keys, m, n, off);
    check_launch();
}

void slot_owners(uint64_t n, uint64_t m, const uint32_t* off, uint32_t* owner, hipStream_t s) {
    DevBuf<uint32_t> flags(m);
    HIP_OK(hipMemsetAsync(flags.get(), 0, m * sizeof(uint32_t), s));
    k_head_flags<<<blocks_for(n, B), B, 0, s>>>(off, n, flags.get());
    check_launch();
    prim::inclusive_max_scan_u32(flags.get(), owner, m, s);
}

void slot_owners(uint64_t n, const DevCsr& csr, uint32_t* owner, hipStream_t s) {
    slot_owners(n, csr.slots, csr.off.get(), owner, s);
}

}  // namespace ursa

using namespace ursa;

const DevCsr& ursa_hip_topo::incoming(hipStream_t s) {
    NoArena keep_out_of_arena;
    std::lock_guard<std::mutex> lk(build_mu);
    if (in.built) return in;
    if (m == 0) {
        build_grouped(n, 0, nullptr, nullptr, in, s);
        return in;
    }
    DevBuf<uint32_t> owner(m), src_by_row(m), dst_by_row(m);
    slot_owners(n, out, owner.get(), s);
    k_reconstruct<<<blocks_for(m, B), B, 0, s>>>(owner.get(), out.tgt.get(), out.eid.get(), m, src_by_row.get(),
                                                 dst_by_row.get());
    check_launch();
    owner.reset();
    build_grouped(n, m, dst_by_row.get(), src_by_row.get(), in, s);
    return in;
}

const DevCsr& ursa_hip_topo::undirected(hipStream_t s) {
    NoArena keep_out_of_arena;
    std::lock_guard<std::mutex> lk(build_mu);
    if (und.built) return und;
    und.off.alloc(n + 1);
    if (m == 0) {
        HIP_OK(hipMemsetAsync(und.off.get(), 0, (n + 1) * sizeof(uint32_t), s));
        und.tgt.alloc(1);
        und.slots = 0;
        HIP_OK(hipStreamSynchronize(s));
        und.built = true;
        return und;
    }
    const int bits = bits_for(n);
    const uint64_t pairs = 2 * m;
    DevBuf<uint64_t> keys(pairs), sorted(pairs);
    {
        DevBuf<uint32_t> owner(m);
        slot_owners(n, out, owner.get(), s);
        k_und_pairs<<<blocks_for(m, B), B, 0, s>>>(owner.get(), out.tgt.get(), m, n, bits, keys.get());
        check_launch();
        HIP_OK(hipStreamSynchronize(s));
    }
    prim::sort_keys_u64(keys.get(), sorted.get(), pairs, bits + bits_for(n + 1), s);
    DevBuf<uint64_t> d_cnt(1);
    prim::unique_u64(sorted.get(), keys.get(), d_cnt.get(), pairs, s);
    uint64_t cnt = 0;
    HIP_OK(hipMemcpy(&cnt, d_cnt.get(), sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (cnt > 0) {
        uint64_t last = 0;
        HIP_OK(hipMemcpy(&last, keys.get() + (cnt - 1), sizeof(uint64_t), hipMemcpyDeviceToHost));
        if (last == (n << bits)) --cnt;
    }
    URSA_REQUIRE(cnt <= 0xFFFFFFFFull, URSA_HIP_ERANGE, "undirected slot count exceeds u32 offsets");
    sorted.reset();
    DevBuf<uint32_t> hi(cnt ? cnt : 1);
    und.tgt.alloc(cnt ? cnt : 1);
    und.slots = cnt;
    if (cnt) {
        k_split_keys<<<blocks_for(cnt, B), B, 0, s>>>(keys.get(), cnt, bits, hi.get(), und.tgt.get());
        check_launch();
    }
    k_offsets_from_sorted<<<blocks_for(cnt + 1, B), B, 0, s>>>(hi.get(), cnt, n, und.off.get());
    check_launch();
    HIP_OK(hipStreamSynchronize(s));
    und.built = true;
    return und;
}

extern "C" {

int ursa_hip_topology_create(uint64_t n, uint64_t m, const uint32_t* src_dense, const uint32_t* dst_dense,
                             ursa_hip_topo** out) {
    return guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        *out = nullptr;
        URSA_REQUIRE(m <= 0xFFFFFFFFull, URSA_HIP_ERANGE,
                     "more than u32::MAX edges is beyond the edge-id cap");
        URSA_REQUIRE(n <= 0xFFFFFFFFull, URSA_HIP_ERANGE, "more than u32::MAX nodes");
        URSA_REQUIRE(m == 0 || (src_dense && dst_dense), URSA_HIP_EINVAL, "src/dst is NULL");
        hipStream_t s = stream();
        auto* t = new ursa_hip_topo;
        try {
            t->n = n;
            t->m = m;
            DevIn<uint32_t> src(src_dense, m, s), dst(dst_dense, m, s);
            if (m) {
                DevBuf<uint32_t> bad(1);
                HIP_OK(hipMemsetAsync(bad.get(), 0, sizeof(uint32_t), s));
                k_check_range<<<blocks_for(m, B), B, 0, s>>>(src.p, dst.p, m, n, bad.get());
                check_launch();
                uint32_t hbad = 0;
                HIP_OK(hipMemcpyAsync(&hbad, bad.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                HIP_OK(hipStreamSynchronize(s));
                URSA_REQUIRE(hbad == 0, URSA_HIP_EINVAL, "dense endpoint id out of range (>= n)");
            }
            build_grouped(n, m, src.p, dst.p, t->out, s);
            t->out_deg.alloc(n ? n : 1);
            if (n) {
                k_degree_of<<<blocks_for(n, B), B, 0, s>>>(t->out.off.get(), n, t->out_deg.get());
                check_launch();
            }
            HIP_OK(hipStreamSynchronize(s));
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

}  // extern "C"

namespace ursa {
namespace {

// Offsets u64 -> u32 with monotonicity, ends and target range checked.
__global__ void k_csr_offsets(const uint64_t* __restrict__ off64, uint64_t n, uint64_t m, uint32_t* __restrict__ off,
                              uint32_t* __restrict__ bad) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = off64[i];
        if (v > m || (i == 0 && v != 0) || (i == n && v != m) || (i < n && off64[i + 1] < v)) atomicOr(bad, 1u);
        off[i] = (uint32_t)v;
    }
}

// Targets in range; edge ids counted (a permutation of 0..m-1 counts each once).
__global__ void k_csr_slots(const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ eid, uint64_t m,
                            uint64_t n, uint32_t* __restrict__ seen, uint32_t* __restrict__ bad) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
        if (tgt[k] >= n || eid[k] >= m) {
            atomicOr(bad, 2u);
            continue;
        }
        atomicAdd(&seen[eid[k]], 1u);
    }
}

__global__ void k_csr_perm(const uint32_t* __restrict__ seen, uint64_t m, uint32_t* __restrict__ bad) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x)
        if (seen[k] != 1) atomicOr(bad, 4u);
}

}  // namespace
}  // namespace ursa

extern "C" {

int ursa_hip_topology_create_from_csr(uint64_t n, uint64_t m, const uint64_t* offsets, const uint32_t* targets,
                                      const uint32_t* edge_ids, ursa_hip_topo** out) {
    return guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        *out = nullptr;
        URSA_REQUIRE(m <= 0xFFFFFFFFull, URSA_HIP_ERANGE, "more than u32::MAX edges is beyond the edge-id cap");
        URSA_REQUIRE(n <= 0xFFFFFFFFull, URSA_HIP_ERANGE, "more than u32::MAX nodes");
        URSA_REQUIRE(offsets != nullptr, URSA_HIP_EINVAL, "offsets is NULL");
        URSA_REQUIRE(m == 0 || (targets && edge_ids), URSA_HIP_EINVAL, "targets/edge_ids is NULL");
        hipStream_t s = stream();
        auto* t = new ursa_hip_topo;
        try {
            t->n = n;
            t->m = m;
            DevIn<uint64_t> off64(offsets, n + 1, s);
            DevIn<uint32_t> tg(targets, m, s), ei(edge_ids, m, s);
            t->out.off.alloc(n + 1);
            t->out.tgt.alloc(m ? m : 1);
            t->out.eid.alloc(m ? m : 1);
            DevBuf<uint32_t> bad(1), seen(m ? m : 1);
            HIP_OK(hipMemsetAsync(bad.get(), 0, sizeof(uint32_t), s));
            k_csr_offsets<<<blocks_for(n + 1, B), B, 0, s>>>(off64.p, n, m, t->out.off.get(), bad.get());
            check_launch();
            if (m) {
                HIP_OK(hipMemsetAsync(seen.get(), 0, m * sizeof(uint32_t), s));
                k_csr_slots<<<blocks_for(m, B), B, 0, s>>>(tg.p, ei.p, m, n, seen.get(), bad.get());
                k_csr_perm<<<blocks_for(m, B), B, 0, s>>>(seen.get(), m, bad.get());
                check_launch();
                HIP_OK(hipMemcpyAsync(t->out.tgt.get(), tg.p, m * 4, hipMemcpyDeviceToDevice, s));
                HIP_OK(hipMemcpyAsync(t->out.eid.get(), ei.p, m * 4, hipMemcpyDeviceToDevice, s));
            }
            uint32_t hbad = 0;
            HIP_OK(hipMemcpyAsync(&hbad, bad.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            URSA_REQUIRE(!(hbad & 1u), URSA_HIP_EINVAL, "offsets must start at 0, end at m and not decrease");
            URSA_REQUIRE(!(hbad & 2u), URSA_HIP_EINVAL, "target or edge id out of range");
            URSA_REQUIRE(!(hbad & 4u), URSA_HIP_EINVAL, "edge_ids is not a permutation of 0..m-1");
            t->out.slots = m;
            t->out.built = true;
            t->out_deg.alloc(n ? n : 1);
            if (n) {
                k_degree_of<<<blocks_for(n, B), B, 0, s>>>(t->out.off.get(), n, t->out_deg.get());
                check_launch();
            }
            HIP_OK(hipStreamSynchronize(s));
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

int ursa_hip_topology_destroy(ursa_hip_topo* t) {
    return guard([&] {
        if (!t) return;
        ensure_device();
        delete t;
    });
}

int ursa_hip_topology_size(const ursa_hip_topo* t, uint64_t* n, uint64_t* m) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        if (n) *n = t->n;
        if (m) *m = t->m;
    });
}

int ursa_hip_topology_incoming(ursa_hip_topo* t) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        t->incoming(stream());
    });
}

int ursa_hip_topology_undirected(ursa_hip_topo* t, uint64_t* slots) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        const DevCsr& u = t->undirected(stream());
        if (slots) *slots = u.slots;
    });
}

int ursa_hip_topology_export(ursa_hip_topo* t, int which, uint64_t* offsets, uint32_t* targets,
                             uint32_t* edge_ids) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(which >= 0 && which <= 2, URSA_HIP_EINVAL, "which must be 0, 1 or 2");
        URSA_REQUIRE(!(which == 2 && edge_ids), URSA_HIP_EINVAL, "the undirected view has no edge ids");
        hipStream_t s = stream();
        const DevCsr& c = which == 0 ? t->out : (which == 1 ? t->incoming(s) : t->undirected(s));
        if (offsets) {
            DevOut<uint64_t> o(offsets, t->n + 1);
            k_widen<<<blocks_for(t->n + 1, B), B, 0, s>>>(c.off.get(), t->n + 1, o.p);
            check_launch();
            o.finish(s);
        }
        if (targets && c.slots) {
            DevOut<uint32_t> o(targets, c.slots);
            HIP_OK(hipMemcpyAsync(o.p, c.tgt.get(), c.slots * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
            o.finish(s);
        }
        if (edge_ids && c.slots) {
            DevOut<uint32_t> o(edge_ids, c.slots);
            HIP_OK(hipMemcpyAsync(o.p, c.eid.get(), c.slots * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
            o.finish(s);
        }
        HIP_OK(hipStreamSynchronize(s));
    });
}

int ursa_hip_copy_probe(const void* src_device, void* dst_device, uint64_t bytes) {
    return guard([&] {
        URSA_REQUIRE(bytes % 16 == 0, URSA_HIP_EINVAL, "bytes must be a multiple of 16");
        URSA_REQUIRE(is_device_ptr(src_device) && is_device_ptr(dst_device), URSA_HIP_EINVAL,
                     "copy probe takes device buffers");
        hipStream_t s = stream();
        timed("copy_probe", 2.0 * (double)bytes, s, [&] {
            k_copy_probe<<<blocks_for(bytes / 16, 256, 1u << 16), 256, 0, s>>>((const uint4*)src_device,
                                                                              (uint4*)dst_device, bytes / 16);
        });
        check_launch();
        HIP_OK(hipStreamSynchronize(s));
    });
}

int ursa_hip_gather_probe(const double* src_device, uint64_t src_len, const uint32_t* idx_device, uint64_t count,
                          double* out_device) {
    return guard([&] {
        URSA_REQUIRE(is_device_ptr(src_device) && is_device_ptr(idx_device) && is_device_ptr(out_device),
                     URSA_HIP_EINVAL, "gather probe takes device buffers");
        URSA_REQUIRE(src_len > 0 && src_len <= 0xFFFFFFFFull, URSA_HIP_EINVAL, "src_len must be in [1, 2^32)");
        hipStream_t s = stream();
        timed("gather_probe", 12.0 * (double)count, s, [&] {
            k_gather_probe<<<blocks_for(count, 256, 1u << 16), 256, 0, s>>>(src_device, idx_device, count,
                                                                            out_device);
        });
        check_launch();
        HIP_OK(hipStreamSynchronize(s));
    });
}

// degree.rs:12-28
int ursa_hip_degree(ursa_hip_topo* t, int direction, uint32_t* out) {
    return guard([&] {
        URSA_REQUIRE(t != nullptr, URSA_HIP_EINVAL, "topology is NULL");
        URSA_REQUIRE(direction >= 0 && direction <= 2, URSA_HIP_EINVAL, "bad direction");
        if (t->n == 0) return;
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        hipStream_t s = stream();
        const uint64_t n = t->n;
        const DevCsr* in = direction == URSA_HIP_OUT ? nullptr : &t->incoming(s);
        DevOut<uint32_t> o(out, n);
        const double bytes = direction == URSA_HIP_BOTH ? 12.0 * n : 8.0 * n;
        timed("degree", bytes, s, [&] {
            if (direction == URSA_HIP_OUT)
                k_degree_of<<<blocks_for(n, B), B, 0, s>>>(t->out.off.get(), n, o.p);
            else if (direction == URSA_HIP_IN)
                k_degree_of<<<blocks_for(n, B), B, 0, s>>>(in->off.get(), n, o.p);
            else
                k_degree_both<<<blocks_for(n, B), B, 0, s>>>(t->out.off.get(), in->off.get(), n, o.p);
        });
        check_launch();
        o.finish(s);
    });
}

}  // extern "C"
