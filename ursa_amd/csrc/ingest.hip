// ingest.hip — IdMap interning of int64 edge endpoints on the device
// (IdMap::from_edge_batches / intern_i64, ursa-core/src/id_map.rs:220-319).
//
// Dense ids are assigned in first-seen order over the edge rows, src before
// dst within a row (id_map.rs:251-254). Endpoint occurrence p = 2*row + side
// (side 0 = src, 1 = dst) is the scan position. On the device:
//   1. stable radix sort of (user id, p) over all 2m occurrences: equal ids
//      form runs with their positions ascending, so a run's first element is
//      the id's first sighting;
//   2. each run's first position, sorted, gives the dense order: the run whose
//      first sighting comes k-th receives dense id k;
//   3. every occurrence is written its run's dense id at its position, and
//      user_ids[dense] = id.
// The result is the reference's numbering exactly (no hash order involved).
#include "../../include/ursa_hip_ext.h"
#include "prim.hpp"
#include "topology.hpp"

namespace ursa {
namespace {

constexpr unsigned IB = 256;

// Occurrences of one batch of m rows whose first row is global row row0.
__global__ void k_occurrences(const int64_t* __restrict__ src, const int64_t* __restrict__ dst, uint64_t m,
                              uint64_t row0, uint64_t* __restrict__ keys, uint32_t* __restrict__ pos) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < m; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t g = row0 + r;
        keys[2 * g] = (uint64_t)src[r];
        keys[2 * g + 1] = (uint64_t)dst[r];
        pos[2 * g] = (uint32_t)(2 * g);
        pos[2 * g + 1] = (uint32_t)(2 * g + 1);
    }
}

// Run heads of the sorted ids: flag, and the head's first position.
__global__ void k_run_heads(const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ spos, uint64_t cnt,
                            uint32_t* __restrict__ head, uint32_t* __restrict__ first_pos) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) {
        const bool h = i == 0 || skeys[i] != skeys[i - 1];
        head[i] = h ? 1u : 0u;
        first_pos[i] = h ? spos[i] : 0xFFFFFFFFu;
    }
}

// run_of[i] = index of i's run (inclusive prefix of heads, minus one).
__global__ void k_minus_one(uint32_t* __restrict__ x, uint64_t cnt) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x)
        x[i] -= 1;
}

// For each run r: its first position and its id, compacted in run order.
__global__ void k_run_table(const uint32_t* __restrict__ head, const uint32_t* __restrict__ run_of,
                            const uint32_t* __restrict__ spos, const uint64_t* __restrict__ skeys, uint64_t cnt,
                            uint32_t* __restrict__ run_first, uint64_t* __restrict__ run_key,
                            uint32_t* __restrict__ run_iota) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) {
        if (!head[i]) continue;
        const uint32_t r = run_of[i];
        run_first[r] = spos[i];
        run_key[r] = skeys[i];
        run_iota[r] = r;
    }
}

// dense_of_run[run sorted k-th by first sighting] = k; user_ids[k] = its id.
__global__ void k_dense_of_run(const uint32_t* __restrict__ runs_by_first, const uint64_t* __restrict__ run_key,
                               uint64_t nruns, uint32_t* __restrict__ dense_of_run, int64_t* __restrict__ users) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nruns; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = runs_by_first[k];
        dense_of_run[r] = (uint32_t)k;
        if (users) users[k] = (int64_t)run_key[r];
    }
}

__global__ void k_scatter_dense(const uint32_t* __restrict__ spos, const uint32_t* __restrict__ run_of,
                                const uint32_t* __restrict__ dense_of_run, uint64_t cnt,
                                uint32_t* __restrict__ src_dense, uint32_t* __restrict__ dst_dense) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = spos[i];
        const uint32_t d = dense_of_run[run_of[i]];
        if (p & 1u)
            dst_dense[p >> 1] = d;
        else
            src_dense[p >> 1] = d;
    }
}

__global__ void k_inclusive_heads(const uint32_t* __restrict__ excl, const uint32_t* __restrict__ head, uint64_t cnt,
                                  uint32_t* __restrict__ incl) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x)
        incl[i] = excl[i] + head[i];
}

}  // namespace

// Interning over a sequence of (src, dst) batches (IdMap::from_edge_batches,
// id_map.rs:220-292): rows are numbered across the batches in order, each
// batch's endpoints are written straight into the occurrence keys at its row
// offset (the batches are never concatenated), and the dense ids of all rows
// come out in batch order.
uint64_t intern_batches(uint64_t nb, const uint64_t* lens, const int64_t* const* srcs, const int64_t* const* dsts,
                        uint32_t* src_out, uint32_t* dst_out, int64_t* users_out) {
    uint64_t m = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        URSA_REQUIRE(lens[b] == 0 || (srcs[b] && dsts[b]), URSA_HIP_EINVAL, "NULL edge batch");
        m += lens[b];
    }
    if (m == 0) return 0;
    URSA_REQUIRE(2 * m <= 0xFFFFFFFFull, URSA_HIP_ERANGE, "too many edge rows for u32 occurrence positions");
    hipStream_t s = stream();
    const uint64_t cnt = 2 * m;
    DevOut<uint32_t> sd(src_out, m), dd(dst_out, m);
    DevBuf<uint64_t> keys(cnt), skeys(cnt);
    DevBuf<uint32_t> pos(cnt), spos(cnt);
    uint64_t row0 = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        if (lens[b] == 0) continue;
        DevIn<int64_t> src(srcs[b], lens[b], s), dst(dsts[b], lens[b], s);
        k_occurrences<<<blocks_for(lens[b], IB, 1u << 16), IB, 0, s>>>(src.p, dst.p, lens[b], row0, keys.get(),
                                                                      pos.get());
        check_launch();
        row0 += lens[b];
    }
    prim::sort_pairs_u64_u32(keys.get(), skeys.get(), pos.get(), spos.get(), cnt, 64, s);
    keys.reset();
    pos.reset();
    DevBuf<uint32_t> head(cnt), first(cnt), excl(cnt + 1), run_of(cnt);
    k_run_heads<<<blocks_for(cnt, IB, 1u << 16), IB, 0, s>>>(skeys.get(), spos.get(), cnt, head.get(), first.get());
    check_launch();
    prim::exclusive_sum_u32(head.get(), excl.get(), cnt, s);
    k_inclusive_heads<<<blocks_for(cnt, IB, 1u << 16), IB, 0, s>>>(excl.get(), head.get(), cnt, run_of.get());
    k_minus_one<<<blocks_for(cnt, IB, 1u << 16), IB, 0, s>>>(run_of.get(), cnt);
    check_launch();
    uint32_t nruns32 = 0;
    HIP_OK(hipMemcpyAsync(&nruns32, excl.get() + cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const uint64_t nruns = nruns32;
    DevBuf<uint32_t> run_first(nruns), run_iota(nruns), sfirst(nruns), runs_by_first(nruns), dense_of_run(nruns);
    DevBuf<uint64_t> run_key(nruns);
    k_run_table<<<blocks_for(cnt, IB, 1u << 16), IB, 0, s>>>(head.get(), run_of.get(), spos.get(), skeys.get(), cnt,
                                                             run_first.get(), run_key.get(), run_iota.get());
    check_launch();
    prim::sort_pairs_u32(run_first.get(), sfirst.get(), run_iota.get(), runs_by_first.get(), nruns,
                         bits_for(cnt), s);
    DevOut<int64_t> users(users_out, nruns);
    k_dense_of_run<<<blocks_for(nruns, IB, 1u << 16), IB, 0, s>>>(runs_by_first.get(), run_key.get(), nruns,
                                                                  dense_of_run.get(), users.p);
    k_scatter_dense<<<blocks_for(cnt, IB, 1u << 16), IB, 0, s>>>(spos.get(), run_of.get(), dense_of_run.get(), cnt,
                                                                 sd.p, dd.p);
    check_launch();
    sd.finish(s);
    dd.finish(s);
    users.finish(s);
    return nruns;
}

uint64_t intern_i64(uint64_t m, const int64_t* src_user, const int64_t* dst_user, uint32_t* src_out,
                    uint32_t* dst_out, int64_t* users_out) {
    return intern_batches(1, &m, &src_user, &dst_user, src_out, dst_out, users_out);
}

}  // namespace ursa

extern "C" int ursa_hip_intern_i64_batches(uint64_t nbatches, const uint64_t* lens, const int64_t* const* srcs,
                                           const int64_t* const* dsts, uint32_t* src_dense, uint32_t* dst_dense,
                                           int64_t* user_ids, uint64_t* n_out) {
    return ursa::guard([&] {
        URSA_REQUIRE(n_out != nullptr, URSA_HIP_EINVAL, "n_out is NULL");
        *n_out = 0;
        if (nbatches == 0) return;
        URSA_REQUIRE(lens && srcs && dsts, URSA_HIP_EINVAL, "NULL batch table");
        uint64_t m = 0;
        for (uint64_t b = 0; b < nbatches; ++b) m += lens[b];
        if (m == 0) return;
        URSA_REQUIRE(src_dense && dst_dense, URSA_HIP_EINVAL, "NULL output buffer");
        *n_out = ursa::intern_batches(nbatches, lens, srcs, dsts, src_dense, dst_dense, user_ids);
    });
}

extern "C" int ursa_hip_intern_i64(uint64_t m, const int64_t* src, const int64_t* dst, uint32_t* src_dense,
                                   uint32_t* dst_dense, int64_t* user_ids, uint64_t* n_out) {
    return ursa::guard([&] {
        URSA_REQUIRE(n_out != nullptr, URSA_HIP_EINVAL, "n_out is NULL");
        *n_out = 0;
        if (m == 0) return;
        URSA_REQUIRE(src && dst && src_dense && dst_dense, URSA_HIP_EINVAL, "NULL edge or output buffer");
        *n_out = ursa::intern_i64(m, src, dst, src_dense, dst_dense, user_ids);
    });
}
