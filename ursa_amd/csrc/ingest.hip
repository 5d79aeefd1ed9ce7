// ingest.hip — IdMap interning of int64 edge endpoints on the device
// (IdMap::from_edge_batches / intern_i64, ursa-core/src/id_map.rs:220-319).
//
// Dense ids are assigned in first-seen order over the edge rows, src before
// dst within a row (id_map.rs:251-254). Endpoint occurrence p = 2*row + side
// (side 0 = src, 1 = dst) is the scan position. On the device:
//   1. every occurrence is inserted into an open-addressing hash table keyed by
//      the user id (linear probing, 16-byte slots: id, first position, dense
//      id); the slot keeps the smallest position that inserted it (atomicMin),
//      i.e. the id's first sighting;
//   2. the occupied slots are compacted and their first positions sorted: the
//      id sighted k-th receives dense id k, and user_ids[k] = id;
//   3. every row looks its two ids up and writes their dense ids in row order.
// The result is the reference's numbering exactly (the table's slot order never
// reaches the output). The first table (at most 2^23 slots, 128 MiB) stays in
// the 256 MB infinity cache; an insert that probes more than TAB_PROBES slots
// marks it too full, and it is rebuilt four times larger, up to twice the
// occurrence count, where every id fits and probing is not limited. The id 0x8000000000000000
// marks empty slots, so that one id is counted beside the table.
#include <vector>

#include "../../include/ursa_hip_ext.h"
#include "prim.hpp"
#include "topology.hpp"

namespace ursa {
namespace {

constexpr unsigned IB = 256;
constexpr unsigned long long EMPTY = 0x8000000000000000ull;  // also a valid id, kept outside the table
constexpr uint64_t FIRST_SLOTS = 1ull << 23;
constexpr uint64_t TAB_PROBES = 128;

struct __align__(16) Slot {
    unsigned long long key;
    uint32_t first;  // smallest occurrence position seen
    uint32_t dense;
};

struct TabCtl {
    uint32_t overflow;     // an insert gave up after its probe limit
    uint32_t empty_first;  // first position of the id EMPTY (or ~0)
    uint32_t empty_dense;
    uint32_t pad;
};

__device__ __forceinline__ uint64_t id_hash(uint64_t x) {
    x ^= x >> 33;
    x *= 0xFF51AFD7ED558CCDull;
    x ^= x >> 33;
    x *= 0xC4CEB9FE1A85EC53ull;
    return x ^ (x >> 33);
}

__global__ void k_tab_clear(Slot* __restrict__ tab, uint64_t cap, TabCtl* ctl) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x)
        tab[i] = Slot{EMPTY, 0xFFFFFFFFu, 0u};
    if (blockIdx.x == 0 && threadIdx.x == 0) *ctl = TabCtl{0u, 0xFFFFFFFFu, 0u, 0u};
}

// Inserts id with occurrence position p, probing at most `probes` slots (the
// caller rebuilds a larger table when any insert gives up).
__device__ __forceinline__ void tab_insert(Slot* __restrict__ tab, uint64_t mask, uint64_t probes, TabCtl* ctl,
                                           uint64_t id, uint32_t p) {
    if (id == EMPTY) {
        if (p < ctl->empty_first) atomicMin(&ctl->empty_first, p);
        return;
    }
    uint64_t h = id_hash(id) & mask;
    for (uint64_t step = 0; step < probes; ++step, h = (h + 1) & mask) {
        unsigned long long k = tab[h].key;
        if (k == EMPTY) {
            k = atomicCAS(&tab[h].key, EMPTY, (unsigned long long)id);
            if (k == EMPTY) k = id;
        }
        if (k == id) {
            if (p < tab[h].first) atomicMin(&tab[h].first, p);
            return;
        }
    }
    ctl->overflow = 1u;
}

// Occurrences of one batch of m rows whose first row is global row row0.
// Rows handled per thread per step: their ids and home slots are loaded
// together before any is resolved (most ids sit in their home slot).
constexpr int ROWS_ILP = 4;

// A thread takes ROWS_ILP rows a step (r, r + stride, ...): loads the 2*ROWS_ILP
// ids and their home slots first, settles ids found at home (atomicMin of the
// first position when it may still fall), and probes for the rest.
__global__ void k_tab_insert(const int64_t* __restrict__ src, const int64_t* __restrict__ dst, uint64_t m,
                             uint64_t row0, Slot* __restrict__ tab, uint64_t mask, uint64_t probes, TabCtl* ctl) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r0 < m; r0 += stride * ROWS_ILP) {
        uint64_t id[2 * ROWS_ILP];
        Slot home[2 * ROWS_ILP];
#pragma unroll
        for (int j = 0; j < ROWS_ILP; ++j) {
            const uint64_t r = r0 + j * stride;
            id[2 * j] = r < m ? (uint64_t)src[r] : EMPTY;
            id[2 * j + 1] = r < m ? (uint64_t)dst[r] : EMPTY;
        }
#pragma unroll
        for (int j = 0; j < 2 * ROWS_ILP; ++j) home[j] = tab[id_hash(id[j]) & mask];
#pragma unroll
        for (int j = 0; j < 2 * ROWS_ILP; ++j) {
            const uint64_t r = r0 + (j >> 1) * stride;
            if (r >= m) continue;
            const uint32_t p = (uint32_t)(2 * (row0 + r)) + (uint32_t)(j & 1);
            if (id[j] != EMPTY && home[j].key == id[j]) {
                if (p < home[j].first) atomicMin(&tab[id_hash(id[j]) & mask].first, p);
            } else {
                tab_insert(tab, mask, probes, ctl, id[j], p);
            }
        }
    }
}

// Occupied slots: (first position, slot index), in any order; the id EMPTY, if
// seen, as slot index `cap`. A workgroup takes COLLECT_CHUNK consecutive slots
// (each wave a quarter of them), counts its occupied slots, and reserves its
// output range with one atomic add.
constexpr uint64_t COLLECT_CHUNK = 4096;
__global__ __launch_bounds__(IB) void k_tab_collect(const Slot* __restrict__ tab, uint64_t cap, const TabCtl* ctl,
                                                    uint32_t* __restrict__ first, uint32_t* __restrict__ slot,
                                                    uint32_t* __restrict__ count) {
    constexpr uint64_t PER_WAVE = COLLECT_CHUNK / (IB / 64);
    __shared__ uint32_t s_cnt[IB / 64];
    __shared__ uint32_t s_base;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const uint64_t w0 = blockIdx.x * COLLECT_CHUNK + wave * PER_WAVE;
    const uint64_t w1 = min(cap, w0 + PER_WAVE);
    uint32_t n = 0;
    for (uint64_t i = w0; i < w1; i += 64) {
        const bool used = i + lane < w1 && tab[i + lane].key != EMPTY;
        n += (uint32_t)__popcll(__ballot(used));
    }
    if (lane == 0) s_cnt[wave] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < IB / 64; ++w) tot += s_cnt[w];
        s_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    uint32_t at = s_base;
    for (int w = 0; w < wave; ++w) at += s_cnt[w];
    for (uint64_t i = w0; i < w1; i += 64) {
        const bool used = i + lane < w1 && tab[i + lane].key != EMPTY;
        const uint64_t mask = __ballot(used);
        if (used) {
            const uint32_t o = at + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            first[o] = tab[i + lane].first;
            slot[o] = (uint32_t)(i + lane);
        }
        at += (uint32_t)__popcll(mask);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && ctl->empty_first != 0xFFFFFFFFu) {
        const uint32_t o = atomicAdd(count, 1u);
        first[o] = ctl->empty_first;
        slot[o] = (uint32_t)cap;
    }
}

// The id sighted k-th (slots sorted by first position) gets dense id k.
__global__ void k_tab_rank(const uint32_t* __restrict__ slot_by_first, uint64_t nids, uint64_t cap,
                           Slot* __restrict__ tab, TabCtl* ctl, int64_t* __restrict__ users) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nids; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = slot_by_first[k];
        uint64_t id;
        if (i == cap) {
            ctl->empty_dense = (uint32_t)k;
            id = EMPTY;
        } else {
            tab[i].dense = (uint32_t)k;
            id = tab[i].key;
        }
        if (users) users[k] = (int64_t)id;
    }
}

__device__ __forceinline__ uint32_t tab_find(const Slot* __restrict__ tab, uint64_t mask, const TabCtl* ctl,
                                             uint64_t id) {
    if (id == EMPTY) return ctl->empty_dense;
    uint64_t h = id_hash(id) & mask;
    for (uint64_t step = 0; step <= mask; ++step, h = (h + 1) & mask) {
        const Slot sl = tab[h];
        if (sl.key == id) return sl.dense;  // every id of the rows was inserted
    }
    return 0xFFFFFFFFu;
}

// ROWS_ILP rows a step as in k_tab_insert: ids and home slots loaded first,
// probes only for ids away from home.
__global__ void k_tab_map(const int64_t* __restrict__ src, const int64_t* __restrict__ dst, uint64_t m,
                          const Slot* __restrict__ tab, uint64_t mask, const TabCtl* ctl,
                          uint32_t* __restrict__ src_dense, uint32_t* __restrict__ dst_dense) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r0 < m; r0 += stride * ROWS_ILP) {
        uint64_t id[2 * ROWS_ILP];
        Slot home[2 * ROWS_ILP];
#pragma unroll
        for (int j = 0; j < ROWS_ILP; ++j) {
            const uint64_t r = r0 + j * stride;
            id[2 * j] = r < m ? (uint64_t)src[r] : EMPTY;
            id[2 * j + 1] = r < m ? (uint64_t)dst[r] : EMPTY;
        }
#pragma unroll
        for (int j = 0; j < 2 * ROWS_ILP; ++j) home[j] = tab[id_hash(id[j]) & mask];
#pragma unroll
        for (int j = 0; j < ROWS_ILP; ++j) {
            const uint64_t r = r0 + j * stride;
            if (r >= m) continue;
            const uint64_t a = id[2 * j], b = id[2 * j + 1];
            src_dense[r] = a != EMPTY && home[2 * j].key == a ? home[2 * j].dense : tab_find(tab, mask, ctl, a);
            dst_dense[r] = b != EMPTY && home[2 * j + 1].key == b ? home[2 * j + 1].dense : tab_find(tab, mask, ctl, b);
        }
    }
}

uint64_t pow2_at_least(uint64_t x) {
    uint64_t c = 1;
    while (c < x) c <<= 1;
    return c;
}

}  // namespace

// Interning over a sequence of (src, dst) batches (IdMap::from_edge_batches,
// id_map.rs:220-292): rows are numbered across the batches in order and each
// batch is inserted and mapped where it lies (the batches are never
// concatenated); the dense ids of all rows come out in batch order.
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
    const uint64_t occ = 2 * m;
    std::vector<DevIn<int64_t>> in_src, in_dst;
    in_src.reserve(nb);
    in_dst.reserve(nb);
    for (uint64_t b = 0; b < nb; ++b) {
        in_src.emplace_back(srcs[b], lens[b], s);
        in_dst.emplace_back(dsts[b], lens[b], s);
    }
    DevBuf<TabCtl> ctl(1);
    // Cache-sized table first; each rebuild is four times larger, and one of
    // twice the occurrence count holds every id at most half full.
    const uint64_t full = std::max<uint64_t>(1024, pow2_at_least(2 * occ));
    uint64_t cap = std::min<uint64_t>(full, FIRST_SLOTS);
    DevBuf<Slot> tab;
    for (;;) {
        tab.alloc(cap);
        k_tab_clear<<<blocks_for(cap, IB, 1u << 14), IB, 0, s>>>(tab.get(), cap, ctl.get());
        check_launch();
        uint64_t row0 = 0;
        for (uint64_t b = 0; b < nb; ++b) {
            if (lens[b] == 0) continue;
            k_tab_insert<<<blocks_for(lens[b], IB, 1u << 16), IB, 0, s>>>(
                in_src[b].p, in_dst[b].p, lens[b], row0, tab.get(), cap - 1, cap == full ? cap : TAB_PROBES, ctl.get());
            check_launch();
            row0 += lens[b];
        }
        TabCtl h{};
        HIP_OK(hipMemcpyAsync(&h, ctl.get(), sizeof(TabCtl), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        if (!h.overflow) break;
        URSA_REQUIRE(cap < full, URSA_HIP_EINTERNAL, "id table overflow at full size");
        cap = std::min<uint64_t>(full, cap * 4);
        tab.reset();
    }
    DevBuf<uint32_t> first(cap + 1), slot(cap + 1), count(1);
    HIP_OK(hipMemsetAsync(count.get(), 0, sizeof(uint32_t), s));
    k_tab_collect<<<(unsigned)((cap + COLLECT_CHUNK - 1) / COLLECT_CHUNK), IB, 0, s>>>(
        tab.get(), cap, ctl.get(), first.get(), slot.get(), count.get());
    check_launch();
    uint32_t nids32 = 0;
    HIP_OK(hipMemcpyAsync(&nids32, count.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const uint64_t nids = nids32;
    DevBuf<uint32_t> sfirst(nids), slot_by_first(nids);
    prim::sort_pairs_u32(first.get(), sfirst.get(), slot.get(), slot_by_first.get(), nids, bits_for(occ), s);
    DevOut<int64_t> users(users_out, nids);
    k_tab_rank<<<blocks_for(nids, IB, 1u << 14), IB, 0, s>>>(slot_by_first.get(), nids, cap, tab.get(), ctl.get(),
                                                            users.p);
    check_launch();
    DevOut<uint32_t> sd(src_out, m), dd(dst_out, m);
    uint64_t row0 = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        if (lens[b] == 0) continue;
        k_tab_map<<<blocks_for(lens[b], IB, 1u << 16), IB, 0, s>>>(in_src[b].p, in_dst[b].p, lens[b], tab.get(),
                                                                   cap - 1, ctl.get(), sd.p + row0, dd.p + row0);
        check_launch();
        row0 += lens[b];
    }
    sd.finish(s);
    dd.finish(s);
    users.finish(s);
    return nids;
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
