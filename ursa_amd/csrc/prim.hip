// prim.hip — thin wrappers over rocPRIM's device-wide algorithms, kept in
// one translation unit so its heavy templates compile once.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "common.hpp"
#include "prim.hpp"

namespace ursa {
namespace prim {

namespace {
template <class F>
void run_twice(F f, hipStream_t s) {
    size_t bytes = 0;
    HIP_OK(f(nullptr, bytes));
    DevBuf<char> tmp(bytes ? bytes : 1);
    HIP_OK(f(tmp.get(), bytes));
    // keep tmp alive until the work has run
    HIP_OK(hipStreamSynchronize(s));
}
}  // namespace

void sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout,
                    size_t n, int end_bit, hipStream_t s) {
    if (n == 0) return;
    run_twice([&](void* t, size_t& b) {
        return rocprim::radix_sort_pairs(t, b, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, s);
    }, s);
}

void sort_pairs_u64_u32(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, size_t n,
                        int end_bit, hipStream_t s) {
    if (n == 0) return;
    run_twice([&](void* t, size_t& b) {
        return rocprim::radix_sort_pairs(t, b, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, s);
    }, s);
}

void sort_keys_u64(const uint64_t* kin, uint64_t* kout, size_t n, int end_bit, hipStream_t s) {
    if (n == 0) return;
    run_twice([&](void* t, size_t& b) {
        return rocprim::radix_sort_keys(t, b, kin, kout, n, 0u, (unsigned)end_bit, s);
    }, s);
}

void inclusive_max_scan_u32(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s) {
    if (n == 0) return;
    run_twice([&](void* t, size_t& b) {
        return rocprim::inclusive_scan(t, b, in, out, n, rocprim::maximum<uint32_t>(), s);
    }, s);
}

void exclusive_sum_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, hipStream_t s) {
    if (n == 0) return;
    run_twice([&](void* t, size_t& b) {
        return rocprim::exclusive_scan(t, b, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s);
    }, s);
}

void exclusive_sum_u32(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s) {
    // Inclusive scan into out[1..n], out[0] = 0.
    HIP_OK(hipMemsetAsync(out, 0, sizeof(uint32_t), s));
    if (n == 0) return;
    run_twice([&](void* t, size_t& b) {
        return rocprim::inclusive_scan(t, b, in, out + 1, n, rocprim::plus<uint32_t>(), s);
    }, s);
}

void unique_u64(const uint64_t* in, uint64_t* out, uint64_t* d_count, size_t n, hipStream_t s) {
    run_twice([&](void* t, size_t& b) {
        return rocprim::unique(t, b, in, out, d_count, n, rocprim::equal_to<uint64_t>(), s);
    }, s);
}

void select_flagged_u32(const uint32_t* in, const uint8_t* flags, uint32_t* out, uint64_t* d_count,
                        size_t n, hipStream_t s) {
    run_twice([&](void* t, size_t& b) {
        return rocprim::select(t, b, in, flags, out, d_count, n, s);
    }, s);
}

}  // namespace prim
}  // namespace ursa
