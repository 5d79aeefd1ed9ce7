// prim.hpp — device-wide sort / scan / select primitives (rocPRIM), used by
// the topology builders. All run on the given stream; temporaries are owned
// by the call.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ursa {
namespace prim {

// Stable LSD radix sort of (key, value) pairs on key bits [0, end_bit).
void sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout,
                    size_t n, int end_bit, hipStream_t s);
// Stable radix sort of (u64 key, u32 value) pairs on key bits [0, end_bit).
void sort_pairs_u64_u32(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, size_t n,
                        int end_bit, hipStream_t s);
// Radix sort of u64 keys on bits [0, end_bit).
void sort_keys_u64(const uint64_t* kin, uint64_t* kout, size_t n, int end_bit, hipStream_t s);
// out[i] = max(in[0..=i]).
void inclusive_max_scan_u32(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s);
// out[i] = sum(in[0..i)) (u64 accumulation); returns nothing, total at out[n] if with_total.
void exclusive_sum_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, hipStream_t s);
// out[i] = sum(in[0..i)) in u32, for i in [0, n]; out has n + 1 entries.
void exclusive_sum_u32(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s);
// Adjacent-duplicate removal; the count is written to *d_count (device).
void unique_u64(const uint64_t* in, uint64_t* out, uint64_t* d_count, size_t n, hipStream_t s);
// Stable compaction of u32 values whose flag is non-zero; count to *d_count.
void select_flagged_u32(const uint32_t* in, const uint8_t* flags, uint32_t* out, uint64_t* d_count,
                        size_t n, hipStream_t s);

}  // namespace prim
}  // namespace ursa
