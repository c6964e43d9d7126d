// sort.hpp — radix sort / scan primitives used by the substep (sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sdr {

size_t sort_pairs_u32_temp(int64_t n);
size_t sort_pairs_u64_temp(int64_t n);
size_t sort_keys_u32_temp(int64_t n);
size_t scan_u32_temp(int64_t n);
size_t scan_over_cap_temp(int64_t n);

hipError_t sort_pairs_u32(void* temp, size_t temp_bytes, const uint32_t* kin, uint32_t* kout,
                          const uint32_t* vin, uint32_t* vout, int64_t n, int end_bit,
                          hipStream_t st);
hipError_t sort_pairs_u64(void* temp, size_t temp_bytes, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, int64_t n, int end_bit,
                          hipStream_t st);
hipError_t sort_keys_u32(void* temp, size_t temp_bytes, const uint32_t* kin, uint32_t* kout,
                         int64_t n, hipStream_t st);
hipError_t exclusive_sum_u32(void* temp, size_t temp_bytes, const uint32_t* in, uint32_t* out,
                             int64_t n, hipStream_t st);

// Exclusive sums over cells of (count > cap) (sizes = false) or of (count > cap ? count :
// 0) (sizes = true), without materialising the summands.
hipError_t exclusive_sum_over_cap(void* temp, size_t temp_bytes, const uint32_t* count, uint32_t cap,
                                  bool sizes, uint32_t* out, int64_t n, hipStream_t st);

}  // namespace sdr
