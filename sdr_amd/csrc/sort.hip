// sort.hip — radix sort / scan primitives (hipCUB) used by the substep.
//
// The particle sort is a stable LSD radix sort of (cell key, slot) pairs over only the
// key bits the grid needs, so particles of one cell keep their previous relative order
// (exact mode: the key carries the reference index, giving reference order per cell).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "sort.hpp"

namespace sdr {

namespace {
struct OverCap {
  uint32_t cap;
  bool sizes;
  __host__ __device__ uint32_t operator()(uint32_t c) const { return c > cap ? (sizes ? c : 1u) : 0u; }
};
using OverCapIt = hipcub::TransformInputIterator<uint32_t, OverCap, const uint32_t*>;
}  // namespace

size_t sort_pairs_u32_temp(int64_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
  return bytes;
}
size_t sort_pairs_u64_temp(int64_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 64);
  return bytes;
}
size_t scan_u32_temp(int64_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                   (int)n);
  return bytes;
}
size_t scan_over_cap_temp(int64_t n) {
  size_t bytes = 0;
  OverCapIt it((const uint32_t*)nullptr, OverCap{0, false});
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, it, (uint32_t*)nullptr, (int)n);
  return bytes;
}
size_t sort_keys_u32_temp(int64_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                    (int)n, 0, 32);
  return bytes;
}

hipError_t sort_pairs_u32(void* temp, size_t temp_bytes, const uint32_t* kin, uint32_t* kout,
                          const uint32_t* vin, uint32_t* vout, int64_t n, int end_bit,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kin, kout, vin, vout, (int)n, 0,
                                            end_bit, st);
}
hipError_t sort_pairs_u64(void* temp, size_t temp_bytes, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, int64_t n, int end_bit,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kin, kout, vin, vout, (int)n, 0,
                                            end_bit, st);
}
hipError_t sort_keys_u32(void* temp, size_t temp_bytes, const uint32_t* kin, uint32_t* kout,
                         int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  return hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, kin, kout, (int)n, 0, 32, st);
}
hipError_t exclusive_sum_u32(void* temp, size_t temp_bytes, const uint32_t* in, uint32_t* out,
                             int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, (int)n, st);
}

hipError_t exclusive_sum_over_cap(void* temp, size_t temp_bytes, const uint32_t* count, uint32_t cap,
                                  bool sizes, uint32_t* out, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  OverCapIt it(count, OverCap{cap, sizes});
  return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, it, out, (int)n, st);
}

}  // namespace sdr
