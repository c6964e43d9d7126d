// bufrsrc.hpp — raw buffer loads and stores through buffer descriptors (gfx950).
//
// A descriptor bounds every access by its record count: a load whose offset is past it
// returns 0 and a store past it is dropped, with no memory traffic.  Kernels use that
// instead of bounds branches: an access that must not happen gets an offset past the
// record count (kOob, or an element index of kFarElem).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdr {

typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
using rsrc_t = __amdgpu_buffer_rsrc_t;

// Offset past every buffer's record count (record counts are kept below 2^31 bytes).
constexpr uint32_t kOob = 0x80000000u;
// an element index whose byte offset (x 2, 4 or 8) is past every plane's records
constexpr uint32_t kFarElem = 0x10000000u;
// Cache-policy bits: sc1 = agent-scope load (misses the XCD's own L2, which another
// XCD's stores do not update); sc0|sc1 = system-scope store, written through the L2 so a
// consumer on another XCD sees it at once instead of when the dirty line is evicted.
constexpr int kAgent = 16;
constexpr int kWriteThrough = 17;

__device__ inline rsrc_t make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <int AUX>
__device__ inline uint64_t bld64(rsrc_t r, uint32_t off) {
  const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX);
  return (uint64_t)v.x | ((uint64_t)v.y << 32);
}
template <int AUX>
__device__ inline void bst64(rsrc_t r, uint32_t off, uint64_t x) {
  u32x2_t v;
  v.x = (uint32_t)x;
  v.y = (uint32_t)(x >> 32);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, AUX);
}
template <class T>
__device__ inline T bld(rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t b = bld64<0>(r, off);
    T v;
    __builtin_memcpy(&v, &b, 8);
    return v;
  } else {
    const uint32_t b = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    T v;
    __builtin_memcpy(&v, &b, 4);
    return v;
  }
}
template <class T>
__device__ inline void bst(rsrc_t r, uint32_t off, T v) {
  if constexpr (sizeof(T) == 8) {
    uint64_t b;
    __builtin_memcpy(&b, &v, 8);
    bst64<0>(r, off, b);
  } else {
    uint32_t b;
    __builtin_memcpy(&b, &v, 4);
    __builtin_amdgcn_raw_buffer_store_b32(b, r, off, 0, 0);
  }
}


}  // namespace sdr
