// sort.hpp — the scans and sorts of the substep (sort.hip), hand-written for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "kernels.hpp"

namespace sdr {

// Bytes of working memory the scans need for n elements (one sum per 4096-element tile).
size_t scan_u32_temp(int64_t n);

// out[i] = in[0] + ... + in[i-1].  Integer sums: the result does not depend on the
// launch shape.  `temp` holds at least scan_u32_temp(n) bytes.
hipError_t exclusive_sum_u32(void* temp, size_t temp_bytes, const uint32_t* in, uint32_t* out,
                             int64_t n, hipStream_t st);

// Exclusive sums over cells of (count > cap) (sizes = false) or of (count > cap ? count :
// 0) (sizes = true), without materialising the summands.
hipError_t exclusive_sum_over_cap(void* temp, size_t temp_bytes, const uint32_t* count, uint32_t cap,
                                  bool sizes, uint32_t* out, int64_t n, hipStream_t st);

// Sorts of small arrays by unsigned 32-bit key (ascending; equal keys keep no particular
// order, so callers pass unique keys): kout (and vout) receive the sorted copy of kin (and
// vin).  `scratch_k`/`scratch_v` hold n entries each (the merge passes ping-pong).
hipError_t sort_keys_u32(const uint32_t* kin, uint32_t* kout, uint32_t* scratch_k, int64_t n, hipStream_t st);
hipError_t sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout,
                          uint32_t* scratch_k, uint32_t* scratch_v, int64_t n, hipStream_t st);

// Working arrays of the particle sort by cell.
struct CellSortBufs {
  uint32_t* key;     // per particle: its clamped cell
  uint32_t* perm;    // per slot of the sorted order: the particle's current index
  uint32_t* skey;    // per slot: the order key (exact mode: the reference rank)
  uint32_t* tperm;   // merge scratch (per slot)
  uint32_t* tskey;   // merge scratch (per slot, exact mode)
  uint32_t* cstart;  // per cell: first slot of its segment
  uint32_t* cend;    // per cell: one past its last slot
  uint32_t* cursor;  // per cell: scatter cursor
  uint32_t* big;     // cells whose segment is too long for one thread's registers
  uint32_t* ctr;     // [0]: entries of `big`
};

// The particles' sorted order by clamped cell (counting sort): cell segments
// [cstart, cend) and perm (slot -> current index).  Inside a segment the order is the
// reference rank in exact mode, the current index otherwise (a stable sort).
template <class T>
void cell_sort(const Parts<T>& p, int64_t n, GridD<T> g, bool exact, const CellSortBufs& b, void* temp,
               size_t temp_bytes, hipStream_t st);

}  // namespace sdr
