// metrics.hpp — launch interface of the metrics and diagnostics reductions (metrics.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace sdr {

// most blocks a reduction launch uses (the per-block partials are that many records)
constexpr int kStatBlocks = 2048;

struct CellStats {
  double psum, pmax, wsum, wpeak;
  unsigned long long nwashed, nfluid, nart, nwet, nacc, maxcount;
};
struct PartStats {
  double vmax, ke;
};
// One frame's reductions:
//   psum / pmax / nfluid   pressure over Fluid cells (sum, max(0, .), count)
//   wsum / wpeak / nacc    wet wall cells at or below the artifact cap
//   nart / nwet            wet wall cells above the cap / all wet wall cells
//   nwashed                wall cells washed so far
//   maxcount               largest per-cell particle count (when counts are given)
//   vmax / ke              max |v|, sum |v|^2 over the particles
struct FrameStats {
  double psum, pmax, wsum, wpeak, vmax, ke;
  unsigned long long nwashed, nfluid, nart, nwet, nacc, maxcount;
};

void launch_wall_cells(const uint8_t* mask, int nx, int ny, int nz, uint8_t* is_wall, unsigned long long* count,
                       hipStream_t st);
// is_wall / washed / wet may be null (diagnostics only); count may be null.
template <class T>
void launch_frame_stats(const Parts<T>& p, int64_t n, GridD<T> g, const uint8_t* cells, const T* pressure,
                        const uint8_t* is_wall, uint8_t* washed, uint8_t* wet, double cap, const uint32_t* count,
                        CellStats* cpart, PartStats* ppart, FrameStats* out, hipStream_t st);

}  // namespace sdr
