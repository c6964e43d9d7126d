// mic_pipe.hpp — launch interface of the pipelined MIC(0) sweeps (mic_pipe.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"
#include "kernels.hpp"

namespace sdr {

// Cross-tile state of one sweep.  Halo words carry `tag` in their high 32 bits; a
// consumer waits until the words it needs carry the current sweep's tag, so tags must
// differ between consecutive sweeps that share the buffers (and the buffers start
// zeroed, tags start at 1).  `error` is set to 1 if any bounded wait expired.
template <class T>
struct MicPipeArgs {
  uint64_t* halo_x;  // [(ntile_i-1) x nz x NJ x V x words(T)]
  uint64_t* halo_z;  // [(ntile_k-1) x NJ x (ntile_i*TI*64) x V x words(T)]
  T* sink;           // [1024] per-thread dummy words for out-of-grid steps
  int* error;
  long spin_limit;
  int ntile_i, ntile_k;
  uint32_t tag;
  uint32_t hx_bytes, hz_bytes;  // sizes of halo_x / halo_z (below 2^31)
  int ncu;       // compute units of the device (bounds the tiles placed by XCD)
  int xcd;       // XCD placement of the tiles (0: none); set by launch_mic_pipe
  // Backward solves of kinds 2, 4, 6, 8, 9 only: with dot_r set, the sweep also leaves one partial of
  // dot(z, dot_r) per workgroup (mic_pipe_dot_partials of them) in dot_partials.
  const T* dot_r;
  double* dot_partials;
  // Forward solves of kinds 2, 4, 6 only: with ax_p set, `in` is r and the sweep first applies
  // p += alpha ax_s, r -= alpha ax_as (alpha from the PCG scalars), leaving one max|r|
  // partial per workgroup in dot_partials.
  T* ax_p;
  const T* ax_s;
  const T* ax_as;
};

// kind 0: k_mic_solve tiles of 64*ti values of i by tk planes; kind 1: k_mic_solve2
// tiles of 64 values of i by tk waves of kt planes; kinds 2, 4, 6: k_mic_sweep3 tiles
// of 64 values of i by tk planes with an I/O wave (halo prefetch depth 8, 4, 2); kinds
// 8, 9: k_mic_sweep4 tiles of tk waves of kt planes with an I/O wave.  The factor (build) always runs the kind-0
// kernel.
struct MicPipeShape {
  int ti, tk, kt, kind, ntile_i, ntile_k, xcd;
  int pf;   // cell ring depth of the k_mic_sweep3 kinds
  int io2;  // k_mic_sweep3 with two I/O waves
};

MicPipeShape mic_pipe_shape(Shear sh, int variant, int ncu);
inline int mic_pipe_dot_partials(const MicPipeShape& s) { return s.ntile_i * s.ntile_k; }
int mic_pipe_variants();

// Sizes (64-bit words) of the halo buffers for a variant, with 2 values per cell
// (enough for every op) of up to 2 words each.
inline void mic_pipe_halo_sizes(Shear sh, MicPipeShape s, size_t* hx, size_t* hz) {
  *hx = (size_t)(s.ntile_i > 1 ? s.ntile_i - 1 : 1) * sh.nz * sh.NJ * 4;
  *hz = (size_t)(s.ntile_k > 1 ? s.ntile_k - 1 : 1) * sh.NJ * (size_t)(s.ntile_i * s.ti * 64) * 4;
}

template <class T>
void launch_mic_pipe(int op, int variant, Shear sh, const uint16_t* flags, T* precon, const T* in,
                     T* out, const PcgScalars* sc, MicPipeArgs<T> pa, hipStream_t st);

}  // namespace sdr
