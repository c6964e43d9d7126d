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
  int ncu;       // compute units of the device
  // Backward solves of the I/O-wave kernel (kind 4) in the fast mode only: with dot set,
  // the sweep also leaves one partial of dot(q, q) = dot(z, r) per workgroup
  // (mic_pipe_dot_partials of them) in dot_partials.
  int dot;
  double* dot_partials;
  // Backward solves of kind 4 in the fast mode: subtract every neighbour's term without
  // the coupling flags (the same values up to the sign of a zero).
  int flagfree;
  // When set, the sweep's tag is read from device memory instead of `tag`: sweeps
  // replayed from a captured graph keep their launch arguments, so launch_tag_next
  // advances this word before each of them.  Device tags lie in [2^31, 2^32), host tags
  // below 2^31, so the two never meet in the halo words.
  const uint32_t* tagp;
};

// tagp[0] = next device tag (2^31 after a wrap); one thread.
void launch_tag_next(uint32_t* tagp, hipStream_t st);
constexpr uint32_t kDeviceTagBase = 0x80000000u;

// kind 0: k_mic_pipe (factor) / k_mic_solve tiles of 64*ti values of i by tk planes;
// kind 4: k_mic_sweep3 tiles of 64 values of i by tk planes with an I/O wave (halo
// prefetch depth 4).  The factor (build) always runs a kind-0 kernel.
struct MicPipeShape {
  int ti, tk, kind, ntile_i, ntile_k;
};

MicPipeShape mic_pipe_shape(Shear sh, int variant, int ncu);
// the default solve shape (k_mic_sweep3, 7 planes per tile)
constexpr int kMicDefaultVariant = 2;
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
