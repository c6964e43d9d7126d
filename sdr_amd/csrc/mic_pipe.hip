// mic_pipe.hip — persistent, pipelined MIC(0) sweeps (pressure.rs:442-570).
//
// The lexicographic factor/solves only allow cell (i,j,k) after (i-1,j,k), (i,j-1,k)
// and (i,j,k-1).  In the sheared solve layout (row J = i + j of plane k) those three
// are (i-1, J-1, k), (i, J-1, k) and (i, J, k-1), so:
//   * one thread owns the pencil (i, k) and walks J (row by row: coalesced loads);
//   * the x-neighbour is the previous step's value of lane i-1 (a wavefront shuffle),
//     the y-neighbour is the thread's own previous step, and the z-neighbour is the
//     previous wave (plane k-1), which runs one step ahead (LDS, one barrier a step);
//   * a workgroup is a tile of TI x 64 values of i by TK planes.  Tiles hand their
//     boundary values to the neighbouring tiles through halo words that carry the
//     sweep's tag next to the payload in one 64-bit atomic store, so a consumer just
//     polls the words it needs: no fences, no counters.  Consumers prefetch their halo
//     words PF steps ahead so the cross-die latency stays off the per-step path.
// Tiles are numbered in dependency order, so with in-order dispatch the lowest running
// tile never waits on an undispatched one; every wait is still bounded and sets an
// error flag instead of hanging.
// Each cell publishes exactly the term the reference subtracts for it, so every cell
// sees the reference operands in the reference order: the result is bit-identical to
// the sequential loops (and to the plane-by-plane kernels).
#include <hip/hip_runtime.h>
#include <limits.h>

#include <type_traits>

#include "bufrsrc.hpp"
#include "kernels.hpp"
#include "mic_pipe.hpp"

namespace sdr {
namespace {

// 64-bit halo word: tag in the high half, a 32-bit piece of the payload in the low half.
template <class T>
struct Words {
  static constexpr int n = sizeof(T) / 4;
};

template <class T>
__device__ inline void put_words(uint64_t* p, T v, uint32_t tag) {
  uint32_t w[Words<T>::n];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int q = 0; q < Words<T>::n; ++q)
    __hip_atomic_store(p + q, ((uint64_t)tag << 32) | w[q], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ inline uint64_t get_word(const uint64_t* p) {
  return __hip_atomic_load((uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T words_value(const uint64_t* w) {
  uint32_t b[Words<T>::n];
#pragma unroll
  for (int q = 0; q < Words<T>::n; ++q) b[q] = (uint32_t)w[q];
  T v;
  __builtin_memcpy(&v, b, sizeof(T));
  return v;
}

// The sweep's tag: the launch argument, or the device word a graph-replayed sweep reads
// (read once per block, uniform).
template <class T>
__device__ inline uint32_t sweep_tag(const MicPipeArgs<T>& pa) {
  if (!pa.tagp) return pa.tag;
  const uint32_t t = __hip_atomic_load(const_cast<uint32_t*>(pa.tagp), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
}

__global__ void k_tag_next(uint32_t* tagp) {
  const uint32_t t = tagp[0] + 1u;
  tagp[0] = t < kDeviceTagBase ? kDeviceTagBase : t;
}

// Value of the neighbouring lane (lane - 1 forward, lane + 1 backward) by a DPP wave
// shift of each 32-bit half; the edge lane gets garbage and is overwritten by the caller.
template <bool REV, class T>
__device__ inline T wave_shift(T v) {
  constexpr int n = sizeof(T) / 4;
  uint32_t w[n];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int q = 0; q < n; ++q)
    w[q] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w[q], REV ? 0x130 : 0x138, 0xf, 0xf, false);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// The same shift with `edge` in the lane that has no source (lane 0 shifting right,
// lane 63 shifting left): DPP keeps the old value there, so no select is needed.
template <bool REV, class T>
__device__ inline T wave_shift_in(T v, T edge) {
  constexpr int n = sizeof(T) / 4;
  uint32_t w[n], e[n];
  __builtin_memcpy(w, &v, sizeof(T));
  __builtin_memcpy(e, &edge, sizeof(T));
#pragma unroll
  for (int q = 0; q < n; ++q)
    w[q] = (uint32_t)__builtin_amdgcn_update_dpp((int)e[q], (int)w[q], REV ? 0x130 : 0x138, 0xf, 0xf, false);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// Values a cell publishes per direction: the build sends (term, modification), the
// solves send one value.
template <int OP>
struct NV {
  static constexpr int n = OP == kMicBuild ? 2 : 1;
};

// Prefetch distance (steps): 8, or 4 where an 8-deep f64 ring would not fit the
// register budget of a 1024-thread workgroup.
template <class T, int TI, int TK>
struct Depth {
  static constexpr int n = (sizeof(T) == 8 && TI * TK >= 16) ? 4 : 8;
};

template <class T, int OP, int TI, int TK>
__global__ void __launch_bounds__(64 * TI * TK)
    k_mic_pipe(Shear sh, const uint16_t* __restrict__ flags, T* __restrict__ precon,
               const T* __restrict__ in, T* __restrict__ out, const PcgScalars* sc,
               MicPipeArgs<T> pa) {
  if (sc->done) return;
  constexpr int V = NV<OP>::n;
  constexpr int W = Words<T>::n;
  constexpr int VW = V * W;
  constexpr bool REV = OP == kMicBwd;
  constexpr int PF = Depth<T, TI, TK>::n;
  __shared__ T zbuf[2][TK][TI * 64][V];
  __shared__ T xbuf[2][TK][TI][V];
  // The step at which some lane gave up waiting (INT_MAX: none).  Set during step s,
  // acted on by every wave after the barrier of step s+1, so all waves break together.
  __shared__ int s_abort_at;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wi = wave % TI, wk = wave / TI;
  const int nti = pa.ntile_i, ntk = pa.ntile_k;
  const int b = blockIdx.x;
  // Dependency order: forward tiles need (it-1, kt) and (it, kt-1), backward tiles the
  // mirror image; block b is always numbered after the tiles it waits on.
  const int it = REV ? nti - 1 - (b % nti) : b % nti;
  const int kt = REV ? ntk - 1 - (b / nti) : b / nti;
  const int i = (it * TI + wi) * 64 + lane;
  const int k = kt * TK + wk;
  const int NJ = sh.NJ;
  const int nsteps = NJ + TK - 1;
  const bool has_xnext = REV ? (it > 0) : (it + 1 < nti);
  const bool has_znext = REV ? (kt > 0) : (kt + 1 < ntk);
  // the halo slot of the boundary between tiles a and a+1 is a
  const int xb_in = REV ? it : it - 1, xb_out = REV ? it - 1 : it;
  const int zb_in = REV ? kt : kt - 1, zb_out = REV ? kt - 1 : kt;
  const int nxp = nti * TI * 64;
  const int dJ = REV ? 1 : -1;  // J offset of the x and y neighbours
  const uint32_t tag = sweep_tag(pa);
  // Which halo inputs this thread consumes.
  const bool x_consumer = (REV ? (lane == 63 && wi == TI - 1 && it + 1 < nti)
                               : (lane == 0 && wi == 0 && it > 0)) && k < sh.nz;
  const bool z_consumer = REV ? (wk == TK - 1 && kt + 1 < ntk) : (wk == 0 && kt > 0);

  if (threadIdx.x == 0) s_abort_at = INT_MAX;
  T xprev[V], yprev[V];
#pragma unroll
  for (int v = 0; v < V; ++v) xprev[v] = yprev[v] = T(0);

  // row J processed at step s, and the halo slots step s reads
  auto row_of = [&](int s) { return REV ? (NJ - 1) - (s - (TK - 1 - wk)) : s - wk; };
  auto xslot = [&](int s) -> int64_t {
    const int Jn = row_of(s) + dJ;
    if (!x_consumer || Jn < 0 || Jn >= NJ) return -1;
    return (((int64_t)xb_in * sh.nz + k) * NJ + Jn) * VW;
  };
  auto zslot = [&](int s) -> int64_t {
    const int Jz = row_of(s);
    if (!z_consumer || Jz < 0 || Jz >= NJ) return -1;
    return (((int64_t)zb_in * NJ + Jz) * nxp + i) * VW;
  };
  // the sheared index of the cell of step s, or -1 outside the grid
  auto cell_of_step = [&](int s) -> int64_t {
    const int Jc = row_of(s), jc = Jc - i;
    if (i >= sh.nx || k >= sh.nz || Jc < 0 || Jc >= NJ || jc < 0 || jc >= sh.ny) return -1;
    return ((int64_t)k * NJ + Jc) * sh.nx + i;
  };
  // Loads and stores of the cell ring are branch-free (an out-of-grid step loads slot 0
  // and stores to this thread's sink word), so the compiler can keep PF steps of loads
  // in flight and wait only for the oldest instead of draining the queue every step.
  T* const sink = pa.sink + threadIdx.x;
  // Prefetch ring: slot u holds the halo words and the cell inputs (flags, precon,
  // right-hand side) of the step s with s % PF == u, loaded PF steps ahead so neither
  // the cross-tile latency nor the HBM latency sits on the per-step critical path.
  uint64_t rx[PF][VW], rz[PF][VW];
  uint32_t cf[PF];
  T cpc[PF], cin[PF];
  // Refill of slot u with the inputs of step s (issued once the slot's previous contents
  // are dead, so each slot keeps its own registers and no copy has to wait for a load).
  auto refill = [&](int u, int s) {
    const int64_t ax = xslot(s), az = zslot(s), tc = cell_of_step(s);
#pragma unroll
    for (int q = 0; q < VW; ++q) {
      if (ax >= 0) rx[u][q] = get_word(pa.halo_x + ax + q);
      if (az >= 0) rz[u][q] = get_word(pa.halo_z + az + q);
    }
    const int64_t tl = tc >= 0 ? tc : 0;
    cf[u] = flags[tl];
    if (OP != kMicBuild) {
      cpc[u] = precon[tl];
      cin[u] = in[tl];
    }
  };
#pragma unroll
  for (int u = 0; u < PF; ++u) {
#pragma unroll
    for (int q = 0; q < VW; ++q) rx[u][q] = rz[u][q] = 0;
    refill(u, u);
  }
  __syncthreads();

  // Wait until every word of a halo value carries this sweep's tag (bounded).
  auto settle = [&](uint64_t* wv, const uint64_t* src, int s) {
    long spins = 0;
    while (true) {
      bool ok = true;
#pragma unroll
      for (int q = 0; q < VW; ++q) ok &= (uint32_t)(wv[q] >> 32) == tag;
      if (ok) return;
      if (++spins > pa.spin_limit ||
          __hip_atomic_load(pa.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        atomicMin(&s_abort_at, s);
        __hip_atomic_store(pa.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int q = 0; q < VW; ++q) wv[q] = get_word(src + q);
    }
  };

  // The step count is padded to a multiple of PF so the unrolled body has no early exit
  // (padding steps lie past the last row: no cell, no halo).
  const int nsteps_pad = (nsteps + PF - 1) / PF * PF;
  for (int s0 = 0; s0 < nsteps_pad; s0 += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int s = s0 + u;
      if (s > 0) {
        __syncthreads();
        if (s_abort_at < s) goto done;
      }
      const int pb = (s + 1) & 1, cb = s & 1;
      // ---- neighbour inputs
      T xin[V], zin[V];
#pragma unroll
      for (int v = 0; v < V; ++v) xin[v] = wave_shift<REV>(xprev[v]);
      if (REV ? lane == 63 : lane == 0) {
        if (REV ? wi != TI - 1 : wi != 0) {
#pragma unroll
          for (int v = 0; v < V; ++v) xin[v] = s > 0 ? xbuf[pb][wk][REV ? wi + 1 : wi - 1][v] : T(0);
        } else {
#pragma unroll
          for (int v = 0; v < V; ++v) xin[v] = T(0);
          const int64_t ax = xslot(s);
          if (ax >= 0) {
            settle(rx[u], pa.halo_x + ax, s);
#pragma unroll
            for (int v = 0; v < V; ++v) xin[v] = words_value<T>(&rx[u][v * W]);
          }
        }
      }
      if (REV ? wk != TK - 1 : wk != 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) zin[v] = s > 0 ? zbuf[pb][REV ? wk + 1 : wk - 1][wi * 64 + lane][v] : T(0);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) zin[v] = T(0);
        const int64_t az = zslot(s);
        if (az >= 0) {
          settle(rz[u], pa.halo_z + az, s);
#pragma unroll
          for (int v = 0; v < V; ++v) zin[v] = words_value<T>(&rz[u][v * W]);
        }
      }
      // this step's cell inputs (the slot is refilled with step s + PF at the end)
      const uint32_t f = cf[u];
      const T pcu = cpc[u], inu = cin[u];
      const int64_t t = cell_of_step(s);
      // ---- the cell; outputs for the x, y and z consumers
      T xo[V], yo[V], zo[V];
#pragma unroll
      for (int v = 0; v < V; ++v) xo[v] = yo[v] = zo[v] = T(0);
      // out-of-grid steps and non-fluid cells compute nothing; their value stays 0
      const bool live = t >= 0 && (f & kFlFluid);
      if (OP != kMicBuild) {
        T res = T(0);
        if (OP == kMicFwd) {
          // q = (((r - Tx) - Ty) - Tz) * precon with T = (plus*precon)(nb) * q(nb)
          T x = inu;
          if (f & kFlMx) x -= xin[0];
          if (f & kFlMy) x -= yprev[0];
          if (f & kFlMz) x -= zin[0];
          const T q = x * pcu;
          res = live ? q : T(0);
          const T pubv = live ? (-pcu) * q : T(0);
          xo[0] = yo[0] = zo[0] = pubv;
        } else {
          // z = (((q - Ux) - Uy) - Uz) * precon with U = (plus*precon)(c) * z(nb)
          T x = inu;
          if (f & kFlPx) x -= (-pcu) * xin[0];
          if (f & kFlPy) x -= (-pcu) * yprev[0];
          if (f & kFlPz) x -= (-pcu) * zin[0];
          const T z = x * pcu;
          res = live ? z : T(0);
          xo[0] = yo[0] = zo[0] = res;
        }
        // non-fluid in-grid cells hold 0 in q / z, so storing 0 there is the identity
        *(t >= 0 ? out + t : sink) = res;
      }
      if (OP == kMicBuild && live) {
        // e = diag - tx^2 - ty^2 - tz^2 - tau*(mx + my + mz); inputs are the
        // (plus*precon, plus*(plus'+plus'')*precon^2) pairs of the -x/-y/-z cells.
        const double TAU = 0.97, SIG = 0.25;
        const double diag = (double)fl_diag(f);
        const double tx = (double)xin[0], ty = (double)yprev[0], tz = (double)zin[0];
        const double mx = (double)xin[1], my = (double)yprev[1], mz = (double)zin[1];
        double e = diag - tx * tx - ty * ty - tz * tz - TAU * (mx + my + mz);
        if (e < SIG * diag) e = diag;
        const T pc = e > 0.0 ? (T)(1.0 / sqrt(e)) : T(0);
        precon[t] = pc;
        const double ax = (f & kFlPx) ? -1.0 : 0.0, ay = (f & kFlPy) ? -1.0 : 0.0,
                     az = (f & kFlPz) ? -1.0 : 0.0, p = (double)pc;
        xo[0] = (T)(ax * p);
        xo[1] = (T)(ax * (ay + az) * p * p);
        yo[0] = (T)(ay * p);
        yo[1] = (T)(ay * (ax + az) * p * p);
        zo[0] = (T)(az * p);
        zo[1] = (T)(az * (ax + ay) * p * p);
      }
      // ---- publish: x -> neighbour lane (shuffle next step) / LDS / halo, y -> own
      // next step, z -> next wave (LDS) / halo
      if (REV ? lane == 0 : lane == 63) {
#pragma unroll
        for (int v = 0; v < V; ++v) xbuf[cb][wk][wi][v] = xo[v];
      }
#pragma unroll
      for (int v = 0; v < V; ++v) zbuf[cb][wk][wi * 64 + lane][v] = zo[v];
      const int J = row_of(s);
      if ((REV ? lane == 0 : lane == 63) && (REV ? wi == 0 : wi == TI - 1) && has_xnext && J >= 0 &&
          J < NJ && k < sh.nz) {
        const int64_t hx = (((int64_t)xb_out * sh.nz + k) * NJ + J) * VW;
#pragma unroll
        for (int v = 0; v < V; ++v) put_words<T>(pa.halo_x + hx + v * W, xo[v], tag);
      }
      if ((REV ? wk == 0 : wk == TK - 1) && has_znext && J >= 0 && J < NJ) {
        const int64_t hz = (((int64_t)zb_out * NJ + J) * nxp + i) * VW;
#pragma unroll
        for (int v = 0; v < V; ++v) put_words<T>(pa.halo_z + hz + v * W, zo[v], tag);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        xprev[v] = xo[v];
        yprev[v] = yo[v];
      }
      refill(u, s + PF);
    }
  }
done:;
}

// ------------------------------------------------------------------ lean solve sweeps
// The forward and backward solves (apply_mic0, pressure.rs:517-570) run ~400 times per
// projection, so their per-step instruction count is what bounds them.  Same tiling and
// pipeline as k_mic_pipe, with:
//   * buffer loads/stores whose offset is pushed past the record count for steps with no
//     cell (the hardware returns 0 / drops the store): no bounds branches, no sink;
//   * 32-bit offsets advanced by a constant per step;
//   * every value a cell publishes in the solves is the same for its x, y and z
//     consumers, so one LDS row per wave serves the z neighbour and the edge lane's x
//     neighbour;
//   * halo traffic only in the waves on a tile face (wave-uniform branches); a wait that
//     expires sets the error flag and the sweep runs on to its end, so no wave leaves
//     the per-step barrier early and every wait after the flag is raised ends at once.
template <int W>
__device__ inline bool tags_ok(const uint64_t (&w)[W], uint32_t tag) {
  bool ok = true;
#pragma unroll
  for (int q = 0; q < W; ++q) ok &= (uint32_t)(w[q] >> 32) == tag;
  return ok;
}

// Prefetch depth of the solve sweeps: 8 steps, 4 where an 8-deep f64 ring would not fit
// the register budget of a 1024-thread workgroup.
template <class T, int TI, int TK>
struct SolveDepth {
  static constexpr int n = (sizeof(T) == 8 && TI * TK >= 16) ? 4 : 8;
};

template <class T, bool REV, int TI, int TK>
__global__ void __launch_bounds__(64 * TI * TK)
    k_mic_solve(Shear sh, const uint16_t* __restrict__ flags, const T* __restrict__ precon,
                const T* __restrict__ in, T* __restrict__ out, const PcgScalars* sc,
                MicPipeArgs<T> pa) {
  if (sc->done) return;
  constexpr int W = Words<T>::n;
  constexpr uint32_t SZ = sizeof(T);
  constexpr int PF = SolveDepth<T, TI, TK>::n;
  static_assert(PF % 2 == 0, "the LDS parity of a step is its ring slot's parity");
  __shared__ T zbuf[2][TK][TI * 64];
  __shared__ T xlds[2][TK];  // x halo values of the edge waves, one per plane

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wi = wave % TI, wk = wave / TI;
  const int nti = pa.ntile_i, ntk = pa.ntile_k;
  const int b = blockIdx.x;
  const int it = REV ? nti - 1 - (b % nti) : b % nti;
  const int kt = REV ? ntk - 1 - (b / nti) : b / nti;
  const int i = (it * TI + wi) * 64 + lane;
  const int k = kt * TK + wk;
  const int NJ = sh.NJ, nx = sh.nx, ny = sh.ny, nz = sh.nz;
  const int nxp = nti * TI * 64;
  const uint32_t ncell = (uint32_t)sh.size();
  const rsrc_t rf = make_rsrc(flags, ncell * 2u), rp = make_rsrc(precon, ncell * SZ),
               ri = make_rsrc(in, ncell * SZ), ro = make_rsrc(out, ncell * SZ),
               rhx = make_rsrc(pa.halo_x, pa.hx_bytes), rhz = make_rsrc(pa.halo_z, pa.hz_bytes);
  const uint32_t tag = sweep_tag(pa);
  const int dJ = REV ? -1 : 1;  // row advance per step; the x/y neighbours sit at J - dJ

  // Where this wave's neighbour values come from and where its boundary values go.
  const bool z_lds = REV ? wk + 1 < TK : wk > 0;
  const bool x_lds = REV ? wi + 1 < TI : wi > 0;
  const bool z_in = !z_lds && (REV ? kt + 1 < ntk : kt > 0);
  const bool x_tile_in = REV ? it + 1 < nti : it > 0;
  const bool x_in = !x_lds && x_tile_in && k < nz;
  // One wave fetches the x halo words of all the tile's edge waves (lane l for the wave
  // of plane l) a step ahead and hands them over through LDS, so a late producer costs
  // the tile one resync rather than one per edge wave.  It is a wave that consumes no z
  // halo.
  const bool x_fetch = x_tile_in && wave == (REV ? 0 : TI * TK - 1);
  const int kl = kt * TK + lane;  // the fetch lane's plane
  const bool x_lane = x_fetch && lane < TK && kl < nz;
  const bool z_out = (REV ? wk == 0 : wk == TK - 1) && (REV ? kt > 0 : kt + 1 < ntk);
  const bool x_out = (REV ? wi == 0 : wi == TI - 1) && (REV ? it > 0 : it + 1 < nti) && k < nz;
  const int edge_in = REV ? 63 : 0, edge_out = REV ? 0 : 63;
  const int zb_in = REV ? kt : kt - 1, zb_out = REV ? kt - 1 : kt;
  const int xb_in = REV ? it : it - 1, xb_out = REV ? it - 1 : it;
  const int zsrc = z_lds ? wk - dJ : wk;
  const int xsrc = x_lds ? (wi - dJ) * 64 + edge_out : 0;
  const bool tok = i < nx && k < nz;
  // cell index of row J: base + J * nx
  const int base = k * NJ * nx + i;
  // halo word indexes (units of 8 bytes) of row J, word q: z  zh + J*zrow + q*nxp,
  // x  xh + J*W + q
  const int zrow = W * nxp;
  const int zh_in = zb_in * NJ * zrow + i, zh_out = zb_out * NJ * zrow + i;
  const int xh_out = (xb_out * nz + k) * NJ * W;
  const int xl_in = (xb_in * nz + kl) * NJ * W;  // fetch lane's x halo row base
  const bool x_edge_out = x_out && lane == edge_out;

  const int nsteps = NJ + TK - 1;
  const int nsteps_pad = (nsteps + PF - 1) / PF * PF;
  const int J0 = REV ? (NJ - 1) + (TK - 1 - wk) : -wk;  // row of step 0
  // A fetch lane at step s serves the edge wave of plane `lane` at step s + 1, whose x
  // neighbour row equals the row that wave is at in step s.
  const int JX0 = REV ? (NJ - 1) + (TK - 1 - lane) : -lane;

  // ring slot u: inputs of the step s with s % PF == u
  uint32_t cf[PF], co[PF];
  T cp[PF], ci[PF];
  uint64_t hz[PF][W], hx[PF][W];
  auto refill = [&](int u, int J, int JX) {
    const bool v = tok && (unsigned)(J - i) < (unsigned)ny;
    const uint32_t c = (uint32_t)(base + J * nx);
    const uint32_t ot = v ? c * SZ : kOob;
    co[u] = ot;
    cf[u] = __builtin_amdgcn_raw_buffer_load_b16(rf, v ? c * 2u : kOob, 0, 0);
    cp[u] = bld<T>(rp, ot);
    ci[u] = bld<T>(ri, ot);
    if (z_in) {
      const uint32_t o = (unsigned)J < (unsigned)NJ ? (uint32_t)(zh_in + J * zrow) * 8u : kOob;
#pragma unroll
      for (int q = 0; q < W; ++q) hz[u][q] = bld64<kAgent>(rhz, o + (uint32_t)(q * nxp) * 8u);
    }
    if (x_fetch) {
      const uint32_t o = x_lane && (unsigned)JX < (unsigned)NJ ? (uint32_t)(xl_in + JX * W) * 8u : kOob;
#pragma unroll
      for (int q = 0; q < W; ++q) hx[u][q] = bld64<kAgent>(rhx, o + (uint32_t)q * 8u);
    }
  };
  // Bounded wait for halo words still carrying an older tag.
  auto settle = [&](uint64_t(&w)[W], rsrc_t r, uint32_t o, uint32_t stride) {
    long spins = 0;
    while (!tags_ok<W>(w, tag)) {
      if (++spins > pa.spin_limit ||
          __hip_atomic_load(pa.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_store(pa.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int q = 0; q < W; ++q) w[q] = bld64<kAgent>(r, o + (uint32_t)q * stride);
    }
  };

  // The word of the current step is stale: the producer is behind this wave's prefetch
  // window.  Wait until the window's newest word has arrived (rows are published in
  // order), then reload the whole window, so the next steps find fresh words instead of
  // each paying one more round trip (which would leave the consumer PF round trips
  // behind its producer for the rest of the sweep).  Word q of halo row R is at byte
  // (hbase + R * rowsz) * 8 + q * stride; slot (u + v) % PF belongs to row Jh + v * dJ.
  auto resync = [&](uint64_t(&ring)[PF][W], int u, int Jh, rsrc_t r, int hbase, int rowsz,
                    uint32_t stride) {
    int Jf = Jh + (PF - 1) * dJ;
    Jf = Jf < 0 ? 0 : (Jf >= NJ ? NJ - 1 : Jf);
    uint64_t w[W];
#pragma unroll
    for (int q = 0; q < W; ++q) w[q] = 0;  // tag 0 is never a sweep's tag
    settle(w, r, (uint32_t)(hbase + Jf * rowsz) * 8u, stride);
#pragma unroll
    for (int v = 0; v < PF; ++v) {
      const int R = Jh + v * dJ;
      const uint32_t o = (unsigned)R < (unsigned)NJ ? (uint32_t)(hbase + R * rowsz) * 8u : kOob;
#pragma unroll
      for (int q = 0; q < W; ++q) ring[(u + v) % PF][q] = bld64<kAgent>(r, o + (uint32_t)q * stride);
    }
    settle(ring[u], r, (uint32_t)(hbase + Jh * rowsz) * 8u, stride);
    // Drain the reloads here (vmcnt(0)): the fast path's waits then keep their precise
    // counts instead of assuming these loads may still be in flight.
    __builtin_amdgcn_s_waitcnt(0x0F70);
  };

  zbuf[1][wk][wi * 64 + lane] = T(0);  // read by step 0
  if (threadIdx.x < TK) xlds[1][threadIdx.x] = T(0);
#pragma unroll
  for (int u = 0; u < PF; ++u) {
#pragma unroll
    for (int q = 0; q < W; ++q) hz[u][q] = hx[u][q] = 0;
    refill(u, J0 + u * dJ, JX0 + u * dJ);
  }
  __syncthreads();

  T prev = T(0);  // this thread's value of the previous step (its x and y publication)
  int J = J0, JX = JX0;
  for (int s0 = 0; s0 < nsteps_pad; s0 += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int cb = u & 1, pb = cb ^ 1;
      const bool row_ok = (unsigned)J < (unsigned)NJ;
      // ---- neighbour values of row J
      T zin = z_lds ? zbuf[pb][zsrc][wi * 64 + lane] : T(0);
      T xe = x_lds ? zbuf[pb][wk][xsrc] : T(0);
      const T xs = wave_shift<REV>(prev);
      if (z_in && row_ok) {
        if (!tags_ok<W>(hz[u], tag)) resync(hz, u, J, rhz, zh_in, zrow, (uint32_t)nxp * 8u);
        zin = words_value<T>(hz[u]);
      }
      if (x_in) xe = xlds[pb][wk];
      const T xin = lane == edge_in ? xe : xs;
      // ---- the cell (a neighbour that does not couple publishes +0, and x - (+0) == x
      // for every x, so the unconditional forward terms equal the reference's guarded ones)
      const uint32_t f = cf[u];
      const T pcu = cp[u], inu = ci[u];
      const bool live = (f & kFlFluid) != 0;
      T res, pub;
      if (!REV) {
        const T q = (((inu - xin) - prev) - zin) * pcu;
        res = live ? q : T(0);
        pub = live ? (-pcu) * q : T(0);
      } else {
        const T npc = -pcu;
        T x = inu;
        x = x - ((f & kFlPx) ? npc * xin : T(0));
        x = x - ((f & kFlPy) ? npc * prev : T(0));
        x = x - ((f & kFlPz) ? npc * zin : T(0));
        const T z = x * pcu;
        res = live ? z : T(0);
        pub = res;
      }
      bst<T>(ro, co[u], res);
      zbuf[cb][wk][wi * 64 + lane] = pub;
      // ---- halo publication
      if (z_out && row_ok) {
        uint32_t wv[W];
        __builtin_memcpy(wv, &pub, sizeof(T));
        const uint32_t o = (uint32_t)(zh_out + J * zrow) * 8u;
#pragma unroll
        for (int q = 0; q < W; ++q)
          bst64<kWriteThrough>(rhz, o + (uint32_t)(q * nxp) * 8u, ((uint64_t)tag << 32) | wv[q]);
      }
      if (x_out && row_ok) {
        uint32_t wv[W];
        __builtin_memcpy(wv, &pub, sizeof(T));
        const uint32_t o = x_edge_out ? (uint32_t)(xh_out + J * W) * 8u : kOob;
#pragma unroll
        for (int q = 0; q < W; ++q) bst64<kWriteThrough>(rhx, o + (uint32_t)q * 8u, ((uint64_t)tag << 32) | wv[q]);
      }
      if (x_fetch) {
        T xv = T(0);
        if (x_lane && (unsigned)JX < (unsigned)NJ) {
          if (!tags_ok<W>(hx[u], tag)) resync(hx, u, JX, rhx, xl_in, W, 8u);
          xv = words_value<T>(hx[u]);
        }
        if (lane < TK) xlds[cb][lane] = xv;
      }
      prev = pub;
      refill(u, J + PF * dJ, JX + PF * dJ);
      J += dJ;
      JX += dJ;
      __syncthreads();
    }
  }
}

// Logical tile (forward dependency order) of workgroup b: i fastest.
__device__ inline void tile_of(int b, int nti, int ntk, int* itl, int* ktl) {
  (void)ntk;
  *itl = b % nti;
  *ktl = b / nti;
}

// ------------------------------------------------------------------ lean solve sweeps with an I/O wave
// The solves again, with the per-step work of the compute waves cut to the cell itself:
//   * TK compute waves (one plane each) plus an I/O wave and a publish wave per
//     workgroup, which do all the cross-tile traffic.  The I/O wave fetches (a step
//     ahead, from its prefetch ring) the z halo row of the tile's boundary plane and the
//     x halo value of every plane; the publish wave publishes (a step late, from LDS)
//     the tile's last plane and the edge lane of every plane.  A wave's wait for a load
//     also waits for the stores it issued before that load, so the halo loads and the
//     write-through halo stores are kept in different waves, and compute waves never
//     wait on halo words and never issue write-through stores.
//   * Cell loads and stores address a buffer descriptor per plane (the plane's sheared
//     slots, NJ * nx values): a lane's offset is (J * nx + i) * size, so rows before the
//     plane (J < 0) wrap to huge unsigned offsets and rows past it exceed the record
//     count, and the hardware drops them; lanes with i >= nx start at an offset that is
//     always out of range.  Slots of a row outside the grid (j < 0 or j >= ny) are
//     loaded from a far offset (not fetched: half the sheared slots), read as flags 0,
//     and compute and store 0; that select is the only bounds test in the step.
// DOT (backward solve): each compute lane also accumulates q * q over its cells in f64,
// and the workgroup leaves one partial.  The preconditioner is applied as M = L L^T (the
// forward solve L q = r, the backward L^T z = q), so dot(z, r) = r^T L^-T L^-1 r = |q|^2:
// the fast reduction mode takes the dot(z, r) of pressure.rs:383 from the backward
// solve's own input, and reads no residual.
// The forward solve needs no flags: a cell is Fluid exactly when its precon is nonzero
// (the factor writes 1/sqrt(e) to Fluid cells only; the others stay +0).  The backward
// solve of the fast mode (FL = false) needs none either: a neighbour that does not
// couple (Air or Solid, or outside the grid) publishes +0, so subtracting its term
// unconditionally gives the same values up to the sign of a zero; the reference mode
// (FL) keeps the coupling flags and the reference bits.
template <class T, bool REV, int TK, bool DOT, int PFH, bool FL = true>
__global__ void __launch_bounds__(64 * (TK + 2))
    k_mic_sweep3(Shear sh, const uint16_t* __restrict__ flags, const T* __restrict__ precon,
                 const T* __restrict__ in, T* __restrict__ out, const PcgScalars* sc,
                 MicPipeArgs<T> pa) {
  static_assert(REV || !DOT, "the dot is taken in the backward solve");
  if (sc->done) return;
  constexpr int W = Words<T>::n;
  constexpr uint32_t SZ = sizeof(T);
  // the cell ring, PF steps deep: 2 loads and a store per step forward, 3 and a store
  // backward; a step waits for the loads of PF steps before (deeper rings, 12 and 16
  // steps, measured no faster)
  constexpr int PF = 8;
  // the step loop is unrolled by U: ring slots are compile-time (u % PF for the cells,
  // (u + 1) % PFH for the halo words) and the LDS parity of step s0 + u is u & 1
  constexpr int U = PF > PFH ? PF : PFH;
  static_assert(U % PF == 0 && U % PFH == 0 && U % 2 == 0, "ring depths must divide the unroll");
  // zbuf[.][w] for w < TK: compute wave w's values of a step; zbuf[.][TK]: the fetched z
  // halo row (the boundary wave's z neighbour)
  __shared__ T zbuf[2][TK + 1][64];
  __shared__ T xlds[2][TK];
  __shared__ double dsum[TK];

  const int lane = threadIdx.x & 63;
  // wave-uniform (scalar) so the per-plane buffer descriptors live in SGPRs
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nti = pa.ntile_i, ntk = pa.ntile_k;
  int itl, ktl;
  tile_of(blockIdx.x, nti, ntk, &itl, &ktl);
  const int it = REV ? nti - 1 - itl : itl;
  const int kt = REV ? ntk - 1 - ktl : ktl;
  const int i = it * 64 + lane;
  const int NJ = sh.NJ, nx = sh.nx, nz = sh.nz;
  const int nxp = nti * 64;
  const int dJ = REV ? -1 : 1;
  const uint32_t tag = sweep_tag(pa);
  // row of compute wave w at step s
  auto row = [&](int w, int s) { return REV ? (NJ - 1) - s + (TK - 1 - w) : s - w; };
  const int nsteps = NJ + TK - 1;  // compute steps; the I/O wave runs one more to publish
  const int nsteps_pad = (nsteps + 1 + U - 1) / U * U;
  const int bwave = REV ? TK - 1 : 0;       // boundary wave: its z neighbour is in the tile before
  const int lwave = REV ? 0 : TK - 1;       // last wave: the tile after takes its z neighbour from it
  const int edge_in = REV ? 63 : 0, edge_out = REV ? 0 : 63;

  if (wave < TK) {
    // ------------------------------------------------------------ compute wave
    const int wk = wave;
    const int k = kt * TK + wk;
    const int plane = k < nz ? k : 0;
    const uint32_t recs = k < nz ? (uint32_t)(NJ * nx) : 0u;
    const T* const pbase = precon + (int64_t)plane * NJ * nx;
    const rsrc_t rf = make_rsrc(flags + (int64_t)plane * NJ * nx, recs * 2u),
                 rp = make_rsrc(pbase, recs * SZ),
                 ri = make_rsrc(in + (int64_t)plane * NJ * nx, recs * SZ),
                 ro = make_rsrc(out + (int64_t)plane * NJ * nx, recs * SZ);
    const int zsrc = wk == bwave ? TK : (REV ? wk + 1 : wk - 1);
    // element offset of (row J, lane i); lanes past the grid start out of range for good
    // (2^28 * size stays below 2^32 and above every plane's record count)
    const uint32_t lbase = i < nx ? (uint32_t)i : kFarElem;
    const int J0 = row(wk, 0);
    // byte offsets (mod 2^32, advanced by a constant per step) of this step's row (oc)
    // and of the row PF steps ahead (of2 in the flags, ofn in the values)
    uint32_t oc = (lbase + (uint32_t)(J0 * nx)) * SZ;
    uint32_t of2 = (lbase + (uint32_t)((J0 + PF * dJ) * nx)) * 2u;
    uint32_t ofn = (lbase + (uint32_t)((J0 + PF * dJ) * nx)) * SZ;
    const uint32_t d2 = (uint32_t)(dJ * nx) * 2u, dS = (uint32_t)(dJ * nx) * SZ;
    // j of the row PF steps ahead: slots outside the grid (j < 0 or j >= ny, half the
    // sheared slots) are not loaded
    int jn = J0 + PF * dJ - i;
    uint32_t cf[PF];
    T cp[PF], ci[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const bool in = (unsigned)(J0 + u * dJ - i) < (unsigned)sh.ny;
      const uint32_t e = in ? lbase + (uint32_t)((J0 + u * dJ) * nx) : kFarElem;
      cf[u] = REV && FL ? __builtin_amdgcn_raw_buffer_load_b16(rf, e * 2u, 0, 0) : 0u;
      cp[u] = bld<T>(rp, e * SZ);
      ci[u] = bld<T>(ri, e * SZ);
      // issue slot by slot, in the order of the step loop: the wait for a slot's loads
      // at the top of the loop then also holds for the ones issued here
      __builtin_amdgcn_sched_barrier(0);
    }
    double dacc = 0.0;  // |q|^2 (DOT)
    zbuf[1][wk][lane] = T(0);
    T prev = T(0);
    __syncthreads();  // the I/O wave's prologue has filled zbuf[1][TK] and xlds[1]
    for (int s0 = 0; s0 < nsteps_pad; s0 += U) {
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const int u = uu % PF;
        const int cb = uu & 1, pb = cb ^ 1;
        const T zin = zbuf[pb][zsrc][lane];
        const T xin = wave_shift_in<REV>(prev, xlds[pb][wk]);  // the edge lane takes the x halo
        const uint32_t f = cf[u];
        const T inu = ci[u];
        const T pcu = cp[u];
        const bool live = REV && FL ? (f & kFlFluid) != 0 : pcu != T(0);
        T res, pub;
        if (!REV) {
          // a neighbour that does not couple publishes +0, and x - (+0) == x for every x
          const T q = (((inu - xin) - prev) - zin) * pcu;
          res = live ? q : T(0);
          pub = live ? (-pcu) * q : T(0);
        } else {
          const T npc = -pcu;
          T x = inu;
          if (FL) {
            x = x - ((f & kFlPx) ? npc * xin : T(0));
            x = x - ((f & kFlPy) ? npc * prev : T(0));
            x = x - ((f & kFlPz) ? npc * zin : T(0));
          } else {
            x = x - npc * xin;
            x = x - npc * prev;
            x = x - npc * zin;
          }
          const T z = x * pcu;
          res = live ? z : T(0);
          pub = res;
        }
        // the value the next step's waves wait for goes to LDS first; the memory
        // operations are issued while that write completes
        zbuf[cb][wk][lane] = pub;
        __builtin_amdgcn_sched_barrier(0);
        prev = pub;
        // slots outside the grid (j < 0 or j >= ny) hold +0 from the solve's set-up and
        // are never stored to
        const bool inc = (unsigned)(jn - PF * dJ) < (unsigned)sh.ny;
        bst<T>(ro, inc ? oc : kFarElem * SZ, res);
        if (DOT) dacc += (double)inu * (double)inu;  // q: 0 outside the grid and off-fluid
        const bool inn = (unsigned)jn < (unsigned)sh.ny;
        const uint32_t en2 = inn ? of2 : kFarElem * 2u, enS = inn ? ofn : kFarElem * SZ;
        if (REV && FL) cf[u] = __builtin_amdgcn_raw_buffer_load_b16(rf, en2, 0, 0);
        cp[u] = bld<T>(rp, enS);
        ci[u] = bld<T>(ri, enS);
        oc += dS;
        of2 += d2;
        ofn += dS;
        jn += dJ;
        __syncthreads();
        // keep each step's instructions in its step: hoisted into an earlier step, the
        // use of a ring slot makes the wave wait there for the newest loads in flight
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (DOT) {
      for (int o = 32; o > 0; o >>= 1) dacc += __shfl_xor(dacc, o);
      if (lane == 0) dsum[wk] = dacc;
    }
    __syncthreads();  // matched by the I/O wave
    if (DOT && wk == 0 && lane == 0) {
      double t = 0.0;
      for (int w = 0; w < TK; ++w) t = t + dsum[w];
      pa.dot_partials[blockIdx.x] = t;
    }
    return;
  }

  const rsrc_t rhx = make_rsrc(pa.halo_x, pa.hx_bytes), rhz = make_rsrc(pa.halo_z, pa.hz_bytes);
  const int zrow = W * nxp;
  // lane l < TK serves plane kt * TK + l for the x halo
  const int kl = kt * TK + lane;
  const bool xl = lane < TK && kl < nz;
  if (wave == TK + 1) {
    // ------------------------------------------------------------ publish wave
    // Publishes the values compute waves wrote at step s - 1 (read from LDS at step s):
    // the last plane's row for the tile after in k, the edge lane of every plane for the
    // tile after in i.  Its stores are kept out of the I/O wave, whose waits for
    // prefetched halo words would otherwise also wait for these stores to complete.
    const bool z_out = REV ? kt > 0 : kt + 1 < ntk;
    const bool x_out = REV ? it > 0 : it + 1 < nti;
    const int zb_out = REV ? kt - 1 : kt, xb_out = REV ? it - 1 : it;
    const int zh_out = zb_out * NJ * zrow + i;
    const int xh_out = (xb_out * nz + (xl ? kl : 0)) * NJ * W;
    const bool zl_ok = kt * TK + lwave < nz;
    auto xpub_off = [&](int R) -> uint32_t {
      return (x_out & xl & ((unsigned)R < (unsigned)NJ)) ? (uint32_t)(xh_out + R * W) * 8u : kOob;
    };
    auto publish_at = [&](uint32_t oz, uint32_t ox, T vz, T vx) {
      uint32_t wz[W], wx[W];
      __builtin_memcpy(wz, &vz, sizeof(T));
      __builtin_memcpy(wx, &vx, sizeof(T));
#pragma unroll
      for (int q = 0; q < W; ++q) {
        bst64<kWriteThrough>(rhz, oz == kOob ? kOob : oz + (uint32_t)(q * nxp) * 8u, ((uint64_t)tag << 32) | wz[q]);
        bst64<kWriteThrough>(rhx, ox == kOob ? kOob : ox + (uint32_t)q * 8u, ((uint64_t)tag << 32) | wx[q]);
      }
    };
    // rows of the step's publish (step s - 1), advanced by dJ per step
    int rzp = row(lwave, -1), rxp = row(lane < TK ? lane : 0, -1);
    const uint32_t dz8 = (uint32_t)(dJ * zrow) * 8u;
    uint32_t ozp = (uint32_t)(zh_out + rzp * zrow) * 8u;
    __syncthreads();  // the prologue barrier
    for (int s0 = 0; s0 < nsteps_pad; s0 += U) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int pb = (u & 1) ^ 1;
        const T vz = zbuf[pb][lwave][lane];
        const T vx = zbuf[pb][lane < TK ? lane : 0][edge_out];
        publish_at((z_out & zl_ok & ((unsigned)rzp < (unsigned)NJ)) ? ozp : kOob, xpub_off(rxp), vz, vx);
        rzp += dJ, rxp += dJ;
        ozp += dz8;
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();  // matches the compute waves' final barrier
    return;
  }

  // -------------------------------------------------------------- I/O wave
  // Fetches the halo inputs: the z halo row of the boundary wave and the x halo value of
  // every plane, PFH steps ahead, into LDS a step ahead of their use.
  {
  // the I/O wave's step is the longest of the workgroup's: let its instructions issue
  // first on the SIMD it shares with a compute wave
  __builtin_amdgcn_s_setprio(3);
  const bool z_in = REV ? kt + 1 < ntk : kt > 0;
  const bool x_in = REV ? it + 1 < nti : it > 0;
  const int zb_in = REV ? kt : kt - 1;
  const int xb_in = REV ? it : it - 1;
  const int zh_in = zb_in * NJ * zrow + i;
  const int xh_in = (xb_in * nz + (xl ? kl : 0)) * NJ * W;
  // Inputs of target step t: the z halo row of the boundary wave at step t, and the x
  // neighbour row of plane `lane` at step t (its row at step t, minus dJ).
  auto zrow_of = [&](int t) { return row(bwave, t); };
  auto xrow_of = [&](int t) { return row(lane, t) - dJ; };
  // (conditions combined with & rather than &&: selects, not branches, in the step)
  auto zoff = [&](int R) -> uint32_t {
    return (z_in & ((unsigned)R < (unsigned)NJ)) ? (uint32_t)(zh_in + R * zrow) * 8u : kOob;
  };
  auto xoff = [&](int R) -> uint32_t {
    return (x_in & xl & ((unsigned)R < (unsigned)NJ)) ? (uint32_t)(xh_in + R * W) * 8u : kOob;
  };
  uint64_t hz[PFH][W], hx[PFH][W];
  uint32_t ozr[PFH], oxr[PFH];  // the offsets each ring slot was loaded from
#pragma unroll
  for (int u = 0; u < PFH; ++u) {
    ozr[u] = oxr[u] = kOob;
#pragma unroll
    for (int q = 0; q < W; ++q) hz[u][q] = hx[u][q] = 0;
  }
  auto fill_at = [&](int u, uint32_t oz, uint32_t ox) {
    ozr[u] = oz;
    oxr[u] = ox;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      hz[u][q] = bld64<kAgent>(rhz, oz + (uint32_t)(q * nxp) * 8u);
      hx[u][q] = bld64<kAgent>(rhx, ox + (uint32_t)q * 8u);
    }
  };
  auto fill = [&](int u, int t) { fill_at(u, zoff(zrow_of(t)), xoff(xrow_of(t))); };
  // A word is ready when it carries this sweep's tag, or when it is not needed (its
  // offset is out of range: the load returned 0, and the value used is 0).
  auto settle = [&](uint64_t(&w)[W], rsrc_t r, uint32_t o, uint32_t stride) {
    if (o == kOob) return;
    long spins = 0;
    while (!tags_ok<W>(w, tag)) {
      if (++spins > pa.spin_limit ||
          __hip_atomic_load(pa.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_store(pa.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int q = 0; q < W; ++q) w[q] = bld64<kAgent>(r, o + (uint32_t)q * stride);
    }
  };
  auto ready = [&](const uint64_t(&w)[W], uint32_t o) -> bool { return (o == kOob) | tags_ok<W>(w, tag); };
  // Take the inputs of target step t from ring slot u into LDS buffer b.  A stale word
  // means the producer is behind the prefetch window: wait for the window's newest word
  // (rows are published in order), then reload the whole window.
  auto take = [&](int u, int t, int b) {
    const uint32_t oz = ozr[u], ox = oxr[u];  // slot u was filled for target step t
    const bool ok = ready(hz[u], oz) & ready(hx[u], ox);  // no branch between the waits
    if (__ballot(!ok) != 0ull) {
      const uint32_t nz_ = zoff(zrow_of(t + PFH - 1)), nx_ = xoff(xrow_of(t + PFH - 1));
      uint64_t wz[W], wx[W];
#pragma unroll
      for (int q = 0; q < W; ++q) wz[q] = wx[q] = 0;  // tag 0 is never a sweep's tag
      settle(wz, rhz, nz_, (uint32_t)nxp * 8u);
      settle(wx, rhx, nx_, 8u);
#pragma unroll
      for (int v = 0; v < PFH; ++v) fill((u + v) % PFH, t + v);
      settle(hz[u], rhz, oz, (uint32_t)nxp * 8u);
      settle(hx[u], rhx, ox, 8u);
      __builtin_amdgcn_s_waitcnt(0x0F70);
    }
    zbuf[b][TK][lane] = oz == kOob ? T(0) : words_value<T>(hz[u]);
    if (lane < TK) xlds[b][lane] = ox == kOob ? T(0) : words_value<T>(hx[u]);
  };
  zbuf[0][TK][lane] = T(0);
  if (lane < TK) xlds[0][lane] = T(0);
  // The prologue issues the loads of steps 0 .. PFH in the order of the step loop, so a
  // wait at the top of the loop holds for every way into it.
#pragma unroll
  for (int u = 0; u < PFH; ++u) {
    fill(u, u);
    __builtin_amdgcn_sched_barrier(0);
  }
  take(0, 0, 1);  // inputs of step 0
  fill(0, PFH);
  __syncthreads();
  __builtin_amdgcn_sched_barrier(0);
  // the row of the step's prefetch (target s + 1 + PFH), advanced by dJ per step; the z
  // offset's row term is advanced with it (no multiply in the step)
  int rzf = zrow_of(1 + PFH), rxf = xrow_of(1 + PFH);
  const uint32_t dz8 = (uint32_t)(dJ * zrow) * 8u;
  uint32_t ozf = (uint32_t)(zh_in + rzf * zrow) * 8u;
  for (int s0 = 0; s0 < nsteps_pad; s0 += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      const int cb = u & 1;
      // inputs of step s + 1 (into buffer cb)
      const int un = (u + 1) % PFH;
      take(un, s + 1, cb);
      fill_at(un, (z_in & ((unsigned)rzf < (unsigned)NJ)) ? ozf : kOob, xoff(rxf));
      rzf += dJ, rxf += dJ;
      ozf += dz8;
      __syncthreads();
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  }
  __syncthreads();  // matches the compute waves' final barrier
}


template <class T, int TK, int PFH>
void launch_sweep3(int op, Shear sh, const uint16_t* flags, const T* precon, const T* in, T* out,
                   const PcgScalars* sc, MicPipeArgs<T> pa, hipStream_t st) {
  const int grid = pa.ntile_i * pa.ntile_k;
  if (op == kMicFwd)
    k_mic_sweep3<T, false, TK, false, PFH><<<grid, 64 * (TK + 2), 0, st>>>(sh, flags, precon, in, out, sc, pa);
  else if (pa.flagfree && pa.dot)
    k_mic_sweep3<T, true, TK, true, PFH, false><<<grid, 64 * (TK + 2), 0, st>>>(sh, flags, precon, in, out, sc, pa);
  else if (pa.flagfree)
    k_mic_sweep3<T, true, TK, false, PFH, false><<<grid, 64 * (TK + 2), 0, st>>>(sh, flags, precon, in, out, sc, pa);
  else
    k_mic_sweep3<T, true, TK, false, PFH><<<grid, 64 * (TK + 2), 0, st>>>(sh, flags, precon, in, out, sc, pa);
}

template <class T, int OP, int TI, int TK>
void launch_one(Shear sh, const uint16_t* flags, T* precon, const T* in, T* out,
                const PcgScalars* sc, MicPipeArgs<T> pa, hipStream_t st) {
  const int grid = pa.ntile_i * pa.ntile_k;
  if constexpr (OP == kMicBuild) {
    k_mic_pipe<T, OP, TI, TK><<<grid, 64 * TI * TK, 0, st>>>(sh, flags, precon, in, out, sc, pa);
  } else {
    k_mic_solve<T, OP == kMicBwd, TI, TK><<<grid, 64 * TI * TK, 0, st>>>(sh, flags, precon, in, out,
                                                                       sc, pa);
  }
}

template <class T, int TI, int TK>
void dispatch_op(int op, Shear sh, const uint16_t* flags, T* precon, const T* in, T* out,
                 const PcgScalars* sc, MicPipeArgs<T> pa, hipStream_t st) {
  if (op == kMicBuild) launch_one<T, kMicBuild, TI, TK>(sh, flags, precon, in, out, sc, pa, st);
  if (op == kMicFwd) launch_one<T, kMicFwd, TI, TK>(sh, flags, precon, in, out, sc, pa, st);
  if (op == kMicBwd) launch_one<T, kMicBwd, TI, TK>(sh, flags, precon, in, out, sc, pa, st);
}

}  // namespace

// Tile shapes.  Variants 0 and 1: k_mic_pipe (factor) / k_mic_solve tiles of 64 values
// of i by 16 and 8 planes, one plane per wave (the factor's shapes: 16 planes for f32,
// 8 for f64, whose 8-deep f64 ring needs the registers).  Variant 2 (the default solve):
// k_mic_sweep3 tiles of 64 values of i by 7 planes plus the I/O and publish waves, halo
// prefetch depth 4 (2 and 8 measured slower).  Other tile shapes (multi-plane waves, 3-15 planes per tile, deeper rings,
// two I/O waves, XCD-placed tiles) were measured slower on 256^3 and 512x512x384 and
// are not built.
namespace {
struct VariantRow {
  int ti, tk, kind;
};
const VariantRow kVariants[] = {{1, 16, 0}, {1, 8, 0}, {1, 7, 4}};
constexpr int kNumVariants = (int)(sizeof kVariants / sizeof kVariants[0]);
}  // namespace

int mic_pipe_variants() { return kNumVariants; }

MicPipeShape mic_pipe_shape(Shear sh, int variant, int ncu) {
  (void)ncu;
  const int v = variant < 0 || variant >= kNumVariants ? 0 : variant;
  const VariantRow& r = kVariants[v];
  MicPipeShape s;
  s.ti = r.ti;
  s.tk = r.tk;
  s.kind = r.kind;
  s.ntile_i = (sh.nx + 64 * s.ti - 1) / (64 * s.ti);
  s.ntile_k = (sh.nz + s.tk - 1) / s.tk;
  return s;
}

void launch_tag_next(uint32_t* tagp, hipStream_t st) { k_tag_next<<<1, 1, 0, st>>>(tagp); }

template <class T>
void launch_mic_pipe(int op, int variant, Shear sh, const uint16_t* flags, T* precon, const T* in,
                     T* out, const PcgScalars* sc, MicPipeArgs<T> pa, hipStream_t st) {
  MicPipeShape s = mic_pipe_shape(sh, variant, pa.ncu);
  pa.ntile_i = s.ntile_i;
  pa.ntile_k = s.ntile_k;
  if (s.kind == 4 && op != kMicBuild) {
    launch_sweep3<T, 7, 4>(op, sh, flags, precon, in, out, sc, pa, st);
    return;
  }
  if (s.tk == 8)
    dispatch_op<T, 1, 8>(op, sh, flags, precon, in, out, sc, pa, st);
  else
    dispatch_op<T, 1, 16>(op, sh, flags, precon, in, out, sc, pa, st);
}

template void launch_mic_pipe<double>(int, int, Shear, const uint16_t*, double*, const double*,
                                      double*, const PcgScalars*, MicPipeArgs<double>, hipStream_t);
template void launch_mic_pipe<float>(int, int, Shear, const uint16_t*, float*, const float*, float*,
                                     const PcgScalars*, MicPipeArgs<float>, hipStream_t);

}  // namespace sdr
