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
                       __HIP_MEMORY_SCOPE_AGENT);
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
  const uint32_t tag = pa.tag;
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
  uint16_t cf[PF];
  T cpc[PF], cin[PF];
#pragma unroll
  for (int u = 0; u < PF; ++u) {
    const int64_t ax = xslot(u), az = zslot(u), tc = cell_of_step(u);
#pragma unroll
    for (int q = 0; q < VW; ++q) {
      rx[u][q] = ax >= 0 ? get_word(pa.halo_x + ax + q) : 0;
      rz[u][q] = az >= 0 ? get_word(pa.halo_z + az + q) : 0;
    }
    const int64_t tl = tc >= 0 ? tc : 0;
    cf[u] = flags[tl];
    if (OP != kMicBuild) {
      cpc[u] = precon[tl];
      cin[u] = in[tl];
    }
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

  for (int s0 = 0; s0 < nsteps; s0 += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int s = s0 + u;
      if (s >= nsteps) break;
      if (s > 0) {
        __syncthreads();
        if (s_abort_at < s) goto done;
      }
      const int pb = (s + 1) & 1, cb = s & 1;
      // ---- neighbour inputs
      T xin[V], zin[V];
#pragma unroll
      for (int v = 0; v < V; ++v) xin[v] = REV ? __shfl_down(xprev[v], 1) : __shfl_up(xprev[v], 1);
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
      // this step's cell inputs, then refill the ring slot with step s + PF
      const uint16_t f = cf[u];
      const T pcu = cpc[u], inu = cin[u];
      const int64_t t = cell_of_step(s);
      {
        const int64_t ax = xslot(s + PF), az = zslot(s + PF), tn = cell_of_step(s + PF);
#pragma unroll
        for (int q = 0; q < VW; ++q) {
          if (ax >= 0) rx[u][q] = get_word(pa.halo_x + ax + q);
          if (az >= 0) rz[u][q] = get_word(pa.halo_z + az + q);
        }
        const int64_t tl = tn >= 0 ? tn : 0;
        cf[u] = flags[tl];
        if (OP != kMicBuild) {
          cpc[u] = precon[tl];
          cin[u] = in[tl];
        }
      }
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
      if (t >= 0) {
        if (f & kFlFluid) {
          if (OP != kMicBuild) {
          } else {
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
        }
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
    }
  }
done:;
}

template <class T, int OP, int TI, int TK>
void launch_one(Shear sh, const uint16_t* flags, T* precon, const T* in, T* out,
                const PcgScalars* sc, MicPipeArgs<T> pa, hipStream_t st) {
  k_mic_pipe<T, OP, TI, TK><<<pa.ntile_i * pa.ntile_k, 64 * TI * TK, 0, st>>>(sh, flags, precon, in,
                                                                             out, sc, pa);
}

template <class T, int TI, int TK>
void dispatch_op(int op, Shear sh, const uint16_t* flags, T* precon, const T* in, T* out,
                 const PcgScalars* sc, MicPipeArgs<T> pa, hipStream_t st) {
  if (op == kMicBuild) launch_one<T, kMicBuild, TI, TK>(sh, flags, precon, in, out, sc, pa, st);
  if (op == kMicFwd) launch_one<T, kMicFwd, TI, TK>(sh, flags, precon, in, out, sc, pa, st);
  if (op == kMicBwd) launch_one<T, kMicBwd, TI, TK>(sh, flags, precon, in, out, sc, pa, st);
}

}  // namespace

MicPipeShape mic_pipe_shape(Shear sh, int variant) {
  // variants: (TI, TK) tiles of 64*TI values of i by TK planes
  static const int table[][2] = {{1, 16}, {1, 8}, {2, 8}, {4, 4}, {1, 4}, {2, 4}};
  int v = variant < 0 || variant >= (int)(sizeof table / sizeof table[0]) ? 0 : variant;
  MicPipeShape s;
  s.ti = table[v][0];
  s.tk = table[v][1];
  s.ntile_i = (sh.nx + 64 * s.ti - 1) / (64 * s.ti);
  s.ntile_k = (sh.nz + s.tk - 1) / s.tk;
  return s;
}

template <class T>
void launch_mic_pipe(int op, int variant, Shear sh, const uint16_t* flags, T* precon, const T* in,
                     T* out, const PcgScalars* sc, MicPipeArgs<T> pa, hipStream_t st) {
  MicPipeShape s = mic_pipe_shape(sh, variant);
  pa.ntile_i = s.ntile_i;
  pa.ntile_k = s.ntile_k;
  switch (variant) {
    case 1: dispatch_op<T, 1, 8>(op, sh, flags, precon, in, out, sc, pa, st); break;
    case 2: dispatch_op<T, 2, 8>(op, sh, flags, precon, in, out, sc, pa, st); break;
    case 3: dispatch_op<T, 4, 4>(op, sh, flags, precon, in, out, sc, pa, st); break;
    case 4: dispatch_op<T, 1, 4>(op, sh, flags, precon, in, out, sc, pa, st); break;
    case 5: dispatch_op<T, 2, 4>(op, sh, flags, precon, in, out, sc, pa, st); break;
    default: dispatch_op<T, 1, 16>(op, sh, flags, precon, in, out, sc, pa, st); break;
  }
}

template void launch_mic_pipe<double>(int, int, Shear, const uint16_t*, double*, const double*,
                                      double*, const PcgScalars*, MicPipeArgs<double>, hipStream_t);
template void launch_mic_pipe<float>(int, int, Shear, const uint16_t*, float*, const float*, float*,
                                     const PcgScalars*, MicPipeArgs<float>, hipStream_t);

}  // namespace sdr
