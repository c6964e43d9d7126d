// kernels.hip — gfx950 kernels of the FLIP/PIC substep (uselessgoddess/sdr hot path).
//
// Each kernel names the reference loop it replaces.  No kernel uses floating-point
// atomics: every float sum is computed by one thread in a fixed order, so results are
// bitwise reproducible; integer atomics (histograms, counters, max on non-negative
// bit patterns) are order-independent.
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include "bufrsrc.hpp"
#include "kernels.hpp"

namespace sdr {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

template <class T>
__device__ inline V3<T> load_pos(const Parts<T>& p, int64_t t) {
  return mk3(p.px[t], p.py[t], p.pz[t]);
}
template <class T>
__device__ inline V3<T> load_vel(const Parts<T>& p, int64_t t) {
  return mk3(p.vx[t], p.vy[t], p.vz[t]);
}

// ------------------------------------------------------------------ sort glue
template <class T>
__global__ void k_gather_parts(Parts<T> s, Parts<T> d, const uint32_t* __restrict__ perm,
                               int64_t n) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  uint32_t q = perm[t];
  d.px[t] = s.px[q];
  d.py[t] = s.py[q];
  d.pz[t] = s.pz[q];
  d.vx[t] = s.vx[q];
  d.vy[t] = s.vy[q];
  d.vz[t] = s.vz[q];
  d.rank[t] = s.rank[q];
}

// ------------------------------------------------------------------ classify (solver.rs:342-371)
// Per cell: the static class (outlet -> Air, !mask -> Solid, else Air) then Fluid if a
// particle whose `cell_of` is this cell lies outside the outlet.  Also counts the
// particles whose `cell_of` is this cell (volume_source, solver.rs:400-409).  Two passes,
// so a cell of thousands of particles costs no more than as many ordinary ones: the
// cells take their static class and a zero count, then every particle adds itself to
// its cell's count (one atomic per run of equal cells in a wave: the particles are in
// cell order) and marks an Air cell Fluid when it lies outside the outlet (every such
// store writes the same value).
__global__ void k_classify_init(const uint8_t* __restrict__ static_cls, int64_t nc, uint8_t* cells,
                                uint32_t* count) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  cells[c] = static_cls[c];
  count[c] = 0;
}
template <class T>
__global__ void k_classify_parts(const uint8_t* __restrict__ static_cls, Parts<T> p, int64_t n, GridD<T> g,
                                 BoxD<T> outlet, uint8_t* cells, uint32_t* count) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int c = -1;
  if (t < n) {
    const V3<T> q = load_pos(p, t);
    c = (int)g.cell_of(q);
    if (c >= 0 && static_cls[c] == kAir && !outlet.contains(q)) cells[c] = kFluid;
  }
  const int prev = __shfl_up(c, 1);
  const bool head = lane == 0 || prev != c;
  const unsigned long long heads = __ballot(head);
  if (head && c >= 0) {
    const unsigned long long after = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    atomicAdd(&count[c], (uint32_t)(end - lane));
  }
}

// ------------------------------------------------------------------ P2G (solver.rs:560-585)
// One thread per face gathers every particle whose clamped trilinear stencil
// (grid.rs:52-88) covers it, from the cell segments that can reach the face: the two
// cells straddling it along the component axis and the 3x3 cells around it across.
// The fast mode visits segments in a fixed cell order; the exact mode merges the
// segments by reference index, reproducing the reference's per-face summation order.
// The weight sum and the normalisation (solver.rs:944-952) follow in the same thread.
template <class T, int COMP, bool EXACT>
__global__ void k_p2g(GridD<T> g, Parts<T> p, const uint32_t* __restrict__ cstart,
                      const uint32_t* __restrict__ cend, T* __restrict__ out) {
  const int fnx = g.nx + (COMP == 0), fny = g.ny + (COMP == 1), fnz = g.nz + (COMP == 2);
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t nf = (int64_t)fnx * fny * fnz;
  if (f >= nf) return;
  const int fi = (int)(f % fnx), fj = (int)((f / fnx) % fny), fk = (int)(f / ((int64_t)fnx * fny));
  const T* vel = COMP == 0 ? p.vx : (COMP == 1 ? p.vy : p.vz);
  // Candidate cell ranges per axis: along the component axis {f-1, f}, across {f-1..f+1}.
  int lo[3], hi[3];
  const int fc[3] = {fi, fj, fk};
  const int nn[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = fc[a] - 1;
    hi[a] = (a == COMP) ? fc[a] : fc[a] + 1;
    if (lo[a] < 0) lo[a] = 0;
    if (hi[a] > nn[a] - 1) hi[a] = nn[a] - 1;
  }
  T acc = T(0), wsum = T(0);
  auto visit = [&](uint32_t t) {
    V3<T> l = g.local(load_pos(p, t));
    T gx = l.x - (COMP == 0 ? T(0) : T(0.5));
    T gy = l.y - (COMP == 1 ? T(0) : T(0.5));
    T gz = l.z - (COMP == 2 ? T(0) : T(0.5));
    int bi, bj, bk;
    T ffx, ffy, ffz;
    clamp_base(gx, fnx, bi, ffx);
    clamp_base(gy, fny, bj, ffy);
    clamp_base(gz, fnz, bk, ffz);
    int di = fi - bi, dj = fj - bj, dk = fk - bk;
    if ((unsigned)di > 1u || (unsigned)dj > 1u || (unsigned)dk > 1u) return;
    T wx = di ? ffx : T(1) - ffx;
    T wy = dj ? ffy : T(1) - ffy;
    T wz = dk ? ffz : T(1) - ffz;
    T w = wx * wy * wz;
    acc += w * vel[t];
    wsum += w;
  };
  if (!EXACT) {
    for (int ck = lo[2]; ck <= hi[2]; ++ck)
      for (int cj = lo[1]; cj <= hi[1]; ++cj)
        for (int ci = lo[0]; ci <= hi[0]; ++ci) {
          int64_t c = ((int64_t)ck * g.ny + cj) * g.nx + ci;
          uint32_t e = cend[c];
          for (uint32_t t = cstart[c]; t < e; ++t) visit(t);
        }
  } else {
    // k-way merge of <= 18 rank-sorted segments, fully unrolled to stay in registers.
    constexpr int S = (COMP == 0 ? 2 : 3) * (COMP == 1 ? 2 : 3) * (COMP == 2 ? 2 : 3);
    uint32_t cur[S], end[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      int ci, cj, ck;
      if (COMP == 0) {
        ci = lo[0] + s % 2;
        cj = lo[1] + (s / 2) % 3;
        ck = lo[2] + s / 6;
      } else if (COMP == 1) {
        ci = lo[0] + s % 3;
        cj = lo[1] + (s / 3) % 2;
        ck = lo[2] + s / 6;
      } else {
        ci = lo[0] + s % 3;
        cj = lo[1] + (s / 3) % 3;
        ck = lo[2] + s / 9;
      }
      bool ok = ci <= hi[0] && cj <= hi[1] && ck <= hi[2];
      int64_t c = ((int64_t)ck * g.ny + cj) * g.nx + ci;
      cur[s] = ok ? cstart[c] : 0u;
      end[s] = ok ? cend[c] : 0u;
    }
    while (true) {
      uint32_t best = 0xffffffffu;
      int bs = -1;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (cur[s] < end[s]) {
          uint32_t r = p.rank[cur[s]];
          if (r < best) {
            best = r;
            bs = s;
          }
        }
      }
      if (bs < 0) break;
      uint32_t t = 0;
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (s == bs) {
          t = cur[s];
          cur[s] = t + 1;
        }
      visit(t);
    }
  }
  out[f] = wsum > T(1e-12) ? acc / wsum : T(0);
}

template <class T>
__global__ void k_add_const(T* a, int64_t n, T c) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] += c;
}

// ------------------------------------------------------------------ enforce_solid_faces (pressure.rs:243-281)
template <class T, int COMP>
__global__ void k_enforce_solid(GridD<T> g, T* __restrict__ a, const uint8_t* __restrict__ cells) {
  const int fnx = g.nx + (COMP == 0), fny = g.ny + (COMP == 1), fnz = g.nz + (COMP == 2);
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (int64_t)fnx * fny * fnz) return;
  int i = (int)(f % fnx), j = (int)((f / fnx) % fny), k = (int)(f / ((int64_t)fnx * fny));
  auto solid = [&](int a_, int b_, int c_) {
    if (a_ < 0 || b_ < 0 || c_ < 0 || a_ >= g.nx || b_ >= g.ny || c_ >= g.nz) return true;
    return cells[((int64_t)c_ * g.ny + b_) * g.nx + a_] == kSolid;
  };
  bool s = COMP == 0 ? (solid(i - 1, j, k) || solid(i, j, k))
                     : (COMP == 1 ? (solid(i, j - 1, k) || solid(i, j, k))
                                  : (solid(i, j, k - 1) || solid(i, j, k)));
  if (s) a[f] = T(0);
}

// ------------------------------------------------------------------ RHS + Laplacian flags
// pressure.rs:152-223 with the volume-control source of solver.rs:392-423, written in
// the sheared solve layout.  One thread per sheared slot; slots outside the grid get
// b = 0 and flags = 0.
template <class T, class I>
__global__ void __launch_bounds__(256) k_rhs_flags(GridD<T> g, Shear sh, CFacesD<T> f,
                                                   const uint8_t* __restrict__ cells,
                                                   const uint32_t* __restrict__ count,
                                                   const double* __restrict__ srcarr, double inv_scale,
                                                   int use_source, double rest, double k_vol, double unit,
                                                   double cap, T* __restrict__ b, uint16_t* __restrict__ flags,
                                                   double* __restrict__ b64) {
  // I: the index type (32-bit when the sheared slots number below 2^31)
  const I t = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
  if (t >= (I)sh.size()) return;
  const I rowi = t / (I)sh.nx;
  int i = (int)(t - rowi * (I)sh.nx);
  int J = (int)(rowi % (I)sh.NJ);
  int k = (int)(rowi / (I)sh.NJ);
  int j = J - i;
  if (j < 0 || j >= sh.ny) {
    b[t] = T(0);
    flags[t] = 0;
    if (b64) b64[t] = 0.0;
    return;
  }
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  auto cell = [&](int a_, int b_, int c_) -> uint8_t {
    if (a_ < 0 || b_ < 0 || c_ < 0 || a_ >= nx || b_ >= ny || c_ >= nz) return kSolid;
    return cells[((int64_t)c_ * ny + b_) * nx + a_];
  };
  int64_t c = ((int64_t)k * ny + j) * nx + i;
  if (cells[c] != kFluid) {
    b[t] = T(0);
    flags[t] = 0;
    if (b64) b64[t] = 0.0;
    return;
  }
  uint16_t fl = kFlFluid;
  int diag = 0;
  const int nb[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k},
                        {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
  const uint16_t bit[6] = {kFlPx, kFlMx, kFlPy, kFlMy, kFlPz, kFlMz};
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    uint8_t kind = cell(nb[d][0], nb[d][1], nb[d][2]);
    if (kind != kSolid) ++diag;
    if (kind == kFluid) fl |= bit[d];
  }
  fl |= (uint16_t)(diag << 8);
  const int64_t ux = nx + 1;
  T du = f.u[((int64_t)k * ny + j) * ux + i + 1] - f.u[((int64_t)k * ny + j) * ux + i];
  T dv = f.v[((int64_t)k * (ny + 1) + j + 1) * nx + i] - f.v[((int64_t)k * (ny + 1) + j) * nx + i];
  T dw = f.w[((int64_t)(k + 1) * ny + j) * nx + i] - f.w[((int64_t)k * ny + j) * nx + i];
  T div = du + dv + dw;
  T s = T(0);
  if (srcarr) {
    s = (T)srcarr[c];
  } else if (use_source) {
    double over = (double)count[c] / rest - 1.0;
    if (over > 0.0) s = (T)fmin(k_vol * over * unit, cap);
  }
  b[t] = T(-inv_scale) * (div - s);
  flags[t] = fl;
  if (b64) {
    // the same right-hand side formed in f64 from the stored faces (refinement path)
    const double du64 = (double)f.u[((int64_t)k * ny + j) * ux + i + 1] - (double)f.u[((int64_t)k * ny + j) * ux + i];
    const double dv64 = (double)f.v[((int64_t)k * (ny + 1) + j + 1) * nx + i] -
                        (double)f.v[((int64_t)k * (ny + 1) + j) * nx + i];
    const double dw64 = (double)f.w[((int64_t)(k + 1) * ny + j) * nx + i] - (double)f.w[((int64_t)k * ny + j) * nx + i];
    double s64 = 0.0;
    if (srcarr) {
      s64 = srcarr[c];
    } else if (use_source) {
      const double over = (double)count[c] / rest - 1.0;
      if (over > 0.0) s64 = fmin(k_vol * over * unit, cap);
    }
    b64[t] = -inv_scale * (((du64 + dv64) + dw64) - s64);
  }
}

// Calls f(t, i, J, k) for every in-grid slot t of the sheared layout.  The in-grid slots
// of a row (k, J) are the contiguous run i in [max(0, J-ny+1), min(nx-1, J)], so a
// wavefront walks one row's run and the slots outside the grid (half of the layout) are
// never touched; rows are strided over the grid's wavefronts.
template <class F>
__device__ inline void for_cells(Shear sh, F&& f) {
  const int lane = threadIdx.x & 63;
  const int nrows = sh.nz * sh.NJ;
  const int wstride = (int)(((int64_t)gridDim.x * blockDim.x) >> 6);
  int row = __builtin_amdgcn_readfirstlane((int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  for (; row < nrows; row += wstride) {
    const int J = row % sh.NJ;
    const int k = row / sh.NJ;
    const int imin = J - sh.ny + 1 > 0 ? J - sh.ny + 1 : 0;
    const int imax = J < sh.nx - 1 ? J : sh.nx - 1;
    const int64_t base = (int64_t)row * sh.nx;
    for (int i = imin + lane; i <= imax; i += 64) f(base + i, i, J, k);
  }
}

// The in-grid run of sheared row `row` (= k * NJ + J): slots base + i, i in [imin, imin + len).
struct RowRun {
  int64_t base;
  int imin, len, J, k;
};
__device__ inline RowRun row_run(Shear sh, int row) {
  RowRun r;
  r.J = row % sh.NJ;
  r.k = row / sh.NJ;
  r.imin = r.J - sh.ny + 1 > 0 ? r.J - sh.ny + 1 : 0;
  const int imax = r.J < sh.nx - 1 ? r.J : sh.nx - 1;
  r.len = imax - r.imin + 1;
  r.base = (int64_t)row * sh.nx;
  return r;
}
// Rows are taken two at a time per wavefront (row and row + stride), so each lane has
// two independent cells' loads in flight: f(a_ok, a_run, a_i, b_ok, b_run, b_i).
template <class F>
__device__ inline void for_row_pairs(Shear sh, F&& f) {
  const int lane = threadIdx.x & 63;
  const int nrows = sh.nz * sh.NJ;
  const int wstride = (int)(((int64_t)gridDim.x * blockDim.x) >> 6);
  int row = __builtin_amdgcn_readfirstlane((int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  for (; row < nrows; row += 2 * wstride) {
    const RowRun a = row_run(sh, row);
    RowRun b = row_run(sh, row + wstride < nrows ? row + wstride : row);
    if (row + wstride >= nrows) b.len = 0;
    const int n = a.len > b.len ? a.len : b.len;
    for (int c = lane; c < n; c += 64) f(c < a.len, a, a.imin + c, c < b.len, b, b.imin + c);
  }
}

// max of the residual norms, propagating NaN (f64::max in the reference ignores NaN, so
// a NaN residual could pass for converged there); equal to fmax for numbers
__device__ inline double nmax(double a, double b) {
  return (a != a || b != b) ? __longlong_as_double(0x7ff8000000000000ll) : fmax(a, b);
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o));
  return v;
}

template <int OPMAX>
__device__ inline void block_partial(double v, double* partials) {
  __shared__ double red[4];
  v = OPMAX ? wave_max(v) : wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    partials[blockIdx.x] = OPMAX ? nmax(nmax(nmax(red[0], red[1]), red[2]), red[3])
                                 : ((red[0] + red[1]) + red[2]) + red[3];
}
// The sum of n partials in k_dot_finish's fast order (lane-strided sums over 256
// threads, a butterfly per wave, the four waves in order), by every thread of a
// 256-thread workgroup, each of which gets it.
__device__ inline double block_combine_sum(const double* partials, int n) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) acc += partials[q];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}
// kOpDenom of the fast mode (pressure.rs:372-373), combined by every workgroup of the
// axpy in k_dot_finish's order: denom = dot(s, As), alpha = sigma / denom with the sigma
// the fused product left in sigma_next, which workgroup 0 commits.  False (for every
// workgroup alike) when the solve stops here: a breakdown (|denom| < 1e-30) or a NaN.
__device__ inline bool denom_step(PcgScalars* sc, const double* den_partials, int nden, double* alpha) {
  const double den = block_combine_sum(den_partials, nden);
  const double sig = sc->sigma_next;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (den != den) {
    if (lead) sc->nan_seen = 1, sc->done = 1;
    return false;
  }
  if (lead) sc->sigma = sig, sc->denom = den;
  if (fabs(den) < 1e-30) {
    if (lead) sc->done = 1;
    return false;
  }
  *alpha = sig / den;
  if (lead) sc->alpha = *alpha;
  return true;
}

// ------------------------------------------------------------------ apply_a (pressure.rs:400-439)
// Matrix-free: diag and the -1 couplings come from the flags.  In-bounds neighbours
// without coupling contribute (+0 * x) exactly as the reference's stored 0.0 does.
template <class T>
__global__ void __launch_bounds__(256) k_spmv(Shear sh, const uint16_t* __restrict__ flags,
                                              const T* __restrict__ x, T* __restrict__ out,
                                              const PcgScalars* sc, double* partials) {
  if (sc->done) return;
  // one more iteration of a running solve (the count a graph replay cannot pass in)
  if (blockIdx.x == 0 && threadIdx.x == 0) const_cast<PcgScalars*>(sc)->iters += 1;
  const int64_t row = sh.nx, plane = (int64_t)sh.nx * sh.NJ;
  double acc = 0.0;  // dot(x, A x) of the stored values (fast reduction mode)
  auto cell = [&](int64_t t, int i, int J, int k, uint16_t f) -> T {
    T v = T(0);
    if (f & kFlFluid) {
      const int j = J - i;
      v = T(fl_diag(f)) * x[t];
      // (i+1, j) -> (i+1, J+1); (i, j+1) -> (i, J+1); (i, j, k+1) -> next plane.
      if (i + 1 < sh.nx) v += ((f & kFlPx) ? T(-1) : T(0)) * x[t + row + 1];
      if (i >= 1) v += ((f & kFlMx) ? T(-1) : T(0)) * x[t - row - 1];
      if (j + 1 < sh.ny) v += ((f & kFlPy) ? T(-1) : T(0)) * x[t + row];
      if (j >= 1) v += ((f & kFlMy) ? T(-1) : T(0)) * x[t - row];
      if (k + 1 < sh.nz) v += ((f & kFlPz) ? T(-1) : T(0)) * x[t + plane];
      if (k >= 1) v += ((f & kFlMz) ? T(-1) : T(0)) * x[t - plane];
      acc += (double)x[t] * (double)v;
    }
    return v;
  };
  for_row_pairs(sh, [&](bool ao, const RowRun& a, int ai, bool bo, const RowRun& b, int bi) {
    const int64_t ta = a.base + ai, tb = b.base + bi;
    const uint16_t fa = ao ? flags[ta] : 0, fb = bo ? flags[tb] : 0;
    const T va = cell(ta, ai, a.J, a.k, fa), vb = cell(tb, bi, b.J, b.k, fb);
    if (ao) out[ta] = va;
    if (bo) out[tb] = vb;
  });
  if (partials) block_partial<0>(acc, partials);
}

// The same product over work units of 64 consecutive slots of a sheared row (unit q is
// chunk q % nchunk of row q / nchunk).  Each wavefront takes a contiguous span of units,
// and the spans of the workgroups that one XCD runs are contiguous too (workgroup b runs
// on XCD b % 8), so the rows that a cell's z neighbours live in are mostly read by the
// same XCD and its L2.  U units are in flight per wavefront step: the flags and all
// seven x values of a lane are loaded before any is used (a neighbour outside the grid,
// and every slot outside it, is addressed past the buffer: no load).  The arithmetic and
// its order are those of k_spmv.
// Work units of the vector kernels: unit q is the 64 slots of a chunk of a sheared row;
// with the solver's unit table (Shear::units) only the chunks that hold in-grid slots
// are units (q = k * units_per_plane + r, table entry r gives row and chunk; a third of
// the chunks of a cube's rows hold none), else chunk q % nchunk of row q / nchunk.
// unit_span gives the calling wavefront its contiguous span [q0, q1) (the spans of
// workgroups on one XCD being contiguous too) and unit_slot the slot of this lane in
// unit qq, or kFarElem outside the grid or the span.
struct UnitLane {
  uint32_t t;  // slot, or kFarElem
  int i, j, k;
  bool v;
};
__host__ __device__ inline int64_t unit_count(Shear sh) {
  return sh.units ? (int64_t)sh.nz * sh.units_per_plane : (int64_t)sh.nz * sh.NJ * ((sh.nx + 63) / 64);
}
__device__ inline void unit_span(Shear sh, int span, int* q0, int* q1) {
  const int nunits = (int)unit_count(sh);
  const int G = gridDim.x, b = blockIdx.x;
  const int lb = (G & 7) == 0 ? (b & 7) * (G >> 3) + (b >> 3) : b;
  const int gw = lb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  *q0 = gw * span;
  *q1 = *q0 + span < nunits ? *q0 + span : nunits;
}
__device__ inline UnitLane unit_slot(Shear sh, int qq, int q1) {
  UnitLane L;
  int r, J;
  if (sh.units) {
    // wave-uniform: the table entry comes through the scalar cache
    const int qc = qq < q1 ? qq : q1 - 1;
    L.k = qc / sh.units_per_plane;
    const uint32_t e = sh.units[qc - L.k * sh.units_per_plane];
    J = (int)(e >> 8);
    L.i = (int)(e & 255u) * 64 + (threadIdx.x & 63);
    r = L.k * sh.NJ + J;
  } else {
    const int nchunk = (sh.nx + 63) >> 6;
    r = qq / nchunk;
    L.i = (qq - r * nchunk) * 64 + (threadIdx.x & 63);
    J = r % sh.NJ;
    L.k = r / sh.NJ;
  }
  L.j = J - L.i;
  L.v = qq < q1 && L.i < sh.nx && (unsigned)L.j < (unsigned)sh.ny;
  L.t = L.v ? (uint32_t)r * (uint32_t)sh.nx + (uint32_t)L.i : kFarElem;
  return L;
}
// unit kernels address with 32-bit byte offsets
template <class T>
bool units_fit(Shear sh) {
  const int64_t nunits = unit_count(sh);
  const int64_t plane = (int64_t)sh.nx * sh.NJ;
  // byte offsets of kFarElem +- a plane stay past the records (kFarElem * 4 = 2^30)
  return (sh.size() + 4 * plane) * (int64_t)sizeof(T) < ((int64_t)1 << 30) && nunits < ((int64_t)1 << 30);
}
// span of whole U-unit steps over nblocks * 4 wavefronts
inline int unit_span_len(Shear sh, int nblocks, int U) {
  const int64_t nunits = unit_count(sh);
  const int64_t waves = (int64_t)nblocks * 4;
  return (int)(((nunits + waves - 1) / waves + U - 1) / U * U);
}

template <class T, int U>
__global__ void __launch_bounds__(256) k_spmv_units(Shear sh, const uint16_t* __restrict__ flags,
                                                    const T* __restrict__ x, T* __restrict__ out,
                                                    const PcgScalars* sc, double* partials,
                                                    int span) {
  if (sc->done) return;
  // one more iteration of a running solve (the count a graph replay cannot pass in)
  if (blockIdx.x == 0 && threadIdx.x == 0) const_cast<PcgScalars*>(sc)->iters += 1;
  constexpr uint32_t SZ = sizeof(T);
  const uint32_t n = (uint32_t)sh.size();
  const rsrc_t rf = make_rsrc(flags, n * 2u), rx = make_rsrc(x, n * SZ), ro = make_rsrc(out, n * SZ);
  const int nx = sh.nx, ny = sh.ny, nz = sh.nz;
  const uint32_t row = (uint32_t)nx, plane = (uint32_t)nx * (uint32_t)sh.NJ;
  int q0, q1;
  unit_span(sh, span, &q0, &q1);
  double acc = 0.0;
  for (int q = q0; q < q1; q += U) {
    UnitLane L[U];
    uint16_t f[U];
    T xc[U], xn[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      L[u] = unit_slot(sh, q + u, q1);
      const uint32_t t = L[u].t;
      const int i = L[u].i, j = L[u].j, k = L[u].k;
      const bool v = L[u].v;
      // Neighbour slots without bounds tests: a neighbour outside the grid is still a
      // slot of the buffer, or past either end of it (an offset that wraps below 0 or
      // passes the record count: the hardware returns 0), and the arithmetic below drops
      // it.  Lanes outside the grid start from kFarElem, whose neighbours stay past the
      // record count (units_fit keeps the buffer below 2^30 - 4 planes bytes).
      (void)v;
      const uint32_t on[6] = {t + row + 1, t - row - 1, t + row, t - row, t + plane, t - plane};
      f[u] = (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rf, t * 2u, 0, 0);
      xc[u] = bld<T>(rx, t * SZ);
#pragma unroll
      for (int d = 0; d < 6; ++d) xn[u][d] = bld<T>(rx, on[d] * SZ);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint16_t fu = f[u];
      const int i = L[u].i, j = L[u].j, k = L[u].k;
      T val = T(fl_diag(fu)) * xc[u];
      // a neighbour inside the grid adds (-1 or +0) * x, one outside it adds nothing
      const T a0 = val + ((fu & kFlPx) ? T(-1) : T(0)) * xn[u][0];
      val = i + 1 < nx ? a0 : val;
      const T a1 = val + ((fu & kFlMx) ? T(-1) : T(0)) * xn[u][1];
      val = i >= 1 ? a1 : val;
      const T a2 = val + ((fu & kFlPy) ? T(-1) : T(0)) * xn[u][2];
      val = j + 1 < ny ? a2 : val;
      const T a3 = val + ((fu & kFlMy) ? T(-1) : T(0)) * xn[u][3];
      val = j >= 1 ? a3 : val;
      const T a4 = val + ((fu & kFlPz) ? T(-1) : T(0)) * xn[u][4];
      val = k + 1 < nz ? a4 : val;
      const T a5 = val + ((fu & kFlMz) ? T(-1) : T(0)) * xn[u][5];
      val = k >= 1 ? a5 : val;
      const bool fl = (fu & kFlFluid) != 0;
      const T res = fl ? val : T(0);
      acc += fl ? (double)xc[u] * (double)res : 0.0;
      bst<T>(ro, L[u].t * SZ, res);
    }
  }
  if (partials) block_partial<0>(acc, partials);
}

// The direction update s = z + beta s (pressure.rs:386-388; s = z on the first step,
// pressure.rs:367) fused with apply_a of the new s and the partials of dot(s, As)
// (pressure.rs:372-373).  A workgroup owns a tile of 64 values of i by kR rows J over a
// chunk of planes and marches up the planes: each plane's tile plus a one-slot halo of
// the new s (formed from z and the old s as it is staged) goes into a ring of LDS planes,
// so z, the old s and the flags are read once per cell (plus the halo) instead of once
// per neighbour, and the new s is stored once, by the tile that owns the cell.  The old
// and new s live in different buffers, since tiles read the old s of their neighbours'
// cells.  The arithmetic and its order are those of k_update_s and k_spmv.
//
// Every access goes through a descriptor of one plane of its buffer (no records for a
// plane outside the grid) at a 32-bit offset, kOob where the slot is outside the grid or
// not the thread's to store: no bounds branches and no 64-bit address arithmetic.
constexpr int kR = 16;
struct Tile25 {
  int nti, ntj, ntk, kc;
};
inline Tile25 tile25(Shear sh) {
  Tile25 t;
  t.kc = sh.nz >= 256 ? 32 : sh.nz >= 128 ? 16 : 8;
  t.nti = (sh.nx + 63) / 64;
  t.ntj = (sh.NJ + kR - 1) / kR;
  t.ntk = (sh.nz + t.kc - 1) / t.kc;
  return t;
}
template <class T>
__global__ void __launch_bounds__(256) k_spmv_update(Shear sh, const uint16_t* __restrict__ flags,
                                                     const T* __restrict__ z, T* s0, T* s1, T* __restrict__ out,
                                                     PcgScalars* sc, double* partials, const double* dot_partials,
                                                     int ndot, Tile25 tl) {
  if (sc->done) return;
  constexpr int NS = ((kR + 2) * 66 + 255) / 256;  // staged slots per thread
  constexpr int NC = kR / 4;                       // owned cells per thread and plane
  constexpr uint32_t SZ = sizeof(T);
  // four plane slots: the staging of plane k + 1 writes the slot plane k - 3 used, which no
  // thread reads after the barrier of plane k - 1, so one barrier per plane suffices
  __shared__ T ring[4][kR + 2][66];
  const int par = sc->spar;
  const bool copy = sc->started == 0;
  // kOpSigmaNew of the step before (pressure.rs:383-385), combined here by every
  // workgroup in the same order: beta = dot(z, r) / sigma; the new sigma waits in
  // sigma_next for the axpy (this kernel's workgroups still read the old one)
  const double sig_new = copy ? sc->sigma : block_combine_sum(dot_partials, ndot);
  if (sig_new != sig_new) {  // a NaN reaches every dot: stop, report it
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->nan_seen = 1, sc->done = 1;
    return;
  }
  const double beta_d = copy ? 0.0 : sig_new / sc->sigma;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->iters += 1;  // one more iteration of a running solve
    if (!copy) sc->beta = beta_d;
    sc->sigma_next = sig_new;
  }
  const T* sin = par ? s1 : s0;
  T* sout = par ? s0 : s1;
  const T beta = (T)beta_d;
  const int nx = sh.nx, ny = sh.ny, nz = sh.nz, NJ = sh.NJ;
  const int64_t plane = (int64_t)NJ * nx;
  // plane k of a buffer of esz-byte elements; no records unless k is in [klo, khi)
  auto prs = [&](const void* base, int k, int klo, int khi, uint32_t esz) {
    const bool in = k >= klo && k < khi;
    return make_rsrc((const char*)base + (in ? (int64_t)k * plane * esz : 0), in ? (uint32_t)(plane * esz) : 0u);
  };
  // workgroup b runs on XCD b % 8: each XCD gets a contiguous range of tiles (i fastest,
  // then J, then the plane chunk), so the halo rows and planes it re-reads are its own
  const int G = gridDim.x, b = blockIdx.x;
  const int lb = (b & 7) * (G >> 3) + (b >> 3);
  const int ti = lb % tl.nti, tj = (lb / tl.nti) % tl.ntj, tk = lb / (tl.nti * tl.ntj);
  const int i0 = ti * 64, J0 = tj * kR, k0 = tk * tl.kc;
  const int k1 = k0 + tl.kc < nz ? k0 + tl.kc : nz;
  const int ilast = i0 + 63 < nx - 1 ? i0 + 63 : nx - 1;
  // a tile with no in-grid slot (j = J - i outside [0, ny) on every row) does nothing
  const bool live = tk < tl.ntk && i0 < nx && J0 + kR - 1 >= i0 && J0 - ilast <= ny - 1;
  double acc = 0.0;
  if (live) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = i0 + lane;
    // staged slots of this thread: load offset (kOob outside the grid) and store offset
    // (kOob unless the tile owns the cell); the ring position is threadIdx.x + 256 q
    uint32_t lo[NS], so[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int e = threadIdx.x + 256 * q;
      const int rr = e / 66, cc = e - rr * 66;
      const int J = J0 - 1 + rr, ii = i0 - 1 + cc, jj = J - ii;
      const bool in = e < (kR + 2) * 66 && (unsigned)ii < (unsigned)nx && (unsigned)jj < (unsigned)ny;
      const bool own = rr >= 1 && rr <= kR && cc >= 1 && cc <= 64;
      lo[q] = in ? (uint32_t)(J * nx + ii) * SZ : kOob;
      so[q] = in && own ? lo[q] : kOob;
    }
    // owned cells: in-plane index (kFarElem outside the grid) and j
    uint32_t ci[NC];
    int cj[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const int J = J0 + w + 4 * m, jj = J - i;
      cj[m] = jj;
      ci[m] = i < nx && (unsigned)jj < (unsigned)ny ? (uint32_t)(J * nx + i) : kFarElem;
    }
    T zr[NS], sr[NS];
    uint16_t fr[NC];
    // issue the loads of plane k: z, the old s, and the owned cells' flags
    auto load_plane = [&](int k) {
      const rsrc_t rz = prs(z, k, 0, nz, SZ), rs = prs(sin, k, 0, copy ? 0 : nz, SZ);
      const rsrc_t rf = prs(flags, k, k0, k1, 2);
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        zr[q] = bld<T>(rz, lo[q]);
        sr[q] = copy ? T(0) : bld<T>(rs, lo[q]);
      }
#pragma unroll
      for (int m = 0; m < NC; ++m) fr[m] = (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rf, ci[m] * 2u, 0, 0);
    };
    // the new s of plane k into ring slot sl; the cells the tile owns are stored
    auto stage = [&](int k, int sl) {
      const rsrc_t rso = prs(sout, k, k0, k1, SZ);
      T* rs = &ring[sl][0][0];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const T v = copy ? zr[q] : zr[q] + beta * sr[q];
        if (q < NS - 1 || threadIdx.x + 256 * q < (kR + 2) * 66) rs[threadIdx.x + 256 * q] = v;
        bst<T>(rso, so[q], v);
      }
    };
    uint16_t fcur[NC];
    load_plane(k0 - 1);
    stage(k0 - 1, 0);
    load_plane(k0);
    stage(k0, 1);
#pragma unroll
    for (int m = 0; m < NC; ++m) fcur[m] = fr[m];
    load_plane(k0 + 1);
    for (int k = k0; k < k1; ++k) {
      const int sm = (k - k0) & 3, s0_ = (k - k0 + 1) & 3, sp = (k - k0 + 2) & 3;
      stage(k + 1, sp);
      uint16_t fn[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) fn[m] = fr[m];
      if (k + 1 < k1) load_plane(k + 2);  // in flight under this plane's product
      __syncthreads();
      const rsrc_t ro = prs(out, k, 0, nz, SZ);
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        // cells outside the grid have flags 0 (no fluid): they add nothing and store nothing
        const int rr = 1 + w + 4 * m, j = cj[m];
        const uint16_t fu = fcur[m];
        const int cc = lane + 1;
        const T xc = ring[s0_][rr][cc];
        T val = T(fl_diag(fu)) * xc;
        // a neighbour inside the grid adds (-1 or +0) * s, one outside it adds nothing
        const T a0 = val + ((fu & kFlPx) ? T(-1) : T(0)) * ring[s0_][rr + 1][cc + 1];
        val = i + 1 < nx ? a0 : val;
        const T a1 = val + ((fu & kFlMx) ? T(-1) : T(0)) * ring[s0_][rr - 1][cc - 1];
        val = i >= 1 ? a1 : val;
        const T a2 = val + ((fu & kFlPy) ? T(-1) : T(0)) * ring[s0_][rr + 1][cc];
        val = j + 1 < ny ? a2 : val;
        const T a3 = val + ((fu & kFlMy) ? T(-1) : T(0)) * ring[s0_][rr - 1][cc];
        val = j >= 1 ? a3 : val;
        const T a4 = val + ((fu & kFlPz) ? T(-1) : T(0)) * ring[sp][rr][cc];
        val = k + 1 < nz ? a4 : val;
        const T a5 = val + ((fu & kFlMz) ? T(-1) : T(0)) * ring[sm][rr][cc];
        val = k >= 1 ? a5 : val;
        const bool fl = (fu & kFlFluid) != 0;
        const T res = fl ? val : T(0);
        acc += fl ? (double)xc * (double)res : 0.0;
        bst<T>(ro, ci[m] * SZ, res);
      }
#pragma unroll
      for (int m = 0; m < NC; ++m) fcur[m] = fn[m];
    }
  }
  block_partial<0>(acc, partials);
}

// ------------------------------------------------------------------ dot (pressure.rs:588-603)
// Exact mode: one workgroup per 4096-element chunk of the reference cell order.  The
// products are formed in parallel into LDS, then one lane folds them left to right,
// acc = ((0 + x0*y0) + x1*y1) + ..., exactly as the reference chunk loop; the fold is a
// chain of dependent adds, so the LDS loads run ahead of it.
template <class T>
__global__ void __launch_bounds__(256) k_dot_exact(Shear sh, const T* __restrict__ a,
                                                   const T* __restrict__ b, double* partials,
                                                   const PcgScalars* sc) {
  if (sc->done) return;
  __shared__ double prod[4096];
  const int64_t n = (int64_t)sh.nx * sh.ny * sh.nz;
  const int64_t c0 = (int64_t)blockIdx.x * 4096;
  const int live = (int)(n - c0 < 4096 ? n - c0 : 4096);
  for (int l = threadIdx.x; l < live; l += 256) {
    const int64_t r = c0 + l;
    const int i = (int)(r % sh.nx);
    const int j = (int)((r / sh.nx) % sh.ny);
    const int k = (int)(r / ((int64_t)sh.nx * sh.ny));
    const int64_t s = sh.at(i, j, k);
    prod[l] = (double)a[s] * (double)b[s];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double acc = 0.0;
#pragma unroll 16
  for (int l = 0; l < live; ++l) acc += prod[l];
  partials[blockIdx.x] = acc;
}

// Fast mode: deterministic fixed tree in f64 (per-thread strided sum, then a wavefront
// butterfly, then the waves of the block in order).

template <class T>
__global__ void __launch_bounds__(256) k_dot_fast(Shear sh, const T* __restrict__ a,
                                                  const T* __restrict__ b, double* partials,
                                                  const PcgScalars* sc) {
  if (sc->done) return;
  double acc = 0.0;
  for_cells(sh, [&](int64_t t, int, int, int) { acc += (double)a[t] * (double)b[t]; });
  block_partial<0>(acc, partials);
}

// Combine partials and apply the scalar step of the PCG loop (pressure.rs:345-389).
__global__ void __launch_bounds__(256) k_dot_finish(double* partials, int nblocks, int exact, int op,
                                                    PcgScalars* sc, double tol, int rel) {
  if (sc->done) return;
  __shared__ double red[4];
  double v;
  if (op == kOpResidual || op == kOpRhsNorm) {
    double m = 0.0;
    for (int q = threadIdx.x; q < nblocks; q += 256) m = nmax(m, partials[q]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    v = nmax(nmax(nmax(red[0], red[1]), red[2]), red[3]);
  } else if (exact) {
    // Iterator::sum over the chunk partials, in chunk order.
    if (threadIdx.x != 0) return;
    v = -0.0;
    for (int q = 0; q < nblocks; ++q) v += partials[q];
  } else {
    double acc = 0.0;
    for (int q = threadIdx.x; q < nblocks; q += 256) acc += partials[q];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    v = ((red[0] + red[1]) + red[2]) + red[3];
  }
  if (threadIdx.x != 0) return;
  if (v != v) {  // a NaN in the vectors reaches every dot and max: stop, report it
    sc->nan_seen = 1;
    sc->done = 1;
    return;
  }
  switch (op) {
    case kOpSigma0:
      sc->sigma = v;
      break;
    case kOpDenom:
      sc->denom = v;
      if (fabs(v) < 1e-30) {
        sc->done = 1;
      } else {
        sc->alpha = sc->sigma / v;
      }
      break;
    case kOpSigmaNew:
      sc->beta = v / sc->sigma;
      sc->sigma = v;
      break;
    case kOpResidual:
      sc->residual = v;
      // the step's direction (written by the fused product) is the current one now
      sc->spar ^= 1;
      sc->started = 1;
      if (v <= sc->tol_eff) sc->done = 1;
      break;
    case kOpRhsNorm:
      sc->residual = v;
      sc->rhs_norm = v;
      sc->tol_eff = rel ? tol * v : tol;
      if (v <= sc->tol_eff) sc->done = 1;
      break;
  }
}

// inf_norm (pressure.rs:605-609): max |x| ignoring NaN (f64::max semantics).
template <class T>
__global__ void __launch_bounds__(256) k_maxabs(Shear sh, const T* __restrict__ a, double* partials,
                                                const PcgScalars* sc) {
  if (sc->done) return;
  double m = 0.0;
  for_cells(sh, [&](int64_t t, int, int, int) { m = nmax(m, fabs((double)a[t])); });
  block_partial<1>(m, partials);
}

// s = z + beta s (pressure.rs:386-388).
template <class T>
__global__ void __launch_bounds__(256) k_update_s(Shear sh, T* __restrict__ s, const T* __restrict__ z,
                                                  const PcgScalars* sc) {
  if (sc->done) return;
  const T beta = (T)sc->beta;
  for_row_pairs(sh, [&](bool ao, const RowRun& a, int ai, bool bo, const RowRun& b, int bi) {
    const int64_t ta = a.base + ai, tb = b.base + bi;
    T za = T(0), sa = T(0), zb = T(0), sb = T(0);
    if (ao) za = z[ta], sa = s[ta];
    if (bo) zb = z[tb], sb = s[tb];
    if (ao) s[ta] = za + beta * sa;
    if (bo) s[tb] = zb + beta * sb;
  });
}

// p += alpha s ; r -= alpha As (pressure.rs:371-375) with max|r| partials, two rows per
// wavefront step (the loads of both cells are issued before either is stored).
template <class T>
__global__ void __launch_bounds__(256) k_axpy_pr2(Shear sh, T* __restrict__ p, T* __restrict__ r,
                                                  const T* s0, const T* __restrict__ as,
                                                  PcgScalars* sc, double* partials,
                                                  const T* s1 = nullptr, const double* den_partials = nullptr,
                                                  int nden = 0) {
  if (sc->done) return;
  // s1: s is the direction the fused product just wrote (sbuf[1 - spar]), and the
  // product's partials of dot(s, As) are combined here (denom_step)
  const T* __restrict__ s = s1 && !sc->spar ? s1 : s0;
  double alpha_d = sc->alpha;
  if (den_partials && !denom_step(sc, den_partials, nden, &alpha_d)) return;
  const T alpha = (T)alpha_d;
  double m = 0.0;
  for_row_pairs(sh, [&](bool ao, const RowRun& a, int ai, bool bo, const RowRun& b, int bi) {
    const int64_t ta = a.base + ai, tb = b.base + bi;
    T pa = T(0), sa = T(0), ra = T(0), qa = T(0), pb = T(0), sb = T(0), rb = T(0), qb = T(0);
    if (ao) pa = p[ta], sa = s[ta], ra = r[ta], qa = as[ta];
    if (bo) pb = p[tb], sb = s[tb], rb = r[tb], qb = as[tb];
    if (ao) {
      p[ta] = pa + alpha * sa;
      const T rv = ra - alpha * qa;
      r[ta] = rv;
      m = nmax(m, fabs((double)rv));
    }
    if (bo) {
      p[tb] = pb + alpha * sb;
      const T rv = rb - alpha * qb;
      r[tb] = rv;
      m = nmax(m, fabs((double)rv));
    }
  });
  block_partial<1>(m, partials);
}

// k_update_s and k_axpy_pr2 over work units (the addressing of k_spmv_units): the
// loads of U units are in flight before the first store.
template <class T, int U>
__global__ void __launch_bounds__(256) k_update_s_units(Shear sh, T* __restrict__ s, const T* __restrict__ z,
                                                        const PcgScalars* sc, int span) {
  if (sc->done) return;
  constexpr uint32_t SZ = sizeof(T);
  const uint32_t n = (uint32_t)sh.size();
  const rsrc_t rs = make_rsrc(s, n * SZ), rz = make_rsrc(z, n * SZ);
  const T beta = (T)sc->beta;
  int q0, q1;
  unit_span(sh, span, &q0, &q1);
  for (int q = q0; q < q1; q += U) {
    uint32_t t[U];
    T zv[U], sv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      t[u] = unit_slot(sh, q + u, q1).t;
      zv[u] = bld<T>(rz, t[u] * SZ);
      sv[u] = bld<T>(rs, t[u] * SZ);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) bst<T>(rs, t[u] * SZ, zv[u] + beta * sv[u]);
  }
}

template <class T, int U>
__global__ void __launch_bounds__(256) k_axpy_pr_units(Shear sh, T* __restrict__ p, T* __restrict__ r,
                                                       const T* __restrict__ s, const T* __restrict__ as,
                                                       const PcgScalars* sc, double* partials, int span) {
  if (sc->done) return;
  constexpr uint32_t SZ = sizeof(T);
  const uint32_t n = (uint32_t)sh.size();
  const rsrc_t rp = make_rsrc(p, n * SZ), rr = make_rsrc(r, n * SZ), rs = make_rsrc(s, n * SZ),
               rq = make_rsrc(as, n * SZ);
  const T alpha = (T)sc->alpha;
  double m = 0.0;
  int q0, q1;
  unit_span(sh, span, &q0, &q1);
  for (int q = q0; q < q1; q += U) {
    uint32_t t[U];
    T pv[U], sv[U], rv[U], qv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      t[u] = unit_slot(sh, q + u, q1).t;
      pv[u] = bld<T>(rp, t[u] * SZ);
      sv[u] = bld<T>(rs, t[u] * SZ);
      rv[u] = bld<T>(rr, t[u] * SZ);
      qv[u] = bld<T>(rq, t[u] * SZ);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bst<T>(rp, t[u] * SZ, pv[u] + alpha * sv[u]);
      const T nr = rv[u] - alpha * qv[u];
      bst<T>(rr, t[u] * SZ, nr);
      m = nmax(m, fabs((double)nr));  // 0 outside the grid
    }
  }
  block_partial<1>(m, partials);
}

// k_axpy_pr_units with s the direction the fused product just wrote (sbuf[1 - spar]).
template <class T, int U>
__global__ void __launch_bounds__(256) k_axpy_pr_par(Shear sh, T* __restrict__ p, T* __restrict__ r,
                                                     const T* s0, const T* s1, const T* __restrict__ as,
                                                     PcgScalars* sc, double* partials, const double* den_partials,
                                                     int nden, int span) {
  if (sc->done) return;
  constexpr uint32_t SZ = sizeof(T);
  const uint32_t n = (uint32_t)sh.size();
  const T* s = sc->spar ? s0 : s1;
  const rsrc_t rp = make_rsrc(p, n * SZ), rr = make_rsrc(r, n * SZ), rs = make_rsrc(s, n * SZ),
               rq = make_rsrc(as, n * SZ);
  double alpha_d;
  if (!denom_step(sc, den_partials, nden, &alpha_d)) return;
  const T alpha = (T)alpha_d;
  double m = 0.0;
  int q0, q1;
  unit_span(sh, span, &q0, &q1);
  for (int q = q0; q < q1; q += U) {
    uint32_t t[U];
    T pv[U], sv[U], rv[U], qv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      t[u] = unit_slot(sh, q + u, q1).t;
      pv[u] = bld<T>(rp, t[u] * SZ);
      sv[u] = bld<T>(rs, t[u] * SZ);
      rv[u] = bld<T>(rr, t[u] * SZ);
      qv[u] = bld<T>(rq, t[u] * SZ);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bst<T>(rp, t[u] * SZ, pv[u] + alpha * sv[u]);
      const T nr = rv[u] - alpha * qv[u];
      bst<T>(rr, t[u] * SZ, nr);
      m = nmax(m, fabs((double)nr));  // 0 outside the grid
    }
  }
  block_partial<1>(m, partials);
}

// ------------------------------------------------------------------ MIC(0) (pressure.rs:442-570)
// One launch per hyperplane sigma = J + k of the sheared layout: every cell of the
// plane reads only cells of earlier planes, in the reference's operand order.
//   build: e = diag - tx^2 - ty^2 - tz^2 - tau*(mx + my + mz); precon = 1/sqrt(e)
//   fwd:   q = (((r - Tx) - Ty) - Tz) * precon, T = (plus[cm]*precon[cm])*q[cm]
//   bwd:   z = (((q - Ux) - Uy) - Uz) * precon, U = (plus[c]*precon[c])*z[cp]
// Non-coupled in-grid neighbours contribute (0*0)*0 = +0 in the reference, and
// subtracting +0 is the identity, so they are skipped here.
template <class T, int OP>
__global__ void k_mic_plane(Shear sh, const uint16_t* __restrict__ flags, T* __restrict__ precon,
                            const T* __restrict__ in, T* __restrict__ out, int sigma,
                            const PcgScalars* sc) {
  if (sc->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (OP == kMicBwd) k = sh.nz - 1 - k;
  if (i >= sh.nx) return;
  int J = OP == kMicBwd ? (sh.NJ - 1) - (sigma - (sh.nz - 1 - k)) : sigma - k;
  if (J < 0 || J >= sh.NJ) return;
  int j = J - i;
  if (j < 0 || j >= sh.ny) return;
  const int64_t row = sh.nx, plane = (int64_t)sh.nx * sh.NJ;
  const int64_t t = ((int64_t)k * sh.NJ + J) * sh.nx + i;
  const uint16_t f = flags[t];
  if (!(f & kFlFluid)) return;  // precon / q / z stay 0 (zero-initialised)
  if (OP == kMicBuild) {
    const double TAU = 0.97, SIG = 0.25;
    double ax = 0, px = 0, axy = 0, axz = 0;
    double ay = 0, py = 0, ayx = 0, ayz = 0;
    double az = 0, pz = 0, azx = 0, azy = 0;
    if (i >= 1) {
      int64_t m = t - row - 1;
      uint16_t g = flags[m];
      ax = (f & kFlMx) ? -1.0 : 0.0;
      px = precon[m];
      axy = (g & kFlPy) ? -1.0 : 0.0;
      axz = (g & kFlPz) ? -1.0 : 0.0;
    }
    if (j >= 1) {
      int64_t m = t - row;
      uint16_t g = flags[m];
      ay = (f & kFlMy) ? -1.0 : 0.0;
      py = precon[m];
      ayx = (g & kFlPx) ? -1.0 : 0.0;
      ayz = (g & kFlPz) ? -1.0 : 0.0;
    }
    if (k >= 1) {
      int64_t m = t - plane;
      uint16_t g = flags[m];
      az = (f & kFlMz) ? -1.0 : 0.0;
      pz = precon[m];
      azx = (g & kFlPx) ? -1.0 : 0.0;
      azy = (g & kFlPy) ? -1.0 : 0.0;
    }
    double tx = ax * px, ty = ay * py, tz = az * pz;
    double diag = (double)fl_diag(f);
    double e = diag - tx * tx - ty * ty - tz * tz -
               TAU * (ax * (axy + axz) * px * px + ay * (ayx + ayz) * py * py +
                      az * (azx + azy) * pz * pz);
    if (e < SIG * diag) e = diag;
    precon[t] = e > 0.0 ? (T)(1.0 / sqrt(e)) : T(0);
  } else if (OP == kMicFwd) {
    T v = in[t];
    if (f & kFlMx) v -= (-precon[t - row - 1]) * out[t - row - 1];
    if (f & kFlMy) v -= (-precon[t - row]) * out[t - row];
    if (f & kFlMz) v -= (-precon[t - plane]) * out[t - plane];
    out[t] = v * precon[t];
  } else {
    const T pc = precon[t];
    T v = in[t];
    if (f & kFlPx) v -= (-pc) * out[t + row + 1];
    if (f & kFlPy) v -= (-pc) * out[t + row];
    if (f & kFlPz) v -= (-pc) * out[t + plane];
    out[t] = v * pc;
  }
}

// ------------------------------------------------------------------ apply_pressure_gradient (pressure.rs:285-332)
template <class T, int COMP>
__global__ void k_gradient(GridD<T> g, Shear sh, T* __restrict__ a, const uint8_t* __restrict__ cells,
                           const T* __restrict__ p, T coef) {
  const int fnx = g.nx + (COMP == 0), fny = g.ny + (COMP == 1), fnz = g.nz + (COMP == 2);
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (int64_t)fnx * fny * fnz) return;
  int i = (int)(f % fnx), j = (int)((f / fnx) % fny), k = (int)(f / ((int64_t)fnx * fny));
  int li = i - (COMP == 0), lj = j - (COMP == 1), lk = k - (COMP == 2);
  if (li < 0 || lj < 0 || lk < 0) return;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  uint8_t l = cells[((int64_t)lk * g.ny + lj) * g.nx + li];
  uint8_t r = cells[((int64_t)k * g.ny + j) * g.nx + i];
  if (l == kSolid || r == kSolid) return;
  if (l == kFluid || r == kFluid) a[f] -= coef * (p[sh.at(i, j, k)] - p[sh.at(li, lj, lk)]);
}

// The refinement path's gradient: the f64 pressure difference applied in f64 to the
// stored face, rounded once.
template <class T, int COMP>
__global__ void k_gradient64(GridD<T> g, Shear sh, T* __restrict__ a, const uint8_t* __restrict__ cells,
                             const double* __restrict__ p, double coef) {
  const int fnx = g.nx + (COMP == 0), fny = g.ny + (COMP == 1), fnz = g.nz + (COMP == 2);
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (int64_t)fnx * fny * fnz) return;
  int i = (int)(f % fnx), j = (int)((f / fnx) % fny), k = (int)(f / ((int64_t)fnx * fny));
  int li = i - (COMP == 0), lj = j - (COMP == 1), lk = k - (COMP == 2);
  if (li < 0 || lj < 0 || lk < 0) return;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  uint8_t l = cells[((int64_t)lk * g.ny + lj) * g.nx + li];
  uint8_t r = cells[((int64_t)k * g.ny + j) * g.nx + i];
  if (l == kSolid || r == kSolid) return;
  if (l == kFluid || r == kFluid)
    a[f] = (T)((double)a[f] - coef * (p[sh.at(i, j, k)] - p[sh.at(li, lj, lk)]));
}

template <class P, class T>
__global__ void k_unshear(Shear sh, const P* __restrict__ p, T* __restrict__ out) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)sh.nx * sh.ny * sh.nz) return;
  int i = (int)(c % sh.nx), j = (int)((c / sh.nx) % sh.ny), k = (int)(c / ((int64_t)sh.nx * sh.ny));
  out[c] = (T)p[sh.at(i, j, k)];
}

// ---- iterative refinement in f64 (f32 storage path)
// r = b - A p in f64 from the flags (the reference apply_a operand order), with max|r|
// partials per block.
__global__ void __launch_bounds__(256) k_residual64(Shear sh, const uint16_t* __restrict__ flags,
                                                    const double* __restrict__ b,
                                                    const double* __restrict__ x, double* __restrict__ r,
                                                    double* partials) {
  __shared__ double red[4];
  const int64_t row = sh.nx, plane = (int64_t)sh.nx * sh.NJ;
  double m = 0.0;
  // in-grid slots only: r's slots outside the grid hold +0 from its allocation on
  for_cells(sh, [&](int64_t t, int i, int J, int k) {
    const uint16_t f = flags[t];
    double v = 0.0;
    if (f & kFlFluid) {
      const int j = J - i;
      double ax = (double)fl_diag(f) * x[t];
      if (i + 1 < sh.nx) ax += ((f & kFlPx) ? -1.0 : 0.0) * x[t + row + 1];
      if (i >= 1) ax += ((f & kFlMx) ? -1.0 : 0.0) * x[t - row - 1];
      if (j + 1 < sh.ny) ax += ((f & kFlPy) ? -1.0 : 0.0) * x[t + row];
      if (j >= 1) ax += ((f & kFlMy) ? -1.0 : 0.0) * x[t - row];
      if (k + 1 < sh.nz) ax += ((f & kFlPz) ? -1.0 : 0.0) * x[t + plane];
      if (k >= 1) ax += ((f & kFlMz) ? -1.0 : 0.0) * x[t - plane];
      v = b[t] - ax;
    }
    r[t] = v;
    m = nmax(m, fabs(v));
  });
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = nmax(nmax(nmax(red[0], red[1]), red[2]), red[3]);
}

// k_residual64 over work units (the addressing of k_spmv_units: U units of 64 slots in
// flight per wave, 32-bit byte offsets, lanes outside the grid at kFarElem, whose loads
// read 0 and whose stores are dropped), the same arithmetic in the same order.
template <int U>
__global__ void __launch_bounds__(256) k_residual64_units(Shear sh, const uint16_t* __restrict__ flags,
                                                          const double* __restrict__ b,
                                                          const double* __restrict__ x, double* __restrict__ r,
                                                          double* partials, int span) {
  const uint32_t n = (uint32_t)sh.size();
  const rsrc_t rf = make_rsrc(flags, n * 2u), rx = make_rsrc(x, n * 8u), rb = make_rsrc(b, n * 8u),
               rr = make_rsrc(r, n * 8u);
  const int nx = sh.nx, ny = sh.ny, nz = sh.nz;
  const uint32_t row = (uint32_t)nx, plane = (uint32_t)nx * (uint32_t)sh.NJ;
  int q0, q1;
  unit_span(sh, span, &q0, &q1);
  double m = 0.0;
  for (int q = q0; q < q1; q += U) {
    UnitLane L[U];
    uint16_t f[U];
    double xc[U], bv[U], xn[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      L[u] = unit_slot(sh, q + u, q1);
      const uint32_t t = L[u].t;
      const uint32_t on[6] = {t + row + 1, t - row - 1, t + row, t - row, t + plane, t - plane};
      f[u] = (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rf, t * 2u, 0, 0);
      xc[u] = bld<double>(rx, t * 8u);
      bv[u] = bld<double>(rb, t * 8u);
#pragma unroll
      for (int d = 0; d < 6; ++d) xn[u][d] = bld<double>(rx, on[d] * 8u);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint16_t fu = f[u];
      const int i = L[u].i, j = L[u].j, k = L[u].k;
      double ax = (double)fl_diag(fu) * xc[u];
      const double a0 = ax + ((fu & kFlPx) ? -1.0 : 0.0) * xn[u][0];
      ax = i + 1 < nx ? a0 : ax;
      const double a1 = ax + ((fu & kFlMx) ? -1.0 : 0.0) * xn[u][1];
      ax = i >= 1 ? a1 : ax;
      const double a2 = ax + ((fu & kFlPy) ? -1.0 : 0.0) * xn[u][2];
      ax = j + 1 < ny ? a2 : ax;
      const double a3 = ax + ((fu & kFlMy) ? -1.0 : 0.0) * xn[u][3];
      ax = j >= 1 ? a3 : ax;
      const double a4 = ax + ((fu & kFlPz) ? -1.0 : 0.0) * xn[u][4];
      ax = k + 1 < nz ? a4 : ax;
      const double a5 = ax + ((fu & kFlMz) ? -1.0 : 0.0) * xn[u][5];
      ax = k >= 1 ? a5 : ax;
      const double v = (fu & kFlFluid) ? bv[u] - ax : 0.0;
      bst<double>(rr, L[u].t * 8u, v);
      m = nmax(m, fabs(v));
    }
  }
  block_partial<1>(m, partials);
}

// The elementwise steps of the reliable updates over work units (as above).
template <class T, int U>
__global__ void __launch_bounds__(256) k_scale_cast_units(Shear sh, const double* __restrict__ r, double inv,
                                                          T* __restrict__ b, int span) {
  const uint32_t n = (uint32_t)sh.size();
  const rsrc_t rr = make_rsrc(r, n * 8u), rb = make_rsrc(b, n * (uint32_t)sizeof(T));
  int q0, q1;
  unit_span(sh, span, &q0, &q1);
  for (int q = q0; q < q1; q += U) {
    uint32_t t[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      t[u] = unit_slot(sh, q + u, q1).t;
      v[u] = bld<double>(rr, t[u] * 8u);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) bst<T>(rb, t[u] * (uint32_t)sizeof(T), (T)(v[u] * inv));
  }
}
template <class T, int U>
__global__ void __launch_bounds__(256) k_scale_inplace_units(Shear sh, T* __restrict__ x, double f, int span) {
  const uint32_t n = (uint32_t)sh.size();
  const rsrc_t rx = make_rsrc(x, n * (uint32_t)sizeof(T));
  int q0, q1;
  unit_span(sh, span, &q0, &q1);
  for (int q = q0; q < q1; q += U) {
    uint32_t t[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      t[u] = unit_slot(sh, q + u, q1).t;
      v[u] = bld<T>(rx, t[u] * (uint32_t)sizeof(T));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) bst<T>(rx, t[u] * (uint32_t)sizeof(T), (T)((double)v[u] * f));
  }
}
template <class T, int U>
__global__ void __launch_bounds__(256) k_accum64_units(Shear sh, double* __restrict__ p, const T* __restrict__ e,
                                                       double scale, int span) {
  const uint32_t n = (uint32_t)sh.size();
  const rsrc_t rp = make_rsrc(p, n * 8u), re = make_rsrc(e, n * (uint32_t)sizeof(T));
  int q0, q1;
  unit_span(sh, span, &q0, &q1);
  for (int q = q0; q < q1; q += U) {
    uint32_t t[U];
    double pv[U];
    T ev[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      t[u] = unit_slot(sh, q + u, q1).t;
      pv[u] = bld<double>(rp, t[u] * 8u);
      ev[u] = bld<T>(re, t[u] * (uint32_t)sizeof(T));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) bst<double>(rp, t[u] * 8u, pv[u] + scale * (double)ev[u]);
  }
}

__global__ void __launch_bounds__(256) k_max_finish(const double* partials, int nblocks, double* out) {
  __shared__ double red[4];
  double m = 0.0;
  for (int q = threadIdx.x; q < nblocks; q += 256) m = nmax(m, partials[q]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = nmax(nmax(nmax(red[0], red[1]), red[2]), red[3]);
}

// rhs of a correction solve: b = r * inv (rounded to the storage type)
template <class T>
__global__ void k_scale_cast(Shear sh, const double* __restrict__ r, double inv, T* __restrict__ b) {
  for_cells(sh, [&](int64_t t, int, int, int) { b[t] = (T)(r[t] * inv); });
}

// x *= f (the search direction carried into new residual units)
template <class T>
__global__ void k_scale_inplace(Shear sh, T* __restrict__ x, double f) {
  for_cells(sh, [&](int64_t t, int, int, int) { x[t] = (T)((double)x[t] * f); });
}

// The PCG scalars after a residual replacement: sigma into the new residual units (or
// 1 for a restart, whose search direction has been cleared), the new stopping bound,
// and the loop running again.
__global__ void k_pcg_rescale(PcgScalars* sc, double f2, double tol_eff, int restart) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sc->sigma = restart ? 1.0 : sc->sigma * f2;
  sc->tol_eff = tol_eff;
  sc->done = 0;
}

// p += scale * e in f64
template <class T>
__global__ void k_accum64(Shear sh, double* __restrict__ p, const T* __restrict__ e, double scale) {
  for_cells(sh, [&](int64_t t, int, int, int) { p[t] += scale * (double)e[t]; });
}

// ------------------------------------------------------------------ extrapolate (solver.rs:608-647, 960-1009)
template <class T, int COMP, class I>
__global__ void __launch_bounds__(256) k_extrap_init(GridD<T> g, const uint8_t* __restrict__ cells,
                                                     uint8_t* known) {
  const int fnx = g.nx + (COMP == 0), fny = g.ny + (COMP == 1), fnz = g.nz + (COMP == 2);
  const I f = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
  if (f >= (I)fnx * (I)fny * (I)fnz) return;
  const I row = f / (I)fnx;
  const int i = (int)(f - row * (I)fnx), j = (int)(row % (I)fny), k = (int)(row / (I)fny);
  auto fluid = [&](int a_, int b_, int c_) {
    return cells[((int64_t)c_ * g.ny + b_) * g.nx + a_] == kFluid;
  };
  bool v;
  if (COMP == 0) v = (i > 0 && fluid(i - 1, j, k)) || (i < g.nx && fluid(i, j, k));
  else if (COMP == 1) v = (j > 0 && fluid(i, j - 1, k)) || (j < g.ny && fluid(i, j, k));
  else v = (k > 0 && fluid(i, j, k - 1)) || (k < g.nz && fluid(i, j, k));
  known[f] = v;
}

// I: the index type (32-bit when the faces number below 2^31: no 64-bit division)
template <class T, class I>
__global__ void __launch_bounds__(256) k_extrap_iter(int nx, int ny, int nz, const T* __restrict__ src,
                                                     T* __restrict__ dst, const uint8_t* __restrict__ ksrc,
                                                     uint8_t* __restrict__ kdst) {
  const I c = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
  if (c >= (I)nx * (I)ny * (I)nz) return;
  if (ksrc[c]) {
    dst[c] = src[c];
    kdst[c] = 1;
    return;
  }
  const I row = c / (I)nx;
  const int i = (int)(c - row * (I)nx), j = (int)(row % (I)ny), k = (int)(row / (I)ny);
  const I sy = (I)nx, sz = (I)nx * (I)ny;
  T sum = T(0);
  int cnt = 0;
  if (i > 0 && ksrc[c - 1]) { sum += src[c - 1]; ++cnt; }
  if (i + 1 < nx && ksrc[c + 1]) { sum += src[c + 1]; ++cnt; }
  if (j > 0 && ksrc[c - sy]) { sum += src[c - sy]; ++cnt; }
  if (j + 1 < ny && ksrc[c + sy]) { sum += src[c + sy]; ++cnt; }
  if (k > 0 && ksrc[c - sz]) { sum += src[c - sz]; ++cnt; }
  if (k + 1 < nz && ksrc[c + sz]) { sum += src[c + sz]; ++cnt; }
  if (cnt > 0) {
    dst[c] = sum / T(cnt);
    kdst[c] = 1;
  } else {
    dst[c] = src[c];
    kdst[c] = 0;
  }
}

// ------------------------------------------------------------------ G2P (solver.rs:650-677)
template <class T>
__global__ void __launch_bounds__(256) k_g2p(Parts<T> p, int64_t n, GridD<T> g, CFacesD<T> f, CFacesD<T> sv, T flip) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  V3<T> x = load_pos(p, t);
  // velocity_at(g, f, x) and the trilinear_diff samples from one set of corner loads
  V3<T> l = g.local(x);
  T pa, pb, pc, dvx, dvy, dvz;
  Corners<T> ku, kv, kw;
  trilinear_and_diff(f.u, sv.u, g.nx + 1, g.ny, g.nz, l.x, l.y - T(0.5), l.z - T(0.5), &pa, &dvx, &ku);
  trilinear_and_diff(f.v, sv.v, g.nx, g.ny + 1, g.nz, l.x - T(0.5), l.y, l.z - T(0.5), &pb, &dvy, &kv);
  trilinear_and_diff(f.w, sv.w, g.nx, g.ny, g.nz + 1, l.x - T(0.5), l.y - T(0.5), l.z, &pc, &dvz, &kw);
  V3<T> pic = mk3(pa, pb, pc);
  V3<T> fv = load_vel(p, t) + mk3(dvx, dvy, dvz);
  V3<T> nv = fv * flip + pic * (T(1) - flip);
  p.vx[t] = nv.x;
  p.vy[t] = nv.y;
  p.vz[t] = nv.z;
}

// ------------------------------------------------------------------ advect (solver.rs:693-711)
template <class T>
__global__ void __launch_bounds__(256) k_advect(Parts<T> p, int64_t n, GridD<T> g, CFacesD<T> f, SdfD<T> sdf, T dt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  V3<T> p0 = load_pos(p, t);
  V3<T> v0 = velocity_at(g, f, p0);
  V3<T> mid = p0 + v0 * (T(0.5) * dt);
  V3<T> vm = velocity_at(g, f, mid);
  V3<T> st = vm * dt;
  T len = len3(st);
  if (len > g.dx) st = st * (g.dx / len);
  V3<T> r = project_inside_or_stay(sdf, g.dx, p0 + st, p0);
  p.px[t] = r.x;
  p.py[t] = r.y;
  p.pz[t] = r.z;
}

// g2p then advect of the same particle in one pass (solver.rs:650-711): the PIC sample
// of g2p is advect's first RK2 sample (same grid, same position), so the particle is
// read once and that sample is taken once; every value is the one the two kernels give.
template <class T>
__global__ void __launch_bounds__(256) k_g2p_advect(Parts<T> p, int64_t n, GridD<T> g, CFacesD<T> f, CFacesD<T> sv, T flip,
                             SdfD<T> sdf, T dt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const V3<T> p0 = load_pos(p, t);
  // velocity_at(g, f, p0) and the trilinear_diff samples from one set of corner loads
  const V3<T> l = g.local(p0);
  T pa, pb, pc, dvx, dvy, dvz;
  Corners<T> ku, kv, kw;
  trilinear_and_diff(f.u, sv.u, g.nx + 1, g.ny, g.nz, l.x, l.y - T(0.5), l.z - T(0.5), &pa, &dvx, &ku);
  trilinear_and_diff(f.v, sv.v, g.nx, g.ny + 1, g.nz, l.x - T(0.5), l.y, l.z - T(0.5), &pb, &dvy, &kv);
  trilinear_and_diff(f.w, sv.w, g.nx, g.ny, g.nz + 1, l.x - T(0.5), l.y - T(0.5), l.z, &pc, &dvz, &kw);
  const V3<T> pic = mk3(pa, pb, pc);
  const V3<T> fv = load_vel(p, t) + mk3(dvx, dvy, dvz);
  const V3<T> nv = fv * flip + pic * (T(1) - flip);
  p.vx[t] = nv.x;
  p.vy[t] = nv.y;
  p.vz[t] = nv.z;
  const V3<T> mid = p0 + pic * (T(0.5) * dt);
  // velocity_at(g, f, mid), in f32 reusing the corners of the first sample where the
  // stencil base has not moved (in f64 the kept corners would spill)
  V3<T> vm;
  if constexpr (sizeof(T) == 4) {
    const V3<T> lm = g.local(mid);
    vm = mk3(trilinear_again(f.u, g.nx + 1, g.ny, g.nz, lm.x, lm.y - T(0.5), lm.z - T(0.5), ku),
             trilinear_again(f.v, g.nx, g.ny + 1, g.nz, lm.x - T(0.5), lm.y, lm.z - T(0.5), kv),
             trilinear_again(f.w, g.nx, g.ny, g.nz + 1, lm.x - T(0.5), lm.y - T(0.5), lm.z, kw));
  } else {
    vm = velocity_at(g, f, mid);
  }
  V3<T> st = vm * dt;
  const T len = len3(st);
  if (len > g.dx) st = st * (g.dx / len);
  const V3<T> r = project_inside_or_stay(sdf, g.dx, p0 + st, p0);
  p.px[t] = r.x;
  p.py[t] = r.y;
  p.pz[t] = r.z;
}

// ------------------------------------------------------------------ enforce_particle_bounds (solver.rs:736-755)
template <class T>
__global__ void k_bounds(Parts<T> p, int64_t n, SdfD<T> sdf, BoxD<T> outlet, uint32_t* keep,
                         uint32_t* cull, unsigned long long* counters) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  V3<T> q = load_pos(p, t);
  bool k_ = false;
  int why = 0;  // 1 drained, 2 culled
  if (!sdf.in_bounds(q)) {
    why = 2;
  } else if (outlet.contains(q)) {
    why = 1;
  } else {
    k_ = sdf.sample(q) < T(0.5) * sdf.dx;
    if (!k_) why = 2;
  }
  keep[t] = k_;
  cull[t] = !k_;
  if (why) atomicAdd(&counters[why - 1], 1ull);
}

// Compaction in place.  The slots' order is not the reference order (that is the rank,
// and every order-dependent stage reads ranks), so only the m removed particles' slots
// change hands: with kept = n - m, the removed slots below kept (holes) and the
// survivors at or above kept (movers) are equally many, and hole j takes mover j, both
// numbered in slot order (pos = exclusive scan of keep).  The removed ranks are listed
// for the rank fix.
template <class T>
__global__ void k_compact_holes(const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos, int64_t n,
                                int64_t kept, const uint32_t* __restrict__ rank, uint32_t* hole_slot,
                                uint32_t* culled_ranks) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || keep[t]) return;
  const uint32_t j = (uint32_t)t - pos[t];  // removed slots before t
  culled_ranks[j] = rank[t];
  if (t < kept) hole_slot[j] = (uint32_t)t;
}
template <class T>
__global__ void k_compact_move(Parts<T> p, const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos,
                               int64_t n, int64_t kept, const uint32_t* __restrict__ hole_slot) {
  const int64_t t = kept + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || !keep[t]) return;
  const uint32_t d = hole_slot[pos[t] - pos[kept]];  // survivors in [kept, t)
  p.px[d] = p.px[t];
  p.py[d] = p.py[t];
  p.pz[d] = p.pz[t];
  p.vx[d] = p.vx[t];
  p.vy[d] = p.vy[t];
  p.vz[d] = p.vz[t];
  p.rank[d] = p.rank[t];
}

// The reference compacts in place (particles.rs:38-49), so a survivor's index drops by
// the number of removed particles before it.
// The number of removed ranks below a survivor's rank, from a bitmap of the removed
// ranks and an exclusive scan of its words' population counts.
__global__ void k_zero_words(uint32_t* bits, int64_t nw) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw) bits[w] = 0;
}
__global__ void k_mark_ranks(const uint32_t* __restrict__ culled, int64_t m, uint32_t* bits) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const uint32_t r = culled[t];
  atomicOr(&bits[r >> 5], 1u << (r & 31));
}
__global__ void k_word_popc(const uint32_t* __restrict__ bits, int64_t nw, uint32_t* pop) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw) pop[w] = (uint32_t)__popc(bits[w]);
}
__global__ void k_rank_fix(uint32_t* rank, int64_t n, const uint32_t* __restrict__ bits,
                           const uint32_t* __restrict__ below) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t r = rank[t];
  const uint32_t w = r >> 5;
  rank[t] = r - (below[w] + (uint32_t)__popc(bits[w] & ((1u << (r & 31)) - 1u)));
}

// ------------------------------------------------------------------ redistribute (solver.rs:454-557)
template <class T>
__global__ void k_hist(Parts<T> p, int64_t n, GridD<T> g, int32_t* cellkey, uint32_t* count) {
  // Particles are still close to cell order (the last sort), so lanes of a wavefront
  // mostly share cells: each run of equal cells adds its length with one atomic.
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int c = -1;
  if (t < n) {
    c = (int)g.cell_of(load_pos(p, t));
    cellkey[t] = c;
  }
  const int prev = __shfl_up(c, 1);
  const bool head = lane == 0 || prev != c;
  const unsigned long long heads = __ballot(head);
  if (head && c >= 0) {
    const unsigned long long after = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    atomicAdd(&count[c], (uint32_t)(end - lane));
  }
}

__global__ void k_active_flags(const int32_t* __restrict__ cellkey, const uint32_t* __restrict__ count,
                               int64_t n, uint32_t cap, uint32_t* flag) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int32_t c = cellkey[t];
  flag[t] = (c >= 0 && count[c] > cap) ? 1u : 0u;
}

__global__ void k_active_compact(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                 int64_t n, const uint32_t* __restrict__ rank, uint32_t* act_rank,
                                 uint32_t* act_slot) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || !flag[t]) return;
  uint32_t q = pos[t];
  act_rank[q] = rank[t];
  act_slot[q] = (uint32_t)t;
}

// The order-dependent decision pass: visits the particles of over-cap cells in
// reference order and mutates `count` as it goes, exactly as the reference loop does.
// Particles outside over-cap cells can never move during a sweep (a receiver only
// rises to the cap), so this set is the whole sweep.  Single thread on purpose.
__global__ void k_redist_decide(const uint32_t* __restrict__ act, int64_t nact,
                                const int32_t* __restrict__ cellkey, uint32_t* count,
                                const uint8_t* __restrict__ mask, int nx, int ny, int nz,
                                uint32_t cap, double gx, double gy, double gz, int32_t* dest,
                                unsigned long long* moved) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  unsigned long long nmoved = 0;
  // dir . gdir for -x,+x,-y,+y,-z,+z (math.rs:42-44 with the unit dirs of solver.rs:492-509)
  double dn[6];
  dn[0] = -1.0 * gx + 0.0 * gy + 0.0 * gz;
  dn[1] = 1.0 * gx + 0.0 * gy + 0.0 * gz;
  dn[2] = 0.0 * gx + -1.0 * gy + 0.0 * gz;
  dn[3] = 0.0 * gx + 1.0 * gy + 0.0 * gz;
  dn[4] = 0.0 * gx + 0.0 * gy + -1.0 * gz;
  dn[5] = 0.0 * gx + 0.0 * gy + 1.0 * gz;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  for (int64_t a = 0; a < nact; ++a) {
    uint32_t slot = act[a];
    int64_t c = cellkey[slot];
    dest[a] = -1;
    if (count[c] <= cap) continue;
    int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sz);
    int64_t nb[6];
    double dd[6];
    int nn = 0;
    if (i > 0) { nb[nn] = c - 1; dd[nn++] = dn[0]; }
    if (i + 1 < nx) { nb[nn] = c + 1; dd[nn++] = dn[1]; }
    if (j > 0) { nb[nn] = c - sy; dd[nn++] = dn[2]; }
    if (j + 1 < ny) { nb[nn] = c + sy; dd[nn++] = dn[3]; }
    if (k > 0) { nb[nn] = c - sz; dd[nn++] = dn[4]; }
    if (k + 1 < nz) { nb[nn] = c + sz; dd[nn++] = dn[5]; }
    bool have = false;
    int64_t best = 0;
    double best_down = -INFINITY;
    uint32_t best_room = 0;
    for (int q = 0; q < nn; ++q) {
      int64_t nc = nb[q];
      if (!mask[nc] || count[nc] >= cap) continue;
      uint32_t room = cap - count[nc];
      bool better = !have || dd[q] > best_down || (dd[q] == best_down && room > best_room);
      if (better) {
        have = true;
        best = nc;
        best_down = dd[q];
        best_room = room;
      }
    }
    if (have) {
      count[c] -= 1;
      count[best] += 1;
      dest[a] = (int32_t)best;
      ++nmoved;
    }
  }
  *moved = nmoved;
}

// One atomic per wavefront for a per-lane flag (a single global counter hit by every
// lane serialises the whole grid on one address).
__device__ inline void wave_count(uint32_t* ctr, bool flag) {
  const unsigned long long m = __ballot(flag);
  const int lane = threadIdx.x & 63;
  if (m && lane == __ffsll((long long)m) - 1) atomicAdd(ctr, (uint32_t)__popcll(m));
}
__device__ inline void wave_count64(unsigned long long* ctr, bool flag) {
  const unsigned long long m = __ballot(flag);
  const int lane = threadIdx.x & 63;
  if (m && lane == __ffsll((long long)m) - 1) atomicAdd(ctr, (unsigned long long)__popcll(m));
}
// Slot of this lane in a list appended to by every lane with flag set (order within
// the wave by lane, between waves arbitrary).
__device__ inline uint32_t wave_append(uint32_t* ctr, bool flag) {
  const unsigned long long m = __ballot(flag);
  const int lane = threadIdx.x & 63;
  const int leader = m ? __ffsll((long long)m) - 1 : 0;
  uint32_t base = 0;
  if (m && lane == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

template <class T>
__global__ void __launch_bounds__(256) k_redist_apply(Parts<T> p, const uint32_t* __restrict__ act,
                               const int32_t* __restrict__ dest, int64_t nact, GridD<T> g,
                               CFacesD<T> f, SdfD<T> sdf, int32_t* cellkey, uint32_t* count) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nact) return;
  int32_t nc = dest[a];
  if (nc < 0) return;
  uint32_t slot = act[a];
  uint64_t key = p.rank[slot];
  int ci = nc % g.nx, cj = (nc / g.nx) % g.ny, ck = nc / (g.nx * g.ny);
  V3<T> jit = mk3((T)jitter_mix(key * 3), (T)jitter_mix(key * 3 + 1), (T)jitter_mix(key * 3 + 2));
  V3<T> q = g.cell_center(ci, cj, ck) + jit * g.dx;
  q = project_inside(sdf, g.dx, q);
  V3<T> v = velocity_at(g, f, q);
  p.px[slot] = q.x;
  p.py[slot] = q.y;
  p.pz[slot] = q.z;
  p.vx[slot] = v.x;
  p.vy[slot] = v.y;
  p.vz[slot] = v.z;
  if (count) {
    // The decision counted the particle in nc; the next sweep's histogram counts it in
    // the cell its projected position lies in (integer atomics: order-free).
    const int32_t c2 = (int32_t)g.cell_of(q);
    if (c2 != nc) {
      atomicSub(&count[nc], 1u);
      if (c2 >= 0) atomicAdd(&count[c2], 1u);
    }
    cellkey[slot] = c2;
  }
}

// ------------------------------------------------------------------ CFL (grid.rs:202-206)
template <class T>
__global__ void k_maxabs_faces(const T* __restrict__ a, int64_t n, unsigned long long* out) {
  __shared__ double red[4];
  double m = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs((double)a[t]));
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(fmax(red[0], red[1]), red[2]), red[3]);
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

// ------------------------------------------------------------------ host <-> device particle layout
template <class T>
__global__ void k_to_aos(Parts<T> p, int64_t n, double* pos, double* vel) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  uint32_t r = p.rank[t];
  if ((int64_t)r >= n) return;  // ranks are dense 0..n-1; never scatter outside the buffer
  pos[3 * (int64_t)r] = p.px[t];
  pos[3 * (int64_t)r + 1] = p.py[t];
  pos[3 * (int64_t)r + 2] = p.pz[t];
  if (!vel) return;  // positions only (frame output)
  vel[3 * (int64_t)r] = p.vx[t];
  vel[3 * (int64_t)r + 1] = p.vy[t];
  vel[3 * (int64_t)r + 2] = p.vz[t];
}

// Positions as f32 triples in reference order (`p.x as f32`, output.rs:104-106).
template <class T>
__global__ void k_to_aos_f32(Parts<T> p, int64_t n, float* pos) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  uint32_t r = p.rank[t];
  if ((int64_t)r >= n) return;
  pos[3 * (int64_t)r] = (float)p.px[t];
  pos[3 * (int64_t)r + 1] = (float)p.py[t];
  pos[3 * (int64_t)r + 2] = (float)p.pz[t];
}

template <class T>
__global__ void k_from_aos(Parts<T> p, int64_t n, int64_t base, const double* pos,
                           const double* vel) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int64_t d = base + t;
  p.px[d] = (T)pos[3 * t];
  p.py[d] = (T)pos[3 * t + 1];
  p.pz[d] = (T)pos[3 * t + 2];
  p.vx[d] = vel ? (T)vel[3 * t] : T(0);
  p.vy[d] = vel ? (T)vel[3 * t + 1] : T(0);
  p.vz[d] = vel ? (T)vel[3 * t + 2] : T(0);
  p.rank[d] = (uint32_t)d;
}

// Initial fill of the benchmark cavities: `per_cell` particles in each listed cell at
// centre + jitter, jitter from the same splitmix64 finaliser the reference uses for
// redistribution, keyed by (seed, cell, slot).
template <class T>
__global__ void k_seed_fill(Parts<T> p, int64_t base, GridD<T> g, const uint32_t* __restrict__ cells,
                            int64_t ncells, uint32_t per_cell, uint64_t seed, T vx, T vy, T vz) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * per_cell) return;
  int64_t q = t / per_cell;
  uint32_t c = cells[q];
  int ci = c % g.nx, cj = (c / g.nx) % g.ny, ck = c / (g.nx * g.ny);
  uint64_t key = seed * 0x100000000ull + (uint64_t)t;
  double jx = jitter_mix(key * 3) / 0.6 * 0.9999, jy = jitter_mix(key * 3 + 1) / 0.6 * 0.9999,
         jz = jitter_mix(key * 3 + 2) / 0.6 * 0.9999;
  V3<T> x = g.cell_center(ci, cj, ck) + mk3((T)jx, (T)jy, (T)jz) * g.dx;
  int64_t d = base + t;
  p.px[d] = x.x;
  p.py[d] = x.y;
  p.pz[d] = x.z;
  p.vx[d] = vx;
  p.vy[d] = vy;
  p.vz[d] = vz;
  p.rank[d] = (uint32_t)d;
}

template <class T>
__global__ void k_static_cls(GridD<T> g, const uint8_t* __restrict__ mask, BoxD<T> outlet,
                             uint8_t* out) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)g.nx * g.ny * g.nz) return;
  int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / ((int64_t)g.nx * g.ny));
  bool o = outlet.contains(g.cell_center(i, j, k));
  out[c] = o ? kAir : (!mask[c] ? kSolid : kAir);
}

// solver.rs:194-199: the static cavity mask from the SDF at each cell centre.
template <class T>
__global__ void k_cavity_mask(GridD<T> g, SdfD<T> sdf, uint8_t* mask) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)g.nx * g.ny * g.nz) return;
  int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / ((int64_t)g.nx * g.ny));
  mask[c] = sdf.sample(g.cell_center(i, j, k)) < T(0);
}

}  // namespace

// ================================================================== launch wrappers
template <class T>
void launch_gather_parts(const Parts<T>& s, const Parts<T>& d, const uint32_t* perm, int64_t n,
                         hipStream_t st) {
  if (n <= 0) return;
  k_gather_parts<T><<<grid_for(n), kBlock, 0, st>>>(s, d, perm, n);
}
template <class T>
void launch_classify(const uint8_t* static_cls, const Parts<T>& p, int64_t n, GridD<T> g, BoxD<T> outlet,
                     uint8_t* cells, uint32_t* count, hipStream_t st) {
  int64_t nc = (int64_t)g.nx * g.ny * g.nz;
  k_classify_init<<<grid_for(nc), kBlock, 0, st>>>(static_cls, nc, cells, count);
  if (n > 0) k_classify_parts<T><<<grid_for(n), kBlock, 0, st>>>(static_cls, p, n, g, outlet, cells, count);
}
template <class T>
void launch_p2g(int comp, GridD<T> g, const Parts<T>& p, const uint32_t* cstart,
                const uint32_t* cend, T* out, bool exact, hipStream_t st) {
  int64_t nf = (int64_t)(g.nx + (comp == 0)) * (g.ny + (comp == 1)) * (g.nz + (comp == 2));
  unsigned gr = grid_for(nf);
  if (exact) {
    if (comp == 0) k_p2g<T, 0, true><<<gr, kBlock, 0, st>>>(g, p, cstart, cend, out);
    if (comp == 1) k_p2g<T, 1, true><<<gr, kBlock, 0, st>>>(g, p, cstart, cend, out);
    if (comp == 2) k_p2g<T, 2, true><<<gr, kBlock, 0, st>>>(g, p, cstart, cend, out);
  } else {
    if (comp == 0) k_p2g<T, 0, false><<<gr, kBlock, 0, st>>>(g, p, cstart, cend, out);
    if (comp == 1) k_p2g<T, 1, false><<<gr, kBlock, 0, st>>>(g, p, cstart, cend, out);
    if (comp == 2) k_p2g<T, 2, false><<<gr, kBlock, 0, st>>>(g, p, cstart, cend, out);
  }
}
template <class T>
void launch_axpy_const(T* a, int64_t n, T c, hipStream_t st) {
  k_add_const<T><<<grid_for(n), kBlock, 0, st>>>(a, n, c);
}
template <class T>
void launch_enforce_solid(GridD<T> g, FacesD<T> f, const uint8_t* cells, hipStream_t st) {
  int64_t nu = (int64_t)(g.nx + 1) * g.ny * g.nz, nv = (int64_t)g.nx * (g.ny + 1) * g.nz,
          nw = (int64_t)g.nx * g.ny * (g.nz + 1);
  k_enforce_solid<T, 0><<<grid_for(nu), kBlock, 0, st>>>(g, f.u, cells);
  k_enforce_solid<T, 1><<<grid_for(nv), kBlock, 0, st>>>(g, f.v, cells);
  k_enforce_solid<T, 2><<<grid_for(nw), kBlock, 0, st>>>(g, f.w, cells);
}
template <class T>
void launch_rhs_flags(GridD<T> g, Shear sh, CFacesD<T> f, const uint8_t* cells,
                      const uint32_t* count, const double* srcarr, double inv_scale, int use_source,
                      double rest, double k_vol, double unit, double cap, T* b, uint16_t* flags,
                      double* b64, hipStream_t st) {
  if (sh.size() < ((int64_t)1 << 31))
    k_rhs_flags<T, uint32_t><<<grid_for(sh.size()), kBlock, 0, st>>>(g, sh, f, cells, count, srcarr, inv_scale,
                                                                      use_source, rest, k_vol, unit, cap, b,
                                                                      flags, b64);
  else
    k_rhs_flags<T, int64_t><<<grid_for(sh.size()), kBlock, 0, st>>>(g, sh, f, cells, count, srcarr, inv_scale,
                                                                     use_source, rest, k_vol, unit, cap, b,
                                                                     flags, b64);
}
int cell_blocks(Shear sh) {
  int64_t waves = (int64_t)sh.nz * sh.NJ;  // one wavefront per row at most
  int64_t b = (waves + 3) / 4;
  if (b > 2048) b = 2048;
  return (int)(b < 1 ? 1 : b);
}
template <class T>
void launch_spmv(Shear sh, const uint16_t* flags, const T* x, T* out, const PcgScalars* sc,
                 double* partials, hipStream_t st) {
  const int nb = cell_blocks(sh);
  if (units_fit<T>(sh)) {
    k_spmv_units<T, 4><<<nb, 256, 0, st>>>(sh, flags, x, out, sc, partials, unit_span_len(sh, nb, 4));
    return;
  }
  k_spmv<T><<<nb, 256, 0, st>>>(sh, flags, x, out, sc, partials);
}
int dot_blocks(Shear sh, bool exact) {
  if (exact) {
    int64_t n = (int64_t)sh.nx * sh.ny * sh.nz;
    return (int)((n + 4095) / 4096);
  }
  return cell_blocks(sh);
}
int maxabs_blocks(Shear sh) { return dot_blocks(sh, false); }
template <class T>
void launch_dot_partials(Shear sh, const T* a, const T* b, double* partials, int nblocks, bool exact,
                         const PcgScalars* sc, hipStream_t st) {
  if (exact)
    k_dot_exact<T><<<nblocks, 256, 0, st>>>(sh, a, b, partials, sc);
  else
    k_dot_fast<T><<<nblocks, 256, 0, st>>>(sh, a, b, partials, sc);
}
void launch_dot_finish(double* partials, int nblocks, bool exact, int op, PcgScalars* sc,
                       double tol, hipStream_t st, int rel) {
  k_dot_finish<<<1, 256, 0, st>>>(partials, nblocks, exact ? 1 : 0, op, sc, tol, rel);
}
template <class T>
void launch_maxabs_partials(Shear sh, const T* a, double* partials, int nblocks,
                            const PcgScalars* sc, hipStream_t st) {
  k_maxabs<T><<<nblocks, 256, 0, st>>>(sh, a, partials, sc);
}
template <class T>
void launch_axpy_pr(Shear sh, T* p, T* r, const T* s, const T* as, const PcgScalars* sc,
                    double* partials, int nblocks, hipStream_t st) {
  if (units_fit<T>(sh)) {
    k_axpy_pr_units<T, 4><<<nblocks, 256, 0, st>>>(sh, p, r, s, as, sc, partials, unit_span_len(sh, nblocks, 4));
    return;
  }
  k_axpy_pr2<T><<<nblocks, 256, 0, st>>>(sh, p, r, s, as, const_cast<PcgScalars*>(sc), partials);  // reads sc only
}
template <class T>
void launch_update_s(Shear sh, T* s, const T* z, const PcgScalars* sc, hipStream_t st) {
  const int nb = cell_blocks(sh);
  if (units_fit<T>(sh)) {
    k_update_s_units<T, 4><<<nb, 256, 0, st>>>(sh, s, z, sc, unit_span_len(sh, nb, 4));
    return;
  }
  k_update_s<T><<<nb, 256, 0, st>>>(sh, s, z, sc);
}
int spmv_update_blocks(Shear sh) {
  const Tile25 t = tile25(sh);
  const int64_t g = (int64_t)t.nti * t.ntj * t.ntk;
  return (int)((g + 7) / 8 * 8);
}
template <class T>
void launch_spmv_update(Shear sh, const uint16_t* flags, const T* z, T* s0, T* s1, T* out, PcgScalars* sc,
                        double* partials, const double* dot_partials, int ndot, hipStream_t st) {
  k_spmv_update<T><<<spmv_update_blocks(sh), 256, 0, st>>>(sh, flags, z, s0, s1, out, sc, partials, dot_partials,
                                                           ndot, tile25(sh));
}
template <class T>
void launch_axpy_pr_par(Shear sh, T* p, T* r, const T* s0, const T* s1, const T* as, PcgScalars* sc,
                        double* partials, const double* den_partials, int nden, int nblocks, hipStream_t st) {
  if (units_fit<T>(sh)) {
    k_axpy_pr_par<T, 4><<<nblocks, 256, 0, st>>>(sh, p, r, s0, s1, as, sc, partials, den_partials, nden,
                                                  unit_span_len(sh, nblocks, 4));
    return;
  }
  k_axpy_pr2<T><<<nblocks, 256, 0, st>>>(sh, p, r, s0, as, sc, partials, s1, den_partials, nden);
}
template <class T>
void launch_mic_plane(int op, Shear sh, const uint16_t* flags, T* precon, const T* in, T* out,
                      int sigma, const PcgScalars* sc, hipStream_t st) {
  dim3 gr((unsigned)((sh.nx + 63) / 64), (unsigned)sh.nz);
  if (op == kMicBuild) k_mic_plane<T, kMicBuild><<<gr, 64, 0, st>>>(sh, flags, precon, in, out, sigma, sc);
  if (op == kMicFwd) k_mic_plane<T, kMicFwd><<<gr, 64, 0, st>>>(sh, flags, precon, in, out, sigma, sc);
  if (op == kMicBwd) k_mic_plane<T, kMicBwd><<<gr, 64, 0, st>>>(sh, flags, precon, in, out, sigma, sc);
}
template <class T>
void launch_gradient(GridD<T> g, Shear sh, FacesD<T> f, const uint8_t* cells, const T* p, T coef,
                     hipStream_t st) {
  int64_t nu = (int64_t)(g.nx + 1) * g.ny * g.nz, nv = (int64_t)g.nx * (g.ny + 1) * g.nz,
          nw = (int64_t)g.nx * g.ny * (g.nz + 1);
  k_gradient<T, 0><<<grid_for(nu), kBlock, 0, st>>>(g, sh, f.u, cells, p, coef);
  k_gradient<T, 1><<<grid_for(nv), kBlock, 0, st>>>(g, sh, f.v, cells, p, coef);
  k_gradient<T, 2><<<grid_for(nw), kBlock, 0, st>>>(g, sh, f.w, cells, p, coef);
}
template <class T>
void launch_unshear(Shear sh, const T* p, T* out, hipStream_t st) {
  k_unshear<T, T><<<grid_for((int64_t)sh.nx * sh.ny * sh.nz), kBlock, 0, st>>>(sh, p, out);
}
template <class T>
void launch_gradient64(GridD<T> g, Shear sh, FacesD<T> f, const uint8_t* cells, const double* p,
                       double coef, hipStream_t st) {
  int64_t nu = (int64_t)(g.nx + 1) * g.ny * g.nz, nv = (int64_t)g.nx * (g.ny + 1) * g.nz,
          nw = (int64_t)g.nx * g.ny * (g.nz + 1);
  k_gradient64<T, 0><<<grid_for(nu), kBlock, 0, st>>>(g, sh, f.u, cells, p, coef);
  k_gradient64<T, 1><<<grid_for(nv), kBlock, 0, st>>>(g, sh, f.v, cells, p, coef);
  k_gradient64<T, 2><<<grid_for(nw), kBlock, 0, st>>>(g, sh, f.w, cells, p, coef);
}
template <class T>
void launch_unshear64(Shear sh, const double* p, T* out, hipStream_t st) {
  k_unshear<double, T><<<grid_for((int64_t)sh.nx * sh.ny * sh.nz), kBlock, 0, st>>>(sh, p, out);
}
void launch_residual64(Shear sh, const uint16_t* flags, const double* b, const double* x, double* r,
                       double* partials, int nblocks, double* out, hipStream_t st) {
  if (units_fit<double>(sh))
    k_residual64_units<4><<<nblocks, 256, 0, st>>>(sh, flags, b, x, r, partials, unit_span_len(sh, nblocks, 4));
  else
    k_residual64<<<nblocks, 256, 0, st>>>(sh, flags, b, x, r, partials);
  k_max_finish<<<1, 256, 0, st>>>(partials, nblocks, out);
}
template <class T>
void launch_scale_cast(Shear sh, const double* r, double inv, T* b, hipStream_t st) {
  const int nb = cell_blocks(sh);
  if (units_fit<double>(sh))
    k_scale_cast_units<T, 4><<<nb, 256, 0, st>>>(sh, r, inv, b, unit_span_len(sh, nb, 4));
  else
    k_scale_cast<T><<<nb, kBlock, 0, st>>>(sh, r, inv, b);
}
template <class T>
void launch_scale_inplace(Shear sh, T* x, double f, hipStream_t st) {
  const int nb = cell_blocks(sh);
  if (units_fit<T>(sh))
    k_scale_inplace_units<T, 4><<<nb, 256, 0, st>>>(sh, x, f, unit_span_len(sh, nb, 4));
  else
    k_scale_inplace<T><<<nb, kBlock, 0, st>>>(sh, x, f);
}
void launch_pcg_rescale(PcgScalars* sc, double f2, double tol_eff, int restart, hipStream_t st) {
  k_pcg_rescale<<<1, 64, 0, st>>>(sc, f2, tol_eff, restart);
}
template <class T>
void launch_accum64(Shear sh, double* p, const T* e, double scale, hipStream_t st) {
  const int nb = cell_blocks(sh);
  if (units_fit<double>(sh))
    k_accum64_units<T, 4><<<nb, 256, 0, st>>>(sh, p, e, scale, unit_span_len(sh, nb, 4));
  else
    k_accum64<T><<<nb, kBlock, 0, st>>>(sh, p, e, scale);
}
template <class T>
void launch_extrap_init(GridD<T> g, int comp, const uint8_t* cells, uint8_t* known, hipStream_t st) {
  int64_t nf = (int64_t)(g.nx + (comp == 0)) * (g.ny + (comp == 1)) * (g.nz + (comp == 2));
  if (nf < ((int64_t)1 << 31)) {
    if (comp == 0) k_extrap_init<T, 0, uint32_t><<<grid_for(nf), kBlock, 0, st>>>(g, cells, known);
    if (comp == 1) k_extrap_init<T, 1, uint32_t><<<grid_for(nf), kBlock, 0, st>>>(g, cells, known);
    if (comp == 2) k_extrap_init<T, 2, uint32_t><<<grid_for(nf), kBlock, 0, st>>>(g, cells, known);
    return;
  }
  if (comp == 0) k_extrap_init<T, 0, int64_t><<<grid_for(nf), kBlock, 0, st>>>(g, cells, known);
  if (comp == 1) k_extrap_init<T, 1, int64_t><<<grid_for(nf), kBlock, 0, st>>>(g, cells, known);
  if (comp == 2) k_extrap_init<T, 2, int64_t><<<grid_for(nf), kBlock, 0, st>>>(g, cells, known);
}
template <class T>
void launch_extrap_iter(int nx, int ny, int nz, const T* src, T* dst, const uint8_t* ksrc,
                        uint8_t* kdst, hipStream_t st) {
  const int64_t n = (int64_t)nx * ny * nz;
  if (n < ((int64_t)1 << 31))
    k_extrap_iter<T, uint32_t><<<grid_for(n), kBlock, 0, st>>>(nx, ny, nz, src, dst, ksrc, kdst);
  else
    k_extrap_iter<T, int64_t><<<grid_for(n), kBlock, 0, st>>>(nx, ny, nz, src, dst, ksrc, kdst);
}
template <class T>
void launch_g2p(const Parts<T>& p, int64_t n, GridD<T> g, CFacesD<T> f, CFacesD<T> saved, T flip,
                hipStream_t st) {
  if (n <= 0) return;
  k_g2p<T><<<grid_for(n), kBlock, 0, st>>>(p, n, g, f, saved, flip);
}
template <class T>
void launch_advect(const Parts<T>& p, int64_t n, GridD<T> g, CFacesD<T> f, SdfD<T> sdf, T dt,
                   hipStream_t st) {
  if (n <= 0) return;
  k_advect<T><<<grid_for(n), kBlock, 0, st>>>(p, n, g, f, sdf, dt);
}
template <class T>
void launch_g2p_advect(const Parts<T>& p, int64_t n, GridD<T> g, CFacesD<T> f, CFacesD<T> saved, T flip,
                       SdfD<T> sdf, T dt, hipStream_t st) {
  if (n <= 0) return;
  k_g2p_advect<T><<<grid_for(n), kBlock, 0, st>>>(p, n, g, f, saved, flip, sdf, dt);
}
template <class T>
void launch_bounds(const Parts<T>& p, int64_t n, SdfD<T> sdf, BoxD<T> outlet, uint32_t* keep,
                   uint32_t* cull, unsigned long long* counters, hipStream_t st) {
  if (n <= 0) return;
  k_bounds<T><<<grid_for(n), kBlock, 0, st>>>(p, n, sdf, outlet, keep, cull, counters);
}
template <class T>
void launch_compact(const Parts<T>& p, const uint32_t* keep, const uint32_t* pos, int64_t n, int64_t kept,
                    uint32_t* hole_slot, uint32_t* culled_ranks, hipStream_t st) {
  if (n <= 0 || kept >= n) return;
  k_compact_holes<T><<<grid_for(n), kBlock, 0, st>>>(keep, pos, n, kept, p.rank, hole_slot, culled_ranks);
  k_compact_move<T><<<grid_for(n - kept), kBlock, 0, st>>>(p, keep, pos, n, kept, hole_slot);
}
void launch_rank_bits(const uint32_t* culled, int64_t m, int64_t nranks, uint32_t* bits, uint32_t* pop,
                      hipStream_t st) {
  const int64_t nw = (nranks + 31) / 32;
  k_zero_words<<<grid_for(nw), kBlock, 0, st>>>(bits, nw);
  if (m > 0) k_mark_ranks<<<grid_for(m), kBlock, 0, st>>>(culled, m, bits);
  k_word_popc<<<grid_for(nw), kBlock, 0, st>>>(bits, nw, pop);
}
void launch_rank_fix(uint32_t* rank, int64_t n, const uint32_t* bits, const uint32_t* below, hipStream_t st) {
  if (n <= 0) return;
  k_rank_fix<<<grid_for(n), kBlock, 0, st>>>(rank, n, bits, below);
}
template <class T>
void launch_hist(const Parts<T>& p, int64_t n, GridD<T> g, int32_t* cellkey, uint32_t* count,
                 hipStream_t st) {
  if (n <= 0) return;
  k_hist<T><<<grid_for(n), kBlock, 0, st>>>(p, n, g, cellkey, count);
}
void launch_active_flags(const int32_t* cellkey, const uint32_t* count, int64_t n, uint32_t cap,
                         uint32_t* flag, hipStream_t st) {
  if (n <= 0) return;
  k_active_flags<<<grid_for(n), kBlock, 0, st>>>(cellkey, count, n, cap, flag);
}
void launch_active_compact(const uint32_t* flag, const uint32_t* pos, int64_t n,
                           const uint32_t* rank, uint32_t* act_rank, uint32_t* act_slot,
                           hipStream_t st) {
  if (n <= 0) return;
  k_active_compact<<<grid_for(n), kBlock, 0, st>>>(flag, pos, n, rank, act_rank, act_slot);
}
void launch_redist_decide(const uint32_t* act, int64_t nact, const int32_t* cellkey,
                          uint32_t* count, const uint8_t* mask, int nx, int ny, int nz, uint32_t cap,
                          double gx, double gy, double gz, int32_t* dest, unsigned long long* moved,
                          hipStream_t st) {
  if (nact <= 0) return;
  k_redist_decide<<<1, 1, 0, st>>>(act, nact, cellkey, count, mask, nx, ny, nz, cap, gx, gy, gz,
                                   dest, moved);
}
template <class T>
void launch_redist_apply(const Parts<T>& p, const uint32_t* act, const int32_t* dest, int64_t nact,
                         GridD<T> g, CFacesD<T> f, SdfD<T> sdf, int32_t* cellkey, uint32_t* count,
                         hipStream_t st) {
  if (nact <= 0) return;
  k_redist_apply<T><<<grid_for(nact), kBlock, 0, st>>>(p, act, dest, nact, g, f, sdf, cellkey, count);
}
template <class T>
void launch_maxabs_faces(const T* a, int64_t n, unsigned long long* out, hipStream_t st) {
  int64_t b = (n + 256 * 8 - 1) / (256 * 8);
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  k_maxabs_faces<T><<<(unsigned)b, 256, 0, st>>>(a, n, out);
}
template <class T>
void launch_to_aos(const Parts<T>& p, int64_t n, double* pos, double* vel, hipStream_t st) {
  if (n <= 0) return;
  k_to_aos<T><<<grid_for(n), kBlock, 0, st>>>(p, n, pos, vel);
}
template <class T>
void launch_to_aos_f32(const Parts<T>& p, int64_t n, float* pos, hipStream_t st) {
  if (n <= 0) return;
  k_to_aos_f32<T><<<grid_for(n), kBlock, 0, st>>>(p, n, pos);
}
template <class T>
void launch_from_aos(const Parts<T>& p, int64_t n, int64_t base, const double* pos,
                     const double* vel, hipStream_t st) {
  if (n <= 0) return;
  k_from_aos<T><<<grid_for(n), kBlock, 0, st>>>(p, n, base, pos, vel);
}
template <class T>
void launch_seed_fill(const Parts<T>& p, int64_t base, GridD<T> g, const uint8_t* /*mask*/,
                      const uint32_t* cells, int64_t ncells, uint32_t per_cell, uint64_t seed, T vx,
                      T vy, T vz, hipStream_t st) {
  int64_t n = ncells * per_cell;
  if (n <= 0) return;
  k_seed_fill<T><<<grid_for(n), kBlock, 0, st>>>(p, base, g, cells, ncells, per_cell, seed, vx, vy,
                                                  vz);
}
template <class T>
void launch_static_cls(GridD<T> g, const uint8_t* mask, BoxD<T> outlet, uint8_t* out, hipStream_t st) {
  k_static_cls<T><<<grid_for((int64_t)g.nx * g.ny * g.nz), kBlock, 0, st>>>(g, mask, outlet, out);
}
template <class T>
void launch_cavity_mask(GridD<T> g, SdfD<T> sdf, uint8_t* mask, hipStream_t st) {
  k_cavity_mask<T><<<grid_for((int64_t)g.nx * g.ny * g.nz), kBlock, 0, st>>>(g, sdf, mask);
}

// ------------------------------------------------------------------ instantiations
// ------------------------------------------------------------------ fills without memset nodes
__global__ void k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = 0u;
}
void launch_zero_u32(uint32_t* p, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  int64_t b = (n + kBlock - 1) / kBlock;
  if (b > 4096) b = 4096;
  k_zero_u32<<<(unsigned)b, kBlock, 0, st>>>(p, n);
}

// The vector kernels' unit table of a plane (one entry per (row J, 64-slot chunk) pair
// holding in-grid slots, in row order, (J << 8) | chunk): row J's entries start after
// the chunk counts of the rows before it.
__device__ __host__ inline int unit_chunks(int J, int nx, int ny) {
  const int imin = J - ny + 1 > 0 ? J - ny + 1 : 0, imax = J < nx - 1 ? J : nx - 1;
  return imax / 64 - imin / 64 + 1;
}
__global__ void k_build_units(int nx, int ny, int NJ, uint32_t* __restrict__ out) {
  const int J = blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= NJ) return;
  int start = 0;
  for (int r = 0; r < J; ++r) start += unit_chunks(r, nx, ny);
  const int imin = J - ny + 1 > 0 ? J - ny + 1 : 0;
  const int n = unit_chunks(J, nx, ny);
  for (int c = 0; c < n; ++c) out[start + c] = ((uint32_t)J << 8) | (uint32_t)(imin / 64 + c);
}
int units_per_plane(int nx, int ny) {
  int n = 0;
  for (int J = 0; J < nx + ny - 1; ++J) n += unit_chunks(J, nx, ny);
  return n;
}
void launch_build_units(int nx, int ny, uint32_t* out, hipStream_t st) {
  const int NJ = nx + ny - 1;
  k_build_units<<<(NJ + 63) / 64, 64, 0, st>>>(nx, ny, NJ, out);
}

#define SDR_INST(T)                                                                                \
  template void launch_gather_parts<T>(const Parts<T>&, const Parts<T>&, const uint32_t*, int64_t,  \
                                       hipStream_t);                                               \
  template void launch_classify<T>(const uint8_t*, const Parts<T>&, int64_t, GridD<T>, BoxD<T>,      \
                                   uint8_t*, uint32_t*, hipStream_t);                              \
  template void launch_p2g<T>(int, GridD<T>, const Parts<T>&, const uint32_t*, const uint32_t*, T*, \
                              bool, hipStream_t);                                                  \
  template void launch_axpy_const<T>(T*, int64_t, T, hipStream_t);                                 \
  template void launch_enforce_solid<T>(GridD<T>, FacesD<T>, const uint8_t*, hipStream_t);         \
  template void launch_rhs_flags<T>(GridD<T>, Shear, CFacesD<T>, const uint8_t*, const uint32_t*,  \
                                    const double*, double, int, double, double, double, double,    \
                                    T*, uint16_t*, double*, hipStream_t);                          \
  template void launch_spmv<T>(Shear, const uint16_t*, const T*, T*, const PcgScalars*,            \
                               double*, hipStream_t);                                              \
  template void launch_dot_partials<T>(Shear, const T*, const T*, double*, int, bool,              \
                                       const PcgScalars*, hipStream_t);                            \
  template void launch_maxabs_partials<T>(Shear, const T*, double*, int, const PcgScalars*,        \
                                          hipStream_t);                                            \
  template void launch_axpy_pr<T>(Shear, T*, T*, const T*, const T*, const PcgScalars*, double*,   \
                                  int, hipStream_t);                                               \
  template void launch_update_s<T>(Shear, T*, const T*, const PcgScalars*, hipStream_t);           \
  template void launch_spmv_update<T>(Shear, const uint16_t*, const T*, T*, T*, T*, PcgScalars*,     \
                                      double*, const double*, int, hipStream_t);                   \
  template void launch_axpy_pr_par<T>(Shear, T*, T*, const T*, const T*, const T*, PcgScalars*,     \
                                      double*, const double*, int, int, hipStream_t);              \
  template void launch_mic_plane<T>(int, Shear, const uint16_t*, T*, const T*, T*, int,            \
                                    const PcgScalars*, hipStream_t);                               \
  template void launch_gradient<T>(GridD<T>, Shear, FacesD<T>, const uint8_t*, const T*, T,        \
                                   hipStream_t);                                                   \
  template void launch_unshear<T>(Shear, const T*, T*, hipStream_t);                               \
  template void launch_gradient64<T>(GridD<T>, Shear, FacesD<T>, const uint8_t*, const double*,     \
                                     double, hipStream_t);                                         \
  template void launch_unshear64<T>(Shear, const double*, T*, hipStream_t);                        \
  template void launch_scale_cast<T>(Shear, const double*, double, T*, hipStream_t);               \
  template void launch_accum64<T>(Shear, double*, const T*, double, hipStream_t);                  \
  template void launch_scale_inplace<T>(Shear, T*, double, hipStream_t);                           \
  template void launch_extrap_init<T>(GridD<T>, int, const uint8_t*, uint8_t*, hipStream_t);       \
  template void launch_extrap_iter<T>(int, int, int, const T*, T*, const uint8_t*, uint8_t*,       \
                                      hipStream_t);                                                \
  template void launch_g2p<T>(const Parts<T>&, int64_t, GridD<T>, CFacesD<T>, CFacesD<T>, T,       \
                              hipStream_t);                                                        \
  template void launch_advect<T>(const Parts<T>&, int64_t, GridD<T>, CFacesD<T>, SdfD<T>, T,       \
                                 hipStream_t);                                                     \
  template void launch_g2p_advect<T>(const Parts<T>&, int64_t, GridD<T>, CFacesD<T>, CFacesD<T>, T, \
                                     SdfD<T>, T, hipStream_t);                                     \
  template void launch_bounds<T>(const Parts<T>&, int64_t, SdfD<T>, BoxD<T>, uint32_t*, uint32_t*, \
                                 unsigned long long*, hipStream_t);                                \
  template void launch_compact<T>(const Parts<T>&, const uint32_t*, const uint32_t*, int64_t, int64_t, \
                                  uint32_t*, uint32_t*, hipStream_t);                              \
  template void launch_hist<T>(const Parts<T>&, int64_t, GridD<T>, int32_t*, uint32_t*,            \
                               hipStream_t);                                                       \
  template void launch_redist_apply<T>(const Parts<T>&, const uint32_t*, const int32_t*, int64_t,  \
                                       GridD<T>, CFacesD<T>, SdfD<T>, int32_t*, uint32_t*,         \
                                       hipStream_t);                                               \
  template void launch_maxabs_faces<T>(const T*, int64_t, unsigned long long*, hipStream_t);       \
  template void launch_to_aos<T>(const Parts<T>&, int64_t, double*, double*, hipStream_t);         \
  template void launch_to_aos_f32<T>(const Parts<T>&, int64_t, float*, hipStream_t);               \
  template void launch_from_aos<T>(const Parts<T>&, int64_t, int64_t, const double*, const double*, \
                                   hipStream_t);                                                   \
  template void launch_seed_fill<T>(const Parts<T>&, int64_t, GridD<T>, const uint8_t*,            \
                                    const uint32_t*, int64_t, uint32_t, uint64_t, T, T, T,         \
                                    hipStream_t);                                                  \
  template void launch_static_cls<T>(GridD<T>, const uint8_t*, BoxD<T>, uint8_t*, hipStream_t);    \
  template void launch_cavity_mask<T>(GridD<T>, SdfD<T>, uint8_t*, hipStream_t);

SDR_INST(double)
SDR_INST(float)

}  // namespace sdr
