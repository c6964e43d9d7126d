// p2g.hip — P2G transfer (solver.rs:560-585, grid.rs:52-88, solver.rs:944-952) as one
// LDS-staged gather over the cell-sorted particles for all three face components.
//
// A face (fi, fj, fk) of component c receives from the particles of the cells [f-1, f]
// along its own axis and [f-1, f+1] across (clipped to the grid), visited slab (k) by
// slab, row (j) by row, cell (i) by cell, particle by particle in sorted order; each
// visit adds w * v and w where w = (wx * wy) * wz from the clamped trilinear base of the
// particle.  This kernel keeps exactly that order, so it reproduces the face-by-face
// gather bit for bit, and it reads each particle once per slab for all three
// components instead of once per visit and component:
//   * a workgroup owns a 16 x 16 column of faces (fi, fj) over 32 levels fk (4-16 on
//     grids too small to fill the chip) and marches up the slabs ck that feed them;
//     waves 0-3 gather the u and v faces of the column, waves 4-7 the w faces (about the
//     same number of candidate particles each);
//   * a slab's particles (18 x 18 cells) are copied into LDS once (structure of arrays)
//     with everything a visit needs precomputed: the clamped trilinear fractions of each
//     axis for the along-axis stencil (l) and the across-axis one (l - 0.5), the three
//     velocities, and a 16-bit code of the particle's cell in the slab and its half-cell
//     bits (which side of the cell centre each across-axis base falls on).  The bases of
//     every stencil follow from that code, so a visit tests the code, then reads the four
//     values its component needs;
//   * a particle of slab ck feeds u/v levels bk and bk + 1 with bk in {ck - 1, ck}, and w
//     levels ck and ck + 1, so each thread keeps rolling accumulator pairs and finalises
//     a level after its last slab: out = sum(w v) / sum(w) if sum(w) > 1e-12, else 0.
// Slabs holding more particles than the LDS staging area are processed in passes over
// consecutive windows of the staged order, which keeps the order.
//
// Heavy cells (P2gHeavy, thresh > 0): a face thread beside a cell of a thousand
// particles would walk all of them while the rest of its workgroup waits.  Cells at or
// over the threshold are reduced beforehand by k_p2g_heavy, one workgroup per cell (for
// each of the 54 faces the cell's particles can reach, the sums of w v and w over the
// cell's particles, each lane in sorted order, then a fixed tree), left out of the
// slab's staging, and added to each face's sums as one term after the slab's other
// particles.  The terms are the same products as the walk's; only the order of the sums
// changes, so such faces agree with the face-by-face gather to rounding.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace sdr {
namespace {

constexpr int kTX = 16, kTY = 16;          // face columns per workgroup
constexpr int kRows = kTY + 2, kCols = kTX + 2;
constexpr int kSlabCells = kRows * kCols;  // 324
constexpr int kZChunk = 32;                // face levels per workgroup (fewer on small grids)
constexpr int kThreads = 512;

template <class T>
struct StageCap {  // about 70 KB of LDS: two workgroups per CU
  static constexpr int n = sizeof(T) == 4 ? 1792 : 896;
};

// staged code: bits 0-4 the cell's column in the slab, 5-9 its row, then the half-cell
// bits: bit 10 + a is 1 when the across-axis base of axis a (grid.rs:26-35 on l - 0.5)
// is the particle's own cell, 0 when it is the cell before
constexpr int kHx = 10, kHy = 11, kHz = 12;
constexpr int kHeavySlab = 32;  // heavy cells per slab taken as one term each (more are walked)
static_assert(kSlabCells <= kThreads, "one slab cell per thread");

template <class T>
__device__ inline T wave_sum_t(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_p2g_fused(GridD<T> g, Parts<T> p, const uint32_t* __restrict__ cstart,
                                                        const uint32_t* __restrict__ cend, T* __restrict__ ou,
                                                        T* __restrict__ ov, T* __restrict__ ow, int zc, P2gHeavy h) {
  constexpr int CAP = StageCap<T>::n;
  // fractions along (a) and across (c) per axis, velocities, cell/half-cell code
  __shared__ T sfax[CAP], sfay[CAP], sfaz[CAP], sfcx[CAP], sfcy[CAP], sfcz[CAP];
  __shared__ T svx[CAP], svy[CAP], svz[CAP];
  __shared__ uint16_t spk[CAP];
  __shared__ uint32_t soff[kSlabCells + 1];  // staged offset of each cell of the slab
  __shared__ uint32_t sglob[kSlabCells];     // its first particle in the sorted arrays
  __shared__ uint32_t wsum_[kThreads / 64];
  __shared__ uint16_t shv_cell[kHeavySlab];  // the slab's heavy cells (slab cell index)
  __shared__ uint32_t shv_idx[kHeavySlab];   // and their entries
  __shared__ uint32_t shv_wave[kThreads / 64];
  const uint32_t hn = h.thresh ? (*h.ctr < h.cap ? *h.ctr : h.cap) : 0u;

  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int ntx = (nx + 1 + kTX - 1) / kTX, nty = (ny + 1 + kTY - 1) / kTY;
  const int b = blockIdx.x;
  const int x0 = (b % ntx) * kTX, y0 = ((b / ntx) % nty) * kTY;
  const int z0 = (b / (ntx * nty)) * zc;
  const int z1 = z0 + zc < nz ? z0 + zc : nz;        // u, v levels [z0, z1)
  const int z1w = z0 + zc < nz + 1 ? z0 + zc : nz + 1;  // w levels [z0, z1w)
  const int col = threadIdx.x & 255;
  const bool wwave = threadIdx.x >= 256;  // waves 4-7: the w faces
  const int tx = col % kTX, ty = col / kTX;
  const int fi = x0 + tx, fj = y0 + ty;
  const bool u_ok = !wwave && fi <= nx && fj < ny;
  const bool v_ok = !wwave && fi < nx && fj <= ny;
  const bool w_ok = wwave && fi < nx && fj < ny;

  // the slab's cells: rows cj in [rlo, rhi], columns ci in [clo, chi]
  const int rlo = y0 - 1 > 0 ? y0 - 1 : 0, rhi = y0 + kTY < ny - 1 ? y0 + kTY : ny - 1;
  const int clo = x0 - 1 > 0 ? x0 - 1 : 0, chi = x0 + kTX < nx - 1 ? x0 + kTX : nx - 1;
  const int nrows = rhi - rlo + 1, ncols = chi - clo + 1, ncell = nrows * ncols;

  // slabs feeding the levels: u/v level L takes ck in [L-1, L+1], w level L [L-1, L]
  const int cs = z0 - 1 > 0 ? z0 - 1 : 0;
  const int ce = z1 < nz - 1 ? z1 : nz - 1;

  // accumulators: u/v of levels ck-1, ck, ck+1 while slab ck is processed; w of ck, ck+1
  T uA0 = T(0), uA1 = T(0), uA2 = T(0), uW0 = T(0), uW1 = T(0), uW2 = T(0);
  T vA0 = T(0), vA1 = T(0), vA2 = T(0), vW0 = T(0), vW1 = T(0), vW2 = T(0);
  T wA0 = T(0), wA1 = T(0), wW0 = T(0), wW1 = T(0);
  auto put = [&](T* out, int fnx, int fny, int L, int lhi, int a, int bb, T acc, T w) {
    if (L < z0 || L >= lhi) return;
    out[((int64_t)L * fny + bb) * fnx + a] = w > T(1e-12) ? acc / w : T(0);
  };

  for (int ck = cs; ck <= ce; ++ck) {
    // ---- cell table of the slab (one cell per thread) and its prefix sum (staged
    // offsets); heavy cells are listed in cell order instead of staged
    bool hv = false;
    uint32_t hidx = 0;
    if ((int)threadIdx.x < ncell) {
      const int t = threadIdx.x;
      const int r = t / ncols, c = t % ncols;
      const int64_t cell = ((int64_t)ck * ny + (rlo + r)) * nx + (clo + c);
      const uint32_t s0 = cstart[cell], e0 = cend[cell];
      if (hn && e0 - s0 >= h.thresh) {
        const uint32_t e = h.hid[cell];  // valid only if the entry names this cell
        if (e < hn && h.hcell[e] == (uint32_t)cell) hv = true, hidx = e;
      }
      sglob[t] = s0;
      soff[t + 1] = e0 - s0;
    }
    if (threadIdx.x == 0) soff[0] = 0;
    int nhs = 0;
    if (hn) {
      const uint64_t bal = __ballot(hv);
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      if (lane == 0) shv_wave[wave] = (uint32_t)__popcll(bal);
      __syncthreads();
      int before = 0;
      for (int w = 0; w < kThreads / 64; ++w) {
        before += w < wave ? (int)shv_wave[w] : 0;
        nhs += (int)shv_wave[w];
      }
      const int slot = before + __popcll(bal & ((1ull << lane) - 1ull));
      if (hv && slot < kHeavySlab) {
        shv_cell[slot] = (uint16_t)threadIdx.x;
        shv_idx[slot] = hidx;
        soff[threadIdx.x + 1] = 0;  // not staged
      }
      nhs = nhs < kHeavySlab ? nhs : kHeavySlab;
    }
    __syncthreads();
    {  // inclusive scan of entries 1..ncell (one per thread), then the wave totals
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      const int e1 = threadIdx.x + 1;
      uint32_t s = e1 <= ncell ? soff[e1] : 0u;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t up = __shfl_up(s, o);
        if (lane >= o) s += up;
      }
      if (lane == 63) wsum_[wave] = s;
      __syncthreads();
      uint32_t base = 0;
      for (int w = 0; w < wave; ++w) base += wsum_[w];
      if (e1 <= ncell) soff[e1] = base + s;
      __syncthreads();
    }
    const uint32_t total = soff[ncell];
    // ---- passes over windows of the staged order
    for (uint32_t w0 = 0; w0 < total; w0 += CAP) {
      const uint32_t w1 = w0 + CAP < total ? w0 + CAP : total;
      for (uint32_t s = w0 + threadIdx.x; s < w1; s += kThreads) {
        int lo = 0, hi = ncell - 1;  // the cell holding staged index s
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (soff[mid] <= s) lo = mid;
          else hi = mid - 1;
        }
        const uint32_t gidx = sglob[lo] + (s - soff[lo]);
        const V3<T> l = g.local(mk3(p.px[gidx], p.py[gidx], p.pz[gidx]));
        const int ci = clo + lo % ncols, cj = rlo + lo / ncols;
        int bx, by, bz, ax_, ay_, az_;
        T cx, cy, cz, fx, fy, fz;
        clamp_base(l.x, nx + 1, ax_, fx);
        clamp_base(l.y, ny + 1, ay_, fy);
        clamp_base(l.z, nz + 1, az_, fz);
        clamp_base(l.x - T(0.5), nx, bx, cx);
        clamp_base(l.y - T(0.5), ny, by, cy);
        clamp_base(l.z - T(0.5), nz, bz, cz);
        const uint32_t q = s - w0;
        sfax[q] = fx, sfay[q] = fy, sfaz[q] = fz;
        sfcx[q] = cx, sfcy[q] = cy, sfcz[q] = cz;
        svx[q] = p.vx[gidx], svy[q] = p.vy[gidx], svz[q] = p.vz[gidx];
        spk[q] = (uint16_t)((ci - clo) | ((cj - rlo) << 5) | ((bx == ci) << kHx) | ((by == cj) << kHy) |
                            ((bz == ck) << kHz));
      }
      __syncthreads();
      // staged run of the row cj's cells i0..i1 of the slab (consecutive in the sorted
      // order), clipped to the window
      auto run = [&](int i0, int i1, int cj, uint32_t& a, uint32_t& e) {
        const int q = (cj - rlo) * ncols;
        a = soff[q + (i0 - clo)], e = soff[q + (i1 - clo) + 1];
        a = a > w0 ? a : w0;
        e = e < w1 ? e : w1;
      };
      // a visit: the particle's along-axis base is its cell, its across-axis base the
      // cell before or the cell itself (half-cell bit); d = face - base must be 0 or 1.
      // Runs are walked four particles at a time with every load issued before the
      // first use; a particle outside the stencil (or past the run) is skipped by
      // selects, so the sums keep the sequential order.
      constexpr int UN = 4;
      if (!wwave) {
        if (u_ok) {  // along x: cells fi-1, fi; across y, z
          const int i0 = fi - 1 > 0 ? fi - 1 : 0, i1 = fi < nx - 1 ? fi : nx - 1;
          const int j0 = fj - 1 > 0 ? fj - 1 : 0, j1 = fj + 1 < ny - 1 ? fj + 1 : ny - 1;
          for (int cj = j0; cj <= j1; ++cj) {
            uint32_t a, e;
            run(i0, i1, cj, a, e);
            for (uint32_t s = a; s < e; s += UN) {
              int pk[UN];
              T fx[UN], fy[UN], fz[UN], v[UN];
#pragma unroll
              for (int u = 0; u < UN; ++u) {
                const uint32_t q = s + u < e ? s + u - w0 : 0u;
                pk[u] = s + u < e ? (int)spk[q] : 0x7fff;
                fx[u] = sfax[q], fy[u] = sfcy[q], fz[u] = sfcz[q], v[u] = svx[q];
              }
#pragma unroll
              for (int u = 0; u < UN; ++u) {
                const int di = fi - (clo + (pk[u] & 31));
                const int dj = fj - (cj - 1 + ((pk[u] >> kHy) & 1));
                const bool ok = (unsigned)di <= 1u && (unsigned)dj <= 1u && pk[u] != 0x7fff;
                const bool lo = ((pk[u] >> kHz) & 1) == 0;  // base ck - 1: levels ck - 1, ck
                const T wxy = (di ? fx[u] : T(1) - fx[u]) * (dj ? fy[u] : T(1) - fy[u]);
                const T wlo = wxy * (T(1) - fz[u]), whi = wxy * fz[u];
                const T vlo = wlo * v[u], vhi = whi * v[u];
                uA0 = ok && lo ? uA0 + vlo : uA0, uW0 = ok && lo ? uW0 + wlo : uW0;
                uA1 = ok ? uA1 + (lo ? vhi : vlo) : uA1, uW1 = ok ? uW1 + (lo ? whi : wlo) : uW1;
                uA2 = ok && !lo ? uA2 + vhi : uA2, uW2 = ok && !lo ? uW2 + whi : uW2;
              }
            }
          }
        }
        if (v_ok) {  // along y: cells fj-1, fj; across x, z
          const int i0 = fi - 1 > 0 ? fi - 1 : 0, i1 = fi + 1 < nx - 1 ? fi + 1 : nx - 1;
          const int j0 = fj - 1 > 0 ? fj - 1 : 0, j1 = fj < ny - 1 ? fj : ny - 1;
          for (int cj = j0; cj <= j1; ++cj) {
            uint32_t a, e;
            run(i0, i1, cj, a, e);
            const int dj = fj - cj;
            for (uint32_t s = a; s < e; s += UN) {
              int pk[UN];
              T fx[UN], fy[UN], fz[UN], v[UN];
#pragma unroll
              for (int u = 0; u < UN; ++u) {
                const uint32_t q = s + u < e ? s + u - w0 : 0u;
                pk[u] = s + u < e ? (int)spk[q] : 0x7fff;
                fx[u] = sfcx[q], fy[u] = sfay[q], fz[u] = sfcz[q], v[u] = svy[q];
              }
#pragma unroll
              for (int u = 0; u < UN; ++u) {
                const int di = fi - (clo + (pk[u] & 31) - 1 + ((pk[u] >> kHx) & 1));
                const bool ok = (unsigned)di <= 1u && pk[u] != 0x7fff;
                const bool lo = ((pk[u] >> kHz) & 1) == 0;
                const T wxy = (di ? fx[u] : T(1) - fx[u]) * (dj ? fy[u] : T(1) - fy[u]);
                const T wlo = wxy * (T(1) - fz[u]), whi = wxy * fz[u];
                const T vlo = wlo * v[u], vhi = whi * v[u];
                vA0 = ok && lo ? vA0 + vlo : vA0, vW0 = ok && lo ? vW0 + wlo : vW0;
                vA1 = ok ? vA1 + (lo ? vhi : vlo) : vA1, vW1 = ok ? vW1 + (lo ? whi : wlo) : vW1;
                vA2 = ok && !lo ? vA2 + vhi : vA2, vW2 = ok && !lo ? vW2 + whi : vW2;
              }
            }
          }
        }
      } else if (w_ok) {  // along z (levels ck, ck+1); across x, y
        const int i0 = fi - 1 > 0 ? fi - 1 : 0, i1 = fi + 1 < nx - 1 ? fi + 1 : nx - 1;
        const int j0 = fj - 1 > 0 ? fj - 1 : 0, j1 = fj + 1 < ny - 1 ? fj + 1 : ny - 1;
        for (int cj = j0; cj <= j1; ++cj) {
          uint32_t a, e;
          run(i0, i1, cj, a, e);
          for (uint32_t s = a; s < e; s += UN) {
            int pk[UN];
            T fx[UN], fy[UN], fz[UN], v[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) {
              const uint32_t q = s + u < e ? s + u - w0 : 0u;
              pk[u] = s + u < e ? (int)spk[q] : 0x7fff;
              fx[u] = sfcx[q], fy[u] = sfcy[q], fz[u] = sfaz[q], v[u] = svz[q];
            }
#pragma unroll
            for (int u = 0; u < UN; ++u) {
              const int di = fi - (clo + (pk[u] & 31) - 1 + ((pk[u] >> kHx) & 1));
              const int dj = fj - (cj - 1 + ((pk[u] >> kHy) & 1));
              const bool ok = (unsigned)di <= 1u && (unsigned)dj <= 1u && pk[u] != 0x7fff;
              const T wxy = (di ? fx[u] : T(1) - fx[u]) * (dj ? fy[u] : T(1) - fy[u]);
              const T wlo = wxy * (T(1) - fz[u]), whi = wxy * fz[u];
              wA0 = ok ? wA0 + wlo * v[u] : wA0, wW0 = ok ? wW0 + wlo : wW0;
              wA1 = ok ? wA1 + whi * v[u] : wA1, wW1 = ok ? wW1 + whi : wW1;
            }
          }
        }
      }
      __syncthreads();
    }
    // ---- the slab's heavy cells, one term each (targets as k_p2g_heavy numbers them)
    for (int e = 0; e < nhs; ++e) {
      const int t = shv_cell[e];
      const int hc = clo + t % ncols, hr = rlo + t / ncols;
      const T* P = (const T*)h.part + (size_t)shv_idx[e] * kP2gHeavyVals;
      if (u_ok) {  // faces hc + dx, hr - 1 + dy, levels ck - 1 .. ck + 1
        const int dx = fi - hc, dy = fj - hr + 1;
        if ((unsigned)dx <= 1u && (unsigned)dy <= 2u) {
          const int k = (dx * 3 + dy) * 3;
          uA0 += P[k], uA1 += P[k + 1], uA2 += P[k + 2];
          uW0 += P[18 + k], uW1 += P[18 + k + 1], uW2 += P[18 + k + 2];
        }
      }
      if (v_ok) {  // faces hc - 1 + dx, hr + dy, levels ck - 1 .. ck + 1
        const int dx = fi - hc + 1, dy = fj - hr;
        if ((unsigned)dx <= 2u && (unsigned)dy <= 1u) {
          const int k = (dx * 2 + dy) * 3;
          vA0 += P[36 + k], vA1 += P[36 + k + 1], vA2 += P[36 + k + 2];
          vW0 += P[54 + k], vW1 += P[54 + k + 1], vW2 += P[54 + k + 2];
        }
      }
      if (w_ok) {  // faces hc - 1 + dx, hr - 1 + dy, levels ck, ck + 1
        const int dx = fi - hc + 1, dy = fj - hr + 1;
        if ((unsigned)dx <= 2u && (unsigned)dy <= 2u) {
          const int k = (dx * 3 + dy) * 2;
          wA0 += P[72 + k], wA1 += P[72 + k + 1];
          wW0 += P[90 + k], wW1 += P[90 + k + 1];
        }
      }
    }
    // ---- finalise the levels whose last slab this was, then roll the windows
    if (u_ok) {
      put(ou, nx + 1, ny, ck - 1, z1, fi, fj, uA0, uW0);
      if (ck == nz - 1) put(ou, nx + 1, ny, ck, z1, fi, fj, uA1, uW1);
    }
    if (v_ok) {
      put(ov, nx, ny + 1, ck - 1, z1, fi, fj, vA0, vW0);
      if (ck == nz - 1) put(ov, nx, ny + 1, ck, z1, fi, fj, vA1, vW1);
    }
    if (w_ok) {
      put(ow, nx, ny, ck, z1w, fi, fj, wA0, wW0);
      if (ck == nz - 1) put(ow, nx, ny, ck + 1, z1w, fi, fj, wA1, wW1);
    }
    uA0 = uA1, uW0 = uW1, uA1 = uA2, uW1 = uW2, uA2 = T(0), uW2 = T(0);
    vA0 = vA1, vW0 = vW1, vA1 = vA2, vW1 = vW2, vA2 = T(0), vW2 = T(0);
    wA0 = wA1, wW0 = wW1, wA1 = T(0), wW1 = T(0);
    __syncthreads();  // the slab's tables are rewritten by the next slab
  }
}

__global__ void k_heavy_reset(uint32_t* ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *ctr = 0;
}
// The cells holding at least h.thresh particles, in no particular order (an entry's
// sums do not depend on its index).
__global__ void k_heavy_list(int64_t ncells, const uint32_t* __restrict__ cstart,
                             const uint32_t* __restrict__ cend, P2gHeavy h) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += stride) {
    if (cend[c] - cstart[c] < h.thresh) continue;
    const uint32_t e = atomicAdd(h.ctr, 1u);
    if (e < h.cap) {
      h.hid[c] = e;
      h.hcell[e] = (uint32_t)c;
    }
  }
}

// One workgroup per heavy cell: for component c and target face k (c = 0, u: faces
// ci + dx, cj - 1 + dy, levels ck - 1 + dz, k = (dx * 3 + dy) * 3 + dz; c = 1, v: faces
// ci - 1 + dx, cj + dy, k = (dx * 2 + dy) * 3 + dz; c = 2, w: faces ci - 1 + dx,
// cj - 1 + dy, levels ck + dz, k = (dx * 3 + dy) * 2 + dz), the sums of w v and of w over
// the cell's particles, with w the product the fused walk forms for that face.  Lane l
// sums particles l, l + 256, ... in sorted order; then a butterfly per wave and the four
// waves in order.
template <class T>
__global__ void __launch_bounds__(256) k_p2g_heavy(GridD<T> g, Parts<T> p, const uint32_t* __restrict__ cstart,
                                                   const uint32_t* __restrict__ cend, P2gHeavy h) {
  __shared__ T red[4][36];
  const uint32_t nh = *h.ctr < h.cap ? *h.ctr : h.cap;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t e = blockIdx.x; e < nh; e += gridDim.x) {
    const uint32_t cell = h.hcell[e];
    const int ci = (int)(cell % (uint32_t)nx), cj = (int)((cell / (uint32_t)nx) % (uint32_t)ny);
    const int ck = (int)(cell / ((uint32_t)nx * (uint32_t)ny));
    const uint32_t s0 = cstart[cell], e0 = cend[cell];
    T* out = (T*)h.part + (size_t)e * kP2gHeavyVals;
#pragma unroll 1
    for (int comp = 0; comp < 3; ++comp) {
      T A[18], W[18];
#pragma unroll
      for (int k = 0; k < 18; ++k) A[k] = T(0), W[k] = T(0);
      for (uint32_t t = s0 + threadIdx.x; t < e0; t += 256) {
        const V3<T> l = g.local(mk3(p.px[t], p.py[t], p.pz[t]));
        int ax_, ay_, az_, bx, by, bz;
        T fx, fy, fz, cx, cy, cz;
        clamp_base(l.x, nx + 1, ax_, fx);
        clamp_base(l.y, ny + 1, ay_, fy);
        clamp_base(l.z, nz + 1, az_, fz);
        clamp_base(l.x - T(0.5), nx, bx, cx);
        clamp_base(l.y - T(0.5), ny, by, cy);
        clamp_base(l.z - T(0.5), nz, bz, cz);
        // across-axis bases: the cell before (0) or the cell itself (1)
        const int hx = bx == ci, hy = by == cj, hz = bz == ck;
        if (comp == 0) {
          const T v = p.vx[t];
#pragma unroll
          for (int dx = 0; dx < 2; ++dx)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
              const int dj = dy - hy;
              const T wxy = (dx ? fx : T(1) - fx) * (dj ? cy : T(1) - cy);
              const T wlo = wxy * (T(1) - cz), whi = wxy * cz;
              const T vlo = wlo * v, vhi = whi * v;
#pragma unroll
              for (int dz = 0; dz < 3; ++dz) {
                const int ez = dz - hz, k = (dx * 3 + dy) * 3 + dz;
                const bool ok = (unsigned)dj <= 1u && (unsigned)ez <= 1u;
                A[k] = ok ? A[k] + (ez ? vhi : vlo) : A[k];
                W[k] = ok ? W[k] + (ez ? whi : wlo) : W[k];
              }
            }
        } else if (comp == 1) {
          const T v = p.vy[t];
#pragma unroll
          for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
              const int di = dx - hx;
              const T wxy = (di ? cx : T(1) - cx) * (dy ? fy : T(1) - fy);
              const T wlo = wxy * (T(1) - cz), whi = wxy * cz;
              const T vlo = wlo * v, vhi = whi * v;
#pragma unroll
              for (int dz = 0; dz < 3; ++dz) {
                const int ez = dz - hz, k = (dx * 2 + dy) * 3 + dz;
                const bool ok = (unsigned)di <= 1u && (unsigned)ez <= 1u;
                A[k] = ok ? A[k] + (ez ? vhi : vlo) : A[k];
                W[k] = ok ? W[k] + (ez ? whi : wlo) : W[k];
              }
            }
        } else {
          const T v = p.vz[t];
#pragma unroll
          for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
              const int di = dx - hx, dj = dy - hy;
              const bool ok = (unsigned)di <= 1u && (unsigned)dj <= 1u;
              const T wxy = (di ? cx : T(1) - cx) * (dj ? cy : T(1) - cy);
              const T wlo = wxy * (T(1) - fz), whi = wxy * fz;
              const int k = (dx * 3 + dy) * 2;
              A[k] = ok ? A[k] + wlo * v : A[k];
              W[k] = ok ? W[k] + wlo : W[k];
              A[k + 1] = ok ? A[k + 1] + whi * v : A[k + 1];
              W[k + 1] = ok ? W[k + 1] + whi : W[k + 1];
            }
        }
      }
#pragma unroll
      for (int k = 0; k < 18; ++k) {
        A[k] = wave_sum_t(A[k]);
        W[k] = wave_sum_t(W[k]);
      }
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 18; ++k) red[wave][k] = A[k], red[wave][18 + k] = W[k];
      }
      __syncthreads();
      if (threadIdx.x < 36) {
        const int q = threadIdx.x;
        out[comp * 36 + q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
      }
      __syncthreads();
    }
  }
}

}  // namespace

template <class T>
void launch_p2g_heavy(GridD<T> g, const Parts<T>& p, const uint32_t* cstart, const uint32_t* cend, P2gHeavy h,
                      hipStream_t st) {
  if (!h.thresh) return;
  const int64_t nc = (int64_t)g.nx * g.ny * g.nz;
  k_heavy_reset<<<1, 64, 0, st>>>(h.ctr);
  const int64_t lb = (nc + 255) / 256;
  k_heavy_list<<<(unsigned)(lb < 4096 ? lb : 4096), 256, 0, st>>>(nc, cstart, cend, h);
  k_p2g_heavy<T><<<1024, 256, 0, st>>>(g, p, cstart, cend, h);
}

template <class T>
bool launch_p2g_fused(GridD<T> g, const Parts<T>& p, const uint32_t* cstart, const uint32_t* cend, FacesD<T> f,
                      P2gHeavy h, hipStream_t st) {
  const int64_t cols = (int64_t)((g.nx + 1 + kTX - 1) / kTX) * ((g.ny + 1 + kTY - 1) / kTY);
  // levels per workgroup: halved from 32 (down to 4) until the launch fills the chip with
  // two workgroups per CU; a face's sums do not depend on which workgroup owns its level
  int zc = kZChunk;
  while (zc > 4 && cols * ((g.nz + 1 + zc - 1) / zc) < 512) zc /= 2;
  const int64_t blocks = cols * ((g.nz + 1 + zc - 1) / zc);
  if (blocks >= (int64_t)1 << 31) return false;
  k_p2g_fused<T><<<(unsigned)blocks, kThreads, 0, st>>>(g, p, cstart, cend, f.u, f.v, f.w, zc, h);
  return true;
}

template bool launch_p2g_fused<float>(GridD<float>, const Parts<float>&, const uint32_t*, const uint32_t*,
                                      FacesD<float>, P2gHeavy, hipStream_t);
template bool launch_p2g_fused<double>(GridD<double>, const Parts<double>&, const uint32_t*, const uint32_t*,
                                       FacesD<double>, P2gHeavy, hipStream_t);
template void launch_p2g_heavy<float>(GridD<float>, const Parts<float>&, const uint32_t*, const uint32_t*, P2gHeavy,
                                      hipStream_t);
template void launch_p2g_heavy<double>(GridD<double>, const Parts<double>&, const uint32_t*, const uint32_t*,
                                       P2gHeavy, hipStream_t);

}  // namespace sdr
