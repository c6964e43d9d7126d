// p2g.hip — P2G transfer (solver.rs:560-585, grid.rs:52-88, solver.rs:944-952) as an
// LDS-staged gather over the cell-sorted particles.
//
// A face (fi, fj, fk) of component COMP receives from the particles of the cells
// [f-1, f] along its own axis and [f-1, f+1] across (clipped to the grid), visited slab
// (k) by slab, row (j) by row, cell (i) by cell, particle by particle in sorted order;
// each visit adds w * v and w where w = (wx * wy) * wz from the clamped trilinear base
// of the particle.  This kernel keeps exactly that order, so it reproduces the
// face-by-face gather bit for bit, but it stages each particle once per slab instead of
// once per visit:
//   * a workgroup owns a 16 x 16 column of faces (fi, fj) over a range of levels fk and
//     marches up the slabs ck that feed them;
//   * a slab's particles (18 x 18 cells) are copied into LDS with their base indices and
//     fractions already computed (the (p - origin) / dx divisions happen once per
//     particle and slab, not once per visit);
//   * a particle of slab ck feeds levels bk and bk + 1, both in {ck - 1, ck, ck + 1}, so
//     each thread keeps three rolling accumulator pairs and finalises a level after its
//     last slab: out = sum(w v) / sum(w) if sum(w) > 1e-12, else 0.
// Slabs holding more particles than the LDS staging area are processed in passes over
// consecutive windows of the staged order, which keeps the order.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace sdr {
namespace {

constexpr int kTX = 16, kTY = 16;         // face columns per workgroup
constexpr int kRows = kTY + 2, kCols = kTX + 2;
constexpr int kSlabCells = kRows * kCols;  // 324
constexpr int kZChunk = 32;                // face levels per workgroup

template <class T>
struct Staged {
  T fx, fy, fz, v;
};
template <class T>
struct StageCap {
  static constexpr int n = sizeof(T) == 4 ? 3072 : 1536;
};

template <class T, int COMP>
__global__ void __launch_bounds__(256) k_p2g_tiled(GridD<T> g, Parts<T> p, const uint32_t* __restrict__ cstart,
                                                   const uint32_t* __restrict__ cend, T* __restrict__ out) {
  constexpr int CAP = StageCap<T>::n;
  __shared__ Staged<T> sp[CAP];
  __shared__ uint32_t sbase[CAP];         // bi | bj << 11 | bk << 22
  __shared__ uint32_t soff[kSlabCells + 1];  // staged offset of each cell of the slab
  __shared__ uint32_t sglob[kSlabCells];     // its first particle in the sorted arrays
  __shared__ uint32_t wsum_[4];

  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int fnx = nx + (COMP == 0), fny = ny + (COMP == 1), fnz = nz + (COMP == 2);
  const int ntx = (fnx + kTX - 1) / kTX, nty = (fny + kTY - 1) / kTY;
  const int b = blockIdx.x;
  const int x0 = (b % ntx) * kTX, y0 = ((b / ntx) % nty) * kTY;
  const int z0 = (b / (ntx * nty)) * kZChunk;
  const int z1 = z0 + kZChunk < fnz ? z0 + kZChunk : fnz;
  const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
  const int fi = x0 + tx, fj = y0 + ty;
  const bool face_ok = fi < fnx && fj < fny;
  const T* vel = COMP == 0 ? p.vx : (COMP == 1 ? p.vy : p.vz);

  // the slab's cells: rows cj in [rlo, rhi], columns ci in [clo, chi]
  const int rlo = y0 - 1 > 0 ? y0 - 1 : 0, rhi = y0 + kTY < ny - 1 ? y0 + kTY : ny - 1;
  const int clo = x0 - 1 > 0 ? x0 - 1 : 0, chi = x0 + kTX < nx - 1 ? x0 + kTX : nx - 1;
  const int nrows = rhi - rlo + 1, ncols = chi - clo + 1, ncell = nrows * ncols;
  // this face's candidate cells (the same clipped ranges as the face-by-face gather)
  const int cilo = (fi - 1 > 0 ? fi - 1 : 0) - clo;
  const int cihi = ((COMP == 0 ? fi : fi + 1) < nx - 1 ? (COMP == 0 ? fi : fi + 1) : nx - 1) - clo;
  const int cjlo = (fj - 1 > 0 ? fj - 1 : 0) - rlo;
  const int cjhi = ((COMP == 1 ? fj : fj + 1) < ny - 1 ? (COMP == 1 ? fj : fj + 1) : ny - 1) - rlo;

  // slabs feeding levels [z0, z1): a level L takes ck in [L-1, L+1] (COMP 2: [L-1, L])
  const int cs = z0 - 1 > 0 ? z0 - 1 : 0;
  const int ce_ = (COMP == 2 ? z1 - 1 : z1);
  const int ce = ce_ < nz - 1 ? ce_ : nz - 1;

  // accumulators of levels ck-1, ck, ck+1 while slab ck is processed
  T A0 = T(0), A1 = T(0), A2 = T(0), W0 = T(0), W1 = T(0), W2 = T(0);
  auto finalize = [&](int L, T a, T w) {
    if (!face_ok || L < z0 || L >= z1) return;
    out[((int64_t)L * fny + fj) * fnx + fi] = w > T(1e-12) ? a / w : T(0);
  };

  for (int ck = cs; ck <= ce; ++ck) {
    // ---- cell table of the slab and its prefix sum (staged offsets)
    for (int t = threadIdx.x; t < ncell; t += blockDim.x) {
      const int r = t / ncols, c = t % ncols;
      const int64_t cell = ((int64_t)ck * ny + (rlo + r)) * nx + (clo + c);
      const uint32_t s0 = cstart[cell], e0 = cend[cell];
      sglob[t] = s0;
      soff[t + 1] = e0 - s0;
    }
    if (threadIdx.x == 0) soff[0] = 0;
    __syncthreads();
    // block prefix sum over <= 324 entries: 256 threads, two entries for some
    {
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      // each thread owns entries 2t+1, 2t+2 (1-based), inclusive scan of pairs
      const int e1 = 2 * threadIdx.x + 1, e2 = e1 + 1;
      const uint32_t a = e1 <= ncell ? soff[e1] : 0u, bb = e2 <= ncell ? soff[e2] : 0u;
      uint32_t s = a + bb;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t up = __shfl_up(s, o);
        if (lane >= o) s += up;
      }
      if (lane == 63) wsum_[wave] = s;
      __syncthreads();
      uint32_t base = 0;
      for (int w = 0; w < wave; ++w) base += wsum_[w];
      const uint32_t incl = base + s;  // sum of entries 1..e2
      if (e1 <= ncell) soff[e1] = incl - bb;
      if (e2 <= ncell) soff[e2] = incl;
      __syncthreads();
    }
    const uint32_t total = soff[ncell];
    // ---- passes over windows of the staged order
    for (uint32_t w0 = 0; w0 < total; w0 += CAP) {
      const uint32_t w1 = w0 + CAP < total ? w0 + CAP : total;
      // stage particles w0..w1 (row r's particles are contiguous in the sorted arrays)
      for (uint32_t s = w0 + threadIdx.x; s < w1; s += blockDim.x) {
        int lo = 0, hi = ncell - 1;  // the cell holding staged index s
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (soff[mid] <= s) lo = mid;
          else hi = mid - 1;
        }
        const uint32_t gidx = sglob[lo] + (s - soff[lo]);
        const V3<T> l = g.local(mk3(p.px[gidx], p.py[gidx], p.pz[gidx]));
        const T gx = l.x - (COMP == 0 ? T(0) : T(0.5));
        const T gy = l.y - (COMP == 1 ? T(0) : T(0.5));
        const T gz = l.z - (COMP == 2 ? T(0) : T(0.5));
        int bi, bj, bk;
        T ffx, ffy, ffz;
        clamp_base(gx, fnx, bi, ffx);
        clamp_base(gy, fny, bj, ffy);
        clamp_base(gz, fnz, bk, ffz);
        sp[s - w0] = Staged<T>{ffx, ffy, ffz, vel[gidx]};
        sbase[s - w0] = (uint32_t)bi | ((uint32_t)bj << 11) | ((uint32_t)bk << 22);
      }
      __syncthreads();
      if (face_ok) {
        for (int r = cjlo; r <= cjhi; ++r) {
          uint32_t a = soff[r * ncols + cilo], e = soff[r * ncols + cihi + 1];
          a = a > w0 ? a : w0;
          e = e < w1 ? e : w1;
          for (uint32_t s = a; s < e; ++s) {
            const uint32_t pb = sbase[s - w0];
            const int di = fi - (int)(pb & 0x7ffu), dj = fj - (int)((pb >> 11) & 0x7ffu);
            if ((unsigned)di > 1u || (unsigned)dj > 1u) continue;
            const Staged<T> q = sp[s - w0];
            const int bk = (int)(pb >> 22);
            const T wx = di ? q.fx : T(1) - q.fx;
            const T wy = dj ? q.fy : T(1) - q.fy;
            const T wxy = wx * wy;
            const T wlo = wxy * (T(1) - q.fz), whi = wxy * q.fz;
            const T vlo = wlo * q.v, vhi = whi * q.v;
            if (bk < ck) {  // levels ck-1, ck
              A0 += vlo, W0 += wlo, A1 += vhi, W1 += whi;
            } else {        // levels ck, ck+1
              A1 += vlo, W1 += wlo, A2 += vhi, W2 += whi;
            }
          }
        }
      }
      __syncthreads();
    }
    // ---- finalise the levels whose last slab this was, then roll the window
    if (COMP == 2) {
      finalize(ck, A1, W1);
      if (ck == nz - 1) finalize(ck + 1, A2, W2);
    } else {
      finalize(ck - 1, A0, W0);
      if (ck == nz - 1) finalize(ck, A1, W1);
    }
    A0 = A1, W0 = W1, A1 = A2, W1 = W2, A2 = T(0), W2 = T(0);
    __syncthreads();  // the slab's tables are rewritten by the next slab
  }
}

}  // namespace

template <class T>
bool launch_p2g_tiled(int comp, GridD<T> g, const Parts<T>& p, const uint32_t* cstart, const uint32_t* cend,
                      T* out, hipStream_t st) {
  // base indices are packed in 11 bits per axis
  if (g.nx + 1 >= 2048 || g.ny + 1 >= 2048 || g.nz + 1 >= 1024) return false;
  const int fnx = g.nx + (comp == 0), fny = g.ny + (comp == 1), fnz = g.nz + (comp == 2);
  const int64_t blocks = (int64_t)((fnx + kTX - 1) / kTX) * ((fny + kTY - 1) / kTY) *
                         ((fnz + kZChunk - 1) / kZChunk);
  if (comp == 0) k_p2g_tiled<T, 0><<<(unsigned)blocks, 256, 0, st>>>(g, p, cstart, cend, out);
  if (comp == 1) k_p2g_tiled<T, 1><<<(unsigned)blocks, 256, 0, st>>>(g, p, cstart, cend, out);
  if (comp == 2) k_p2g_tiled<T, 2><<<(unsigned)blocks, 256, 0, st>>>(g, p, cstart, cend, out);
  return true;
}

template bool launch_p2g_tiled<float>(int, GridD<float>, const Parts<float>&, const uint32_t*, const uint32_t*,
                                      float*, hipStream_t);
template bool launch_p2g_tiled<double>(int, GridD<double>, const Parts<double>&, const uint32_t*,
                                       const uint32_t*, double*, hipStream_t);

}  // namespace sdr
