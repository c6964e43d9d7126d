// metrics.hip — clinical metrics and solver diagnostics on the device
// (metrics.rs:119-316, solver.rs:225-233, 758-891, particles.rs:52-58).
//
// A frame costs one pass over the particles (wet-wall marking, speed and kinetic
// energy) and one pass over the cells (coverage, fluid and membrane pressure).  Counts
// and maxima are exact; sums are per-thread sequential, then a fixed wave / block /
// block-order tree, so every result is deterministic.
#include <hip/hip_runtime.h>

#include "metrics.hpp"

namespace sdr {
namespace {

constexpr int kThreads = 256;

__device__ inline double wsum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline double wmax_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ inline unsigned long long wsum_u(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline unsigned long long wmax_u(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}

unsigned blocks_for(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  if (b > kStatBlocks) b = kStatBlocks;
  return (unsigned)(b < 1 ? 1 : b);
}

// A "wall" cell is a cavity cell with a face neighbour outside the cavity or outside
// the domain (metrics.rs:140-182).
__global__ void __launch_bounds__(kThreads) k_wall_cells(const uint8_t* __restrict__ mask, int nx, int ny,
                                                         int nz, uint8_t* __restrict__ is_wall,
                                                         unsigned long long* count) {
  const int64_t nc = (int64_t)nx * ny * nz;
  unsigned long long cnt = 0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((int64_t)nx * ny));
    uint8_t w = 0;
    if (mask[c]) {
      auto out = [&](int a, int b, int e) {
        if (a < 0 || b < 0 || e < 0 || a >= nx || b >= ny || e >= nz) return true;
        return mask[((int64_t)e * ny + b) * nx + a] == 0;
      };
      w = (out(i - 1, j, k) || out(i + 1, j, k) || out(i, j - 1, k) || out(i, j + 1, k) ||
           out(i, j, k - 1) || out(i, j, k + 1))
              ? 1
              : 0;
    }
    is_wall[c] = w;
    cnt += w;
  }
  cnt = wsum_u(cnt);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(count, cnt);
}

// record, first half (metrics.rs:197-205): wall cells holding a particle this frame are
// wet, and washed from now on.  Idempotent byte stores: any order gives the same bytes.
// Also the particle half of the frame: max |v| (solver.rs:884-890) and sum |v|^2
// (particles.rs:52-58), per block.
template <class T>
__global__ void __launch_bounds__(kThreads) k_mark_particles(Parts<T> p, int64_t n, GridD<T> g,
                                                             const uint8_t* __restrict__ is_wall,
                                                             uint8_t* __restrict__ washed,
                                                             uint8_t* __restrict__ wet,
                                                             PartStats* __restrict__ partials) {
  double vmax = 0.0, ke = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    if (is_wall) {
      const int64_t c = g.cell_of(mk3(p.px[t], p.py[t], p.pz[t]));
      if (c >= 0 && is_wall[c]) {
        washed[c] = 1;
        wet[c] = 1;
      }
    }
    // Vec3::length / length_squared (math.rs:52-63): x*x + y*y + z*z, left to right
    const double vx = (double)p.vx[t], vy = (double)p.vy[t], vz = (double)p.vz[t];
    const double l2 = vx * vx + vy * vy + vz * vz;
    vmax = fmax(vmax, sqrt(l2));
    ke += l2;
  }
  __shared__ double sm[kThreads / 64], ss[kThreads / 64];
  vmax = wmax_d(vmax);
  ke = wsum_d(ke);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = vmax, ss[threadIdx.x >> 6] = ke;
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[blockIdx.x].vmax = fmax(fmax(fmax(sm[0], sm[1]), sm[2]), sm[3]);
    partials[blockIdx.x].ke = ((ss[0] + ss[1]) + ss[2]) + ss[3];
  }
}

// record, second half (metrics.rs:206-245, 460-479; solver.rs:763-765, 859-882): per cell,
// coverage (washed), fluid pressure max / sum / count and, on the frame's wet wall
// cells, the membrane load (cells above `cap` counted as artifacts).  The wet flags are
// cleared for the next frame.
template <class T>
__global__ void __launch_bounds__(kThreads) k_cell_stats(const uint8_t* __restrict__ cells,
                                                         const T* __restrict__ pressure,
                                                         const uint8_t* __restrict__ washed,
                                                         uint8_t* __restrict__ wet, int64_t nc, double cap,
                                                         CellStats* __restrict__ partials) {
  double psum = 0.0, pmax = 0.0, wsum = 0.0, wpeak = 0.0;
  unsigned long long nwashed = 0, nfluid = 0, nart = 0, nwet = 0, nacc = 0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const double pc = (double)pressure[c];
    if (cells[c] == kFluid) {
      psum += pc;
      pmax = fmax(pmax, pc);
      ++nfluid;
    }
    if (washed) {
      nwashed += washed[c];
      if (wet[c]) {
        wet[c] = 0;
        ++nwet;
        if (pc > cap) {
          ++nart;
        } else {
          wsum += pc;
          wpeak = fmax(wpeak, pc);
          ++nacc;
        }
      }
    }
  }
  __shared__ double sd[4][kThreads / 64];
  __shared__ unsigned long long su[5][kThreads / 64];
  const int w = threadIdx.x >> 6;
  psum = wsum_d(psum), pmax = wmax_d(pmax), wsum = wsum_d(wsum), wpeak = wmax_d(wpeak);
  nwashed = wsum_u(nwashed), nfluid = wsum_u(nfluid), nart = wsum_u(nart), nwet = wsum_u(nwet), nacc = wsum_u(nacc);
  if ((threadIdx.x & 63) == 0) {
    sd[0][w] = psum, sd[1][w] = pmax, sd[2][w] = wsum, sd[3][w] = wpeak;
    su[0][w] = nwashed, su[1][w] = nfluid, su[2][w] = nart, su[3][w] = nwet, su[4][w] = nacc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    CellStats s;
    s.psum = ((sd[0][0] + sd[0][1]) + sd[0][2]) + sd[0][3];
    s.pmax = fmax(fmax(fmax(sd[1][0], sd[1][1]), sd[1][2]), sd[1][3]);
    s.wsum = ((sd[2][0] + sd[2][1]) + sd[2][2]) + sd[2][3];
    s.wpeak = fmax(fmax(fmax(sd[3][0], sd[3][1]), sd[3][2]), sd[3][3]);
    s.nwashed = su[0][0] + su[0][1] + su[0][2] + su[0][3];
    s.nfluid = su[1][0] + su[1][1] + su[1][2] + su[1][3];
    s.nart = su[2][0] + su[2][1] + su[2][2] + su[2][3];
    s.nwet = su[3][0] + su[3][1] + su[3][2] + su[3][3];
    s.nacc = su[4][0] + su[4][1] + su[4][2] + su[4][3];
    s.maxcount = 0;
    partials[blockIdx.x] = s;
  }
}

// max over cells of a per-cell particle count (solver.rs:770-782)
__global__ void __launch_bounds__(kThreads) k_max_count(const uint32_t* __restrict__ count, int64_t nc,
                                                        CellStats* __restrict__ partials) {
  unsigned long long m = 0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    m = count[c] > m ? count[c] : m;
  __shared__ unsigned long long sm[kThreads / 64];
  m = wmax_u(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long r = sm[0];
    for (int q = 1; q < kThreads / 64; ++q) r = sm[q] > r ? sm[q] : r;
    partials[blockIdx.x].maxcount = r;
  }
}

// Block partials -> one record, combined in block order by thread 0.
__global__ void k_stats_finish(const CellStats* __restrict__ cp, int ncb, const PartStats* __restrict__ pp,
                               int npb, FrameStats* out) {
  if (threadIdx.x != 0) return;
  FrameStats r{};
  for (int b = 0; b < ncb; ++b) {
    const CellStats& s = cp[b];
    r.psum += s.psum;
    r.pmax = fmax(r.pmax, s.pmax);
    r.wsum += s.wsum;
    r.wpeak = fmax(r.wpeak, s.wpeak);
    r.nwashed += s.nwashed;
    r.nfluid += s.nfluid;
    r.nart += s.nart;
    r.nwet += s.nwet;
    r.nacc += s.nacc;
    r.maxcount = s.maxcount > r.maxcount ? s.maxcount : r.maxcount;
  }
  for (int b = 0; b < npb; ++b) {
    r.vmax = fmax(r.vmax, pp[b].vmax);
    r.ke += pp[b].ke;
  }
  *out = r;
}

}  // namespace

void launch_wall_cells(const uint8_t* mask, int nx, int ny, int nz, uint8_t* is_wall, unsigned long long* count,
                       hipStream_t st) {
  k_wall_cells<<<blocks_for((int64_t)nx * ny * nz), kThreads, 0, st>>>(mask, nx, ny, nz, is_wall, count);
}

template <class T>
void launch_frame_stats(const Parts<T>& p, int64_t n, GridD<T> g, const uint8_t* cells, const T* pressure,
                        const uint8_t* is_wall, uint8_t* washed, uint8_t* wet, double cap, const uint32_t* count,
                        CellStats* cpart, PartStats* ppart, FrameStats* out, hipStream_t st) {
  const int64_t nc = (int64_t)g.nx * g.ny * g.nz;
  const unsigned pb = blocks_for(n), cb = blocks_for(nc);
  k_mark_particles<T><<<pb, kThreads, 0, st>>>(p, n, g, is_wall, washed, wet, ppart);
  k_cell_stats<T><<<cb, kThreads, 0, st>>>(cells, pressure, washed, wet, nc, cap, cpart);
  if (count) k_max_count<<<cb, kThreads, 0, st>>>(count, nc, cpart);
  k_stats_finish<<<1, 64, 0, st>>>(cpart, (int)cb, ppart, n > 0 ? (int)pb : 0, out);
}

template void launch_frame_stats<float>(const Parts<float>&, int64_t, GridD<float>, const uint8_t*, const float*,
                                        const uint8_t*, uint8_t*, uint8_t*, double, const uint32_t*, CellStats*,
                                        PartStats*, FrameStats*, hipStream_t);
template void launch_frame_stats<double>(const Parts<double>&, int64_t, GridD<double>, const uint8_t*,
                                         const double*, const uint8_t*, uint8_t*, uint8_t*, double,
                                         const uint32_t*, CellStats*, PartStats*, FrameStats*, hipStream_t);

}  // namespace sdr
