// surface.hip — particle surfacing on the device.
//
// Two stages, both deterministic and free of float atomics:
//   * the particle distance field: every node of a node grid takes the distance to the
//     nearest particle, clamped at the reach 2 r, minus the ball radius r, so the zero
//     level set bounds the union of balls of radius r around the particles.  Particles
//     are read through the solver's cell segments (cell-sorted SoA), each node visiting
//     the cells its reach overlaps; a minimum does not depend on visit order, so the
//     field is the same bits as a brute-force CPU scan in the same f64 arithmetic.
//   * naive surface nets (surface_nets.rs:48-221) over any node field: one vertex per
//     straddling cell at the mean of its edge crossings, and two triangles per
//     straddling grid edge whose four cells all carry a vertex, oriented along the
//     edge's sign change.  The sequential numbering of the reference (vertices in cell
//     order, triangles in node order with x, y, z edges per node) becomes two exclusive
//     scans: cells that straddle -> vertex slots, triangles per node -> triangle slots.
// Every per-element value is computed operand for operand as the reference does (no
// contraction), so vertices and triangles are the CPU path's bits.  The one difference
// is on the grid's max faces: a grid edge lying in the plane j = ny - 1 (k = nz - 1) has
// cells on one side only; the reference then indexes a cell outside the grid (wrapping
// into the next row, or past the end), here the ring is incomplete and emits nothing,
// as on the min faces.
// Bounds: the field is an FP64 distance scan (vector ALU plus L2-resident particle
// reads); the nets are HBM-bound passes over the node field.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device.hpp"
#include "sort.hpp"
#include "surface.hpp"

namespace sdr {
namespace {

constexpr int kBlock = 256;
constexpr int kMeasureTile = 4 * kBlock;  // triangles per block of the area/volume sums

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

using D3 = V3<double>;

__device__ inline D3 cross3(D3 a, D3 o) {  // math.rs:47-53
  return mk3(a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x);
}

// floor(v) clamped to a bin index in [0, n - 1] (NaN -> 0).
__device__ inline int bin_clamp(double v, int n) {
  const double f = floor(v);
  if (!(f >= 0.0)) return 0;
  if (f > (double)(n - 1)) return n - 1;
  return (int)f;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_particle_field(const T* __restrict__ px, const T* __restrict__ py, const T* __restrict__ pz,
                     const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend, BinGrid bg,
                     NodeGrid ng, double r, double* __restrict__ phi) {
  const int64_t nn = (int64_t)ng.nx * ng.ny * ng.nz;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nn) return;
  const int i = (int)(t % ng.nx), j = (int)((t / ng.nx) % ng.ny), k = (int)(t / ((int64_t)ng.nx * ng.ny));
  const double x = ng.ox + (double)i * ng.h, y = ng.oy + (double)j * ng.h, z = ng.oz + (double)k * ng.h;
  const double reach = 2.0 * r;
  const int i0 = bin_clamp(((x - reach) - bg.ox) / bg.dx, bg.nx), i1 = bin_clamp(((x + reach) - bg.ox) / bg.dx, bg.nx);
  const int j0 = bin_clamp(((y - reach) - bg.oy) / bg.dx, bg.ny), j1 = bin_clamp(((y + reach) - bg.oy) / bg.dx, bg.ny);
  const int k0 = bin_clamp(((z - reach) - bg.oz) / bg.dx, bg.nz), k1 = bin_clamp(((z + reach) - bg.oz) / bg.dx, bg.nz);
  double best = INFINITY;
  for (int kk = k0; kk <= k1; ++kk)
    for (int jj = j0; jj <= j1; ++jj) {
      const int64_t row = ((int64_t)kk * bg.ny + jj) * bg.nx;
      for (int ii = i0; ii <= i1; ++ii) {  // an empty cell's segment is [0, 0)
        const uint32_t a = cstart[row + ii], b = cend[row + ii];
        for (uint32_t q = a; q < b; ++q) {
          const double ex = x - (double)px[q], ey = y - (double)py[q], ez = z - (double)pz[q];
          const double d2 = ex * ex + ey * ey + ez * ez;
          best = d2 < best ? d2 : best;
        }
      }
    }
  const double d = sqrt(best);
  phi[t] = (d < reach ? d : reach) - r;
}

// Corner c of a unit cube: bit 0 = x, bit 1 = y, bit 2 = z (surface_nets.rs:16-25).
// Its 12 edges as corner pairs, x edges then y then z (surface_nets.rs:28-42).
__constant__ unsigned char kEdgeA[12] = {0, 2, 4, 6, 0, 1, 4, 5, 0, 1, 2, 3};
__constant__ unsigned char kEdgeB[12] = {1, 3, 5, 7, 2, 3, 6, 7, 4, 5, 6, 7};

__device__ inline void corners(const double* __restrict__ phi, const NodeGrid& ng, int i, int j, int k,
                               double v[8]) {
  const int64_t sx = 1, sy = ng.nx, sz = (int64_t)ng.nx * ng.ny;
  const int64_t base = (int64_t)k * sz + (int64_t)j * sy + i;
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = phi[base + (c & 1) * sx + ((c >> 1) & 1) * sy + ((c >> 2) & 1) * sz];
}

__device__ inline bool straddles(const double v[8]) {
  int neg = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) neg += v[c] < 0.0;
  return neg != 0 && neg != 8;
}

__global__ void __launch_bounds__(kBlock) k_cell_flags(const double* __restrict__ phi, NodeGrid ng,
                                                       uint32_t* __restrict__ flag) {
  const int cx = ng.nx - 1, cy = ng.ny - 1, cz = ng.nz - 1;
  const int64_t nc = (int64_t)cx * cy * cz;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t > nc) return;
  if (t == nc) {  // the scan's trailing zero: vscan[nc] is the vertex count
    flag[t] = 0;
    return;
  }
  const int i = (int)(t % cx), j = (int)((t / cx) % cy), k = (int)(t / ((int64_t)cx * cy));
  double v[8];
  corners(phi, ng, i, j, k, v);
  flag[t] = straddles(v) ? 1u : 0u;
}

// surface_nets.rs:80-118: the vertex of a straddling cell, written at its scan slot.
__global__ void __launch_bounds__(kBlock)
    k_cell_vertices(const double* __restrict__ phi, NodeGrid ng, const uint32_t* __restrict__ vscan,
                    uint32_t* __restrict__ cellv, double* __restrict__ verts) {
  const int cx = ng.nx - 1, cy = ng.ny - 1, cz = ng.nz - 1;
  const int64_t nc = (int64_t)cx * cy * cz;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nc) return;
  const int i = (int)(t % cx), j = (int)((t / cx) % cy), k = (int)(t / ((int64_t)cx * cy));
  double v[8];
  corners(phi, ng, i, j, k, v);
  if (!straddles(v)) {
    cellv[t] = 0xffffffffu;
    return;
  }
  D3 sum = mk3(0.0, 0.0, 0.0);
  int cnt = 0;
  for (int e = 0; e < 12; ++e) {
    const int a = kEdgeA[e], b = kEdgeB[e];
    const double va = v[a], vb = v[b];
    if ((va < 0.0) != (vb < 0.0)) {
      const double s = va / (va - vb);  // crossing parameter in [0, 1]
      const D3 pa = mk3((double)(a & 1), (double)((a >> 1) & 1), (double)((a >> 2) & 1));
      const D3 pb = mk3((double)(b & 1), (double)((b >> 1) & 1), (double)((b >> 2) & 1));
      sum = sum + (pa + (pb - pa) * s);
      ++cnt;
    }
  }
  const double n = (double)cnt;
  const D3 local = mk3(sum.x / n, sum.y / n, sum.z / n);
  const uint32_t slot = vscan[t];
  verts[3 * (int64_t)slot + 0] = ng.ox + ((double)i + local.x) * ng.h;
  verts[3 * (int64_t)slot + 1] = ng.oy + ((double)j + local.y) * ng.h;
  verts[3 * (int64_t)slot + 2] = ng.oz + ((double)k + local.z) * ng.h;
  cellv[t] = slot;
}

// surface_nets.rs:197-221: two triangles of a quad, wound so the geometric normal of
// (a, b, c) agrees with the desired axis direction.
__device__ inline void emit_quad(const double* __restrict__ verts, uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                 D3 desired, uint32_t* __restrict__ out) {
  const D3 pa = mk3(verts[3 * (int64_t)a], verts[3 * (int64_t)a + 1], verts[3 * (int64_t)a + 2]);
  const D3 pb = mk3(verts[3 * (int64_t)b], verts[3 * (int64_t)b + 1], verts[3 * (int64_t)b + 2]);
  const D3 pc = mk3(verts[3 * (int64_t)c], verts[3 * (int64_t)c + 1], verts[3 * (int64_t)c + 2]);
  const D3 n = cross3(pb - pa, pc - pa);
  if (dot3(n, desired) >= 0.0) {
    out[0] = a, out[1] = b, out[2] = c;
    out[3] = a, out[4] = c, out[5] = d;
  } else {
    out[0] = a, out[1] = c, out[2] = b;
    out[3] = a, out[4] = d, out[5] = c;
  }
}

// surface_nets.rs:122-192 per node: count (WRITE = false) or write (WRITE = true) the
// triangles of its x, y and z edges.  A ring cell outside the grid counts as missing.
template <bool WRITE>
__global__ void __launch_bounds__(kBlock)
    k_quads(const double* __restrict__ phi, NodeGrid ng, const uint32_t* __restrict__ cellv,
            const double* __restrict__ verts, uint32_t* __restrict__ tcount, const uint32_t* __restrict__ tscan,
            uint32_t* __restrict__ tris) {
  const int nx = ng.nx, ny = ng.ny, nz = ng.nz;
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  const int64_t nn = (int64_t)nx * ny * nz;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t > nn) return;
  if (t == nn) {
    if (!WRITE) tcount[t] = 0;  // the scan's trailing zero: tscan[nn] is the triangle count
    return;
  }
  const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
  auto get = [&](int ci, int cj, int ck) -> uint32_t {
    if (ci < 0 || cj < 0 || ck < 0 || ci >= cx || cj >= cy || ck >= cz) return 0xffffffffu;
    return cellv[((int64_t)ck * cy + cj) * cx + ci];
  };
  const double v0 = phi[t];
  uint32_t cnt = 0;
  uint32_t* out = WRITE ? tris + 3 * (int64_t)tscan[t] : nullptr;
  // desired normal: +axis when the field rises along the edge, else -axis (with +0
  // in the other components, as the reference's Vec3::new(-1.0, 0.0, 0.0))
  auto edge = [&](double v1, uint32_t r0, uint32_t r1, uint32_t r2, uint32_t r3, D3 up, D3 down) {
    if ((v0 < 0.0) == (v1 < 0.0)) return;
    if (r0 == 0xffffffffu || r1 == 0xffffffffu || r2 == 0xffffffffu || r3 == 0xffffffffu) return;
    if (WRITE) emit_quad(verts, r0, r1, r2, r3, v0 < v1 ? up : down, out + 6 * (int64_t)(cnt / 2));
    cnt += 2;
  };
  if (i + 1 < nx && j >= 1 && k >= 1)
    edge(phi[t + 1], get(i, j - 1, k - 1), get(i, j, k - 1), get(i, j, k), get(i, j - 1, k), mk3(1.0, 0.0, 0.0),
         mk3(-1.0, 0.0, 0.0));
  if (j + 1 < ny && i >= 1 && k >= 1)
    edge(phi[t + nx], get(i - 1, j, k - 1), get(i, j, k - 1), get(i, j, k), get(i - 1, j, k), mk3(0.0, 1.0, 0.0),
         mk3(0.0, -1.0, 0.0));
  if (k + 1 < nz && i >= 1 && j >= 1)
    edge(phi[t + (int64_t)nx * ny], get(i - 1, j - 1, k), get(i, j - 1, k), get(i, j, k), get(i - 1, j, k),
         mk3(0.0, 0.0, 1.0), mk3(0.0, 0.0, -1.0));
  if (!WRITE) tcount[t] = cnt;
}

// mesh.rs:52-70 per triangle, summed in a fixed tree: block b sums triangles
// [b * tile, (b + 1) * tile), four per thread in order, then pairwise in LDS.
__global__ void __launch_bounds__(kBlock) k_measure(const double* __restrict__ verts, const uint32_t* __restrict__ tris,
                                                    int64_t nt, double* __restrict__ partial) {
  __shared__ double sa[kBlock], sv[kBlock];
  double area = 0.0, vol = 0.0;
  const int64_t base = (int64_t)blockIdx.x * kMeasureTile + 4 * (int64_t)threadIdx.x;
  for (int q = 0; q < 4; ++q) {
    const int64_t f = base + q;
    if (f >= nt) break;
    const uint32_t ia = tris[3 * f], ib = tris[3 * f + 1], ic = tris[3 * f + 2];
    const D3 a = mk3(verts[3 * (int64_t)ia], verts[3 * (int64_t)ia + 1], verts[3 * (int64_t)ia + 2]);
    const D3 b = mk3(verts[3 * (int64_t)ib], verts[3 * (int64_t)ib + 1], verts[3 * (int64_t)ib + 2]);
    const D3 c = mk3(verts[3 * (int64_t)ic], verts[3 * (int64_t)ic + 1], verts[3 * (int64_t)ic + 2]);
    const D3 n = cross3(b - a, c - a);
    area += 0.5 * sqrt(dot3(n, n));
    vol += dot3(a, cross3(b, c)) / 6.0;
  }
  sa[threadIdx.x] = area, sv[threadIdx.x] = vol;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sa[threadIdx.x] += sa[threadIdx.x + w], sv[threadIdx.x] += sv[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[2 * blockIdx.x] = sa[0], partial[2 * blockIdx.x + 1] = sv[0];
}

// One block folds the per-block partials: thread t sums partials t, t + 256, ... in
// order, then pairwise in LDS.
__global__ void __launch_bounds__(kBlock) k_measure_finish(const double* __restrict__ partial, int64_t nb,
                                                           double* __restrict__ out) {
  __shared__ double sa[kBlock], sv[kBlock];
  double area = 0.0, vol = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += kBlock) area += partial[2 * b], vol += partial[2 * b + 1];
  sa[threadIdx.x] = area, sv[threadIdx.x] = vol;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sa[threadIdx.x] += sa[threadIdx.x + w], sv[threadIdx.x] += sv[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sa[0], out[1] = sv[0];
}

}  // namespace

template <class T>
void launch_particle_field(const T* px, const T* py, const T* pz, const uint32_t* cstart, const uint32_t* cend,
                           BinGrid bg, NodeGrid ng, double r, double* phi, hipStream_t st) {
  const int64_t nn = ng.nodes();
  if (nn == 0) return;
  k_particle_field<T><<<blocks_for(nn), kBlock, 0, st>>>(px, py, pz, cstart, cend, bg, ng, r, phi);
}
template void launch_particle_field<float>(const float*, const float*, const float*, const uint32_t*,
                                           const uint32_t*, BinGrid, NodeGrid, double, double*, hipStream_t);
template void launch_particle_field<double>(const double*, const double*, const double*, const uint32_t*,
                                            const uint32_t*, BinGrid, NodeGrid, double, double*, hipStream_t);

void SurfaceNets::run(const double* phi, const NodeGrid& ng, hipStream_t st) {
  if (ng.nx < 2 || ng.ny < 2 || ng.nz < 2) throw Error(SDR_E_INVALID, "surface nets need at least 2 nodes per axis");
  const int64_t nn = ng.nodes(), nc = ng.cells();
  // triangle slots (up to 6 per node) and the scans' trailing entries stay within u32
  if (nn > 700000000) throw Error(SDR_E_INVALID, "node grid too large for 32-bit mesh indices");
  cellv.ensure((size_t)(nc + 1) * 4);
  vscan.ensure((size_t)(nc + 1) * 4);
  tcount.ensure((size_t)(nn + 1) * 4);
  tscan.ensure((size_t)(nn + 1) * 4);
  temp.ensure(std::max(scan_u32_temp(nc + 1), scan_u32_temp(nn + 1)));
  host2.ensure(4 * sizeof(double));

  // vertices: flags -> slots -> positions
  k_cell_flags<<<blocks_for(nc + 1), kBlock, 0, st>>>(phi, ng, cellv.as<uint32_t>());
  HIPCHK(exclusive_sum_u32(temp.p, temp.bytes, cellv.as<uint32_t>(), vscan.as<uint32_t>(), nc + 1, st));
  uint32_t* hcount = (uint32_t*)host2.p;
  HIPCHK(hipMemcpyAsync(hcount, vscan.as<uint32_t>() + nc, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  nverts = hcount[0];
  verts.ensure((size_t)nverts * 3 * sizeof(double));
  if (nc > 0)
    k_cell_vertices<<<blocks_for(nc), kBlock, 0, st>>>(phi, ng, vscan.as<uint32_t>(), cellv.as<uint32_t>(),
                                                       verts.as<double>());

  // triangles: counts per node -> slots -> indices
  k_quads<false><<<blocks_for(nn + 1), kBlock, 0, st>>>(phi, ng, cellv.as<uint32_t>(), verts.as<double>(),
                                                        tcount.as<uint32_t>(), nullptr, nullptr);
  HIPCHK(exclusive_sum_u32(temp.p, temp.bytes, tcount.as<uint32_t>(), tscan.as<uint32_t>(), nn + 1, st));
  HIPCHK(hipMemcpyAsync(hcount, tscan.as<uint32_t>() + nn, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  ntris = hcount[0];
  tris.ensure((size_t)ntris * 3 * sizeof(uint32_t));
  if (ntris > 0)
    k_quads<true><<<blocks_for(nn), kBlock, 0, st>>>(phi, ng, cellv.as<uint32_t>(), verts.as<double>(), nullptr,
                                                     tscan.as<uint32_t>(), tris.as<uint32_t>());

  // surface area and signed volume (mesh.rs:52-70)
  const int64_t nb = ((int64_t)ntris + kMeasureTile - 1) / kMeasureTile;
  partial.ensure((size_t)(nb + 1) * 2 * sizeof(double));
  double* hm = (double*)host2.p;
  if (ntris > 0) {
    k_measure<<<(unsigned)nb, kBlock, 0, st>>>(verts.as<double>(), tris.as<uint32_t>(), (int64_t)ntris,
                                                partial.as<double>());
    k_measure_finish<<<1, kBlock, 0, st>>>(partial.as<double>(), nb, partial.as<double>() + 2 * nb);
    HIPCHK(hipMemcpyAsync(hm, partial.as<double>() + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    area = hm[0], volume = hm[1];
  } else {
    area = 0.0, volume = 0.0;
  }
}

}  // namespace sdr
