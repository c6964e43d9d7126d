// mesh_sdf.hip — `Sdf::from_mesh` (sdf.rs:47-86) on the device: the signed distance of
// every node of a grid to a closed triangle mesh (sdf.rs:187-272).
//
// One thread per node walks every triangle in mesh order, keeping the smallest squared
// point-triangle distance (Ericson's closest-point regions, sdf.rs:201-249) and summing
// the signed solid angles (van Oosterom & Strackee, sdf.rs:254-272); the node is inside
// (negative) when the generalised winding number exceeds 1/2.  A workgroup stages 256
// triangles at a time in LDS, which all its threads then read by broadcast.  The
// arithmetic follows the reference operand for operand (no contraction), so distances
// are the CPU path's bits; the solid angles use the device atan2, which can move the
// winding sum in its last bits and so only matters for a node on the surface itself.
// FP64 vector work: O(nodes x triangles), compute-bound.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "mesh_sdf.hpp"

namespace sdr {
namespace {

constexpr int kTile = 256;

using D3 = V3<double>;

__device__ inline double len2(D3 a) { return dot3(a, a); }
__device__ inline D3 cross3(D3 a, D3 o) {  // math.rs:47-53
  return mk3(a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x);
}

// sdf.rs:201-249
__device__ inline double point_triangle_distance_sq(D3 p, D3 a, D3 b, D3 c) {
  const D3 ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return len2(ap);
  const D3 bp = p - b;
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return len2(bp);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double v = d1 / (d1 - d3);
    return len2(p - (a + ab * v));
  }
  const D3 cp = p - c;
  const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return len2(cp);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    return len2(p - (a + ac * w));
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return len2(p - (b + (c - b) * w));
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  return len2(p - (a + ab * v + ac * w));
}

// sdf.rs:254-272
__device__ inline double solid_angle(D3 p, D3 a, D3 b, D3 c) {
  const D3 av = a - p, bv = b - p, cv = c - p;
  const double la = sqrt(len2(av)), lb = sqrt(len2(bv)), lc = sqrt(len2(cv));
  if (la < 1e-12 || lb < 1e-12 || lc < 1e-12) return 0.0;
  const double numerator = dot3(av, cross3(bv, cv));
  const double denominator = la * lb * lc + dot3(av, bv) * lc + dot3(bv, cv) * la + dot3(cv, av) * lb;
  return 2.0 * atan2(numerator, denominator);
}

__global__ void __launch_bounds__(kTile) k_mesh_sdf(const MeshTri* __restrict__ tris, int64_t nt, int nx, int ny,
                                                    int nz, double ox, double oy, double oz, double dx,
                                                    double* __restrict__ out) {
  __shared__ MeshTri st[kTile];
  const int64_t nn = (int64_t)nx * ny * nz;
  const int64_t t = (int64_t)blockIdx.x * kTile + threadIdx.x;
  const bool valid = t < nn;
  const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
  // origin + Vec3::new(i * dx, j * dx, k * dx) (sdf.rs:69-71)
  const D3 p = mk3(ox, oy, oz) + mk3((double)i * dx, (double)j * dx, (double)k * dx);
  double min_d2 = __builtin_inf(), winding = 0.0;
  for (int64_t c0 = 0; c0 < nt; c0 += kTile) {
    __syncthreads();
    if (c0 + threadIdx.x < nt) st[threadIdx.x] = tris[c0 + threadIdx.x];
    __syncthreads();
    const int m = nt - c0 < kTile ? (int)(nt - c0) : kTile;
    if (!valid) continue;
    for (int q = 0; q < m; ++q) {
      const MeshTri& tr = st[q];
      const D3 a = mk3(tr.ax, tr.ay, tr.az), b = mk3(tr.bx, tr.by, tr.bz), c = mk3(tr.cx, tr.cy, tr.cz);
      const double d2 = point_triangle_distance_sq(p, a, b, c);
      if (d2 < min_d2) min_d2 = d2;
      winding += solid_angle(p, a, b, c);
    }
  }
  if (!valid) return;
  winding /= 4.0 * M_PI;
  const double dist = sqrt(min_d2);
  out[t] = winding > 0.5 ? -dist : dist;
}

}  // namespace

void launch_mesh_sdf(const MeshTri* tris, int64_t nt, int nx, int ny, int nz, double ox, double oy, double oz,
                     double dx, double* out, hipStream_t st) {
  const int64_t nn = (int64_t)nx * ny * nz;
  const int64_t blocks = (nn + kTile - 1) / kTile;
  k_mesh_sdf<<<(unsigned)blocks, kTile, 0, st>>>(tris, nt, nx, ny, nz, ox, oy, oz, dx, out);
}

}  // namespace sdr
