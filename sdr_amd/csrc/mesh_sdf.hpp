// mesh_sdf.hpp — launch interface of the mesh signed-distance build (mesh_sdf.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdr {

// One triangle's corners, f64 (72 bytes).
struct MeshTri {
  double ax, ay, az, bx, by, bz, cx, cy, cz;
};

// Node values of the signed distance to the closed mesh `tris` on the node grid
// origin + (i, j, k) * dx, x fastest (nx * ny * nz values into out).
void launch_mesh_sdf(const MeshTri* tris, int64_t nt, int nx, int ny, int nz, double ox, double oy, double oz,
                     double dx, double* out, hipStream_t st);

}  // namespace sdr
