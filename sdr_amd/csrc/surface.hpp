// surface.hpp — particle surfacing on the device: the particle distance field on a node
// grid and naive surface nets over a node field (surface.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "host.hpp"

namespace sdr {

// A regular node grid: node (i, j, k) sits at origin + (i h, j h, k h).
struct NodeGrid {
  int nx, ny, nz;
  double ox, oy, oz, h;
  int64_t nodes() const { return (int64_t)nx * ny * nz; }
  int64_t cells() const { return (int64_t)(nx - 1) * (ny - 1) * (nz - 1); }
};

// The MAC grid the particles are binned on (cell segments [cstart, cend) of the
// cell-sorted particle arrays).
struct BinGrid {
  int nx, ny, nz;
  double ox, oy, oz, dx;
};

// phi(node) = min(min_j |node - x_j|, 2 r) - r over the particles x_j: the distance to
// the union of balls of radius r, clamped at r outside.  Each node visits the cells
// that its reach 2 r overlaps.
template <class T>
void launch_particle_field(const T* px, const T* py, const T* pz, const uint32_t* cstart, const uint32_t* cend,
                           BinGrid bg, NodeGrid ng, double r, double* phi, hipStream_t st);

// Mesh of a node field's zero level set, resident on the device (naive surface nets,
// surface_nets.rs:48-221): vertices (xyz f64, one per straddling cell, in cell order)
// and triangles (u32 triples, in node order, x/y/z edges per node).
struct SurfaceNets {
  Buf cellv;   // per cell: vertex index or UINT32_MAX (flags before the scan)
  Buf vscan;   // exclusive scan of the cell flags (cells + 1)
  Buf tcount;  // per node: triangles its three edges emit (nodes + 1)
  Buf tscan;   // exclusive scan of tcount
  Buf temp;    // scan scratch
  Buf verts, tris, partial;
  PinnedBuf host2;
  uint64_t nverts = 0, ntris = 0;
  double area = 0, volume = 0;
  // Polygonise the zero set of `phi` (device, ng.nodes() values) on stream st; leaves
  // the counts, area and signed volume on the host (one synchronisation).
  void run(const double* phi, const NodeGrid& ng, hipStream_t st);
};

}  // namespace sdr
