// device.hpp — device-side geometry helpers shared by the substep kernels.
//
// Every helper restates the reference arithmetic operand-for-operand (the f64 path is
// compiled with -ffp-contract=off so no a*b+c is fused), because the f64 device path
// is required to reproduce the reference bits wherever no reduction order is involved.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdr {

enum : uint8_t { kSolid = 0, kFluid = 1, kAir = 2 };

// Matrix-free 7-point Laplacian flags of one cell (pressure.rs:171-223): bit 0 = the
// cell is Fluid; bits 1..6 = the +x,-x,+y,-y,+z,-z neighbour is inside the grid and
// Fluid (the reference's plus_* coupling is -1 exactly then); bits 8..10 = diag, the
// count of non-Solid neighbours (out-of-grid counts as Solid).
enum : uint16_t {
  kFlFluid = 1u << 0,
  kFlPx = 1u << 1,
  kFlMx = 1u << 2,
  kFlPy = 1u << 3,
  kFlMy = 1u << 4,
  kFlPz = 1u << 5,
  kFlMz = 1u << 6,
};
__host__ __device__ inline int fl_diag(uint16_t f) { return (f >> 8) & 7; }

template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
__device__ inline V3<T> mk3(T x, T y, T z) {
  return V3<T>{x, y, z};
}
template <class T>
__device__ inline V3<T> operator+(V3<T> a, V3<T> b) {
  return mk3(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class T>
__device__ inline V3<T> operator-(V3<T> a, V3<T> b) {
  return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class T>
__device__ inline V3<T> operator*(V3<T> a, T s) {
  return mk3(a.x * s, a.y * s, a.z * s);
}
template <class T>
__device__ inline V3<T> operator/(V3<T> a, T s) {
  return mk3(a.x / s, a.y / s, a.z / s);
}
template <class T>
__device__ inline T dot3(V3<T> a, V3<T> b) {  // math.rs:42-44
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
__device__ inline T len3(V3<T> a) {  // math.rs:61-63
  return sqrt(dot3(a, a));
}
template <class T>
__device__ inline V3<T> normalize_or_zero(V3<T> a) {  // math.rs:73-80
  T l = len3(a);
  if (l > T(1e-12)) return a / l;
  return mk3(T(0), T(0), T(0));
}

// Rust `f64 as usize` for a value already known to be >= 0 (or NaN): saturating,
// NaN -> 0.  Returned as int64 so out-of-grid values stay out of grid.
template <class T>
__device__ inline int64_t as_index(T v) {
  if (!(v > T(0))) return 0;
  if (v >= T(4.0e18)) return INT64_MAX;
  return (int64_t)v;
}

// grid.rs:26-35 / sdf.rs:136-145 clamped base + fraction.
template <class T>
__device__ inline void clamp_base(T g, int n, int& i, T& f) {
  if (g <= T(0)) {
    i = 0;
    f = T(0);
  } else if (g >= T(n - 1)) {
    i = n - 2;
    f = T(1);
  } else {
    T fl = floor(g);
    i = (int)as_index(fl);
    f = g - fl;
  }
}

// grid.rs:18-47.
template <class T>
__device__ inline T trilinear(const T* __restrict__ a, int nx, int ny, int nz, T gx, T gy, T gz) {
  int i, j, k;
  T fx, fy, fz;
  clamp_base(gx, nx, i, fx);
  clamp_base(gy, ny, j, fy);
  clamp_base(gz, nz, k, fz);
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const T* p = a + (int64_t)k * sz + (int64_t)j * sy + i;
  T c00 = p[0] * (T(1) - fx) + p[1] * fx;
  T c10 = p[sy] * (T(1) - fx) + p[sy + 1] * fx;
  T c01 = p[sz] * (T(1) - fx) + p[sz + 1] * fx;
  T c11 = p[sz + sy] * (T(1) - fx) + p[sz + sy + 1] * fx;
  T c0 = c00 * (T(1) - fy) + c10 * fy;
  T c1 = c01 * (T(1) - fy) + c11 * fy;
  return c0 * (T(1) - fz) + c1 * fz;
}

// Trilinear sample of (a - b), with the difference taken per corner exactly as the
// reference's `diff` arrays hold it (solver.rs:954-956 then grid.rs:18-47).
template <class T>
__device__ inline T trilinear_diff(const T* __restrict__ a, const T* __restrict__ b, int nx, int ny,
                                   int nz, T gx, T gy, T gz) {
  int i, j, k;
  T fx, fy, fz;
  clamp_base(gx, nx, i, fx);
  clamp_base(gy, ny, j, fy);
  clamp_base(gz, nz, k, fz);
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const int64_t o = (int64_t)k * sz + (int64_t)j * sy + i;
  const T* p = a + o;
  const T* q = b + o;
  T c00 = (p[0] - q[0]) * (T(1) - fx) + (p[1] - q[1]) * fx;
  T c10 = (p[sy] - q[sy]) * (T(1) - fx) + (p[sy + 1] - q[sy + 1]) * fx;
  T c01 = (p[sz] - q[sz]) * (T(1) - fx) + (p[sz + 1] - q[sz + 1]) * fx;
  T c11 = (p[sz + sy] - q[sz + sy]) * (T(1) - fx) + (p[sz + sy + 1] - q[sz + sy + 1]) * fx;
  T c0 = c00 * (T(1) - fy) + c10 * fy;
  T c1 = c01 * (T(1) - fy) + c11 * fy;
  return c0 * (T(1) - fz) + c1 * fz;
}

// The eight corner values of a trilinear stencil and the offset of its base corner.
template <class T>
struct Corners {
  int64_t o;
  T c[8];  // x fastest: (0,0,0), (1,0,0), (0,1,0), (1,1,0), (0,0,1), ...
};

// trilinear(a, ...) and trilinear_diff(a, b, ...) at one point from one set of corner
// loads of a (the arithmetic of the two functions, value for value); a's corners are
// left in *keep for trilinear_again.
template <class T>
__device__ inline void trilinear_and_diff(const T* __restrict__ a, const T* __restrict__ b, int nx, int ny,
                                          int nz, T gx, T gy, T gz, T* va, T* vd, Corners<T>* keep) {
  int i, j, k;
  T fx, fy, fz;
  clamp_base(gx, nx, i, fx);
  clamp_base(gy, ny, j, fy);
  clamp_base(gz, nz, k, fz);
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const int64_t o = (int64_t)k * sz + (int64_t)j * sy + i;
  const T* p = a + o;
  const T* q = b + o;
  const T a000 = p[0], a100 = p[1], a010 = p[sy], a110 = p[sy + 1];
  const T a001 = p[sz], a101 = p[sz + 1], a011 = p[sz + sy], a111 = p[sz + sy + 1];
  keep->o = o;
  keep->c[0] = a000, keep->c[1] = a100, keep->c[2] = a010, keep->c[3] = a110;
  keep->c[4] = a001, keep->c[5] = a101, keep->c[6] = a011, keep->c[7] = a111;
  const T b000 = q[0], b100 = q[1], b010 = q[sy], b110 = q[sy + 1];
  const T b001 = q[sz], b101 = q[sz + 1], b011 = q[sz + sy], b111 = q[sz + sy + 1];
  {
    T c00 = a000 * (T(1) - fx) + a100 * fx;
    T c10 = a010 * (T(1) - fx) + a110 * fx;
    T c01 = a001 * (T(1) - fx) + a101 * fx;
    T c11 = a011 * (T(1) - fx) + a111 * fx;
    T c0 = c00 * (T(1) - fy) + c10 * fy;
    T c1 = c01 * (T(1) - fy) + c11 * fy;
    *va = c0 * (T(1) - fz) + c1 * fz;
  }
  T c00 = (a000 - b000) * (T(1) - fx) + (a100 - b100) * fx;
  T c10 = (a010 - b010) * (T(1) - fx) + (a110 - b110) * fx;
  T c01 = (a001 - b001) * (T(1) - fx) + (a101 - b101) * fx;
  T c11 = (a011 - b011) * (T(1) - fx) + (a111 - b111) * fx;
  T c0 = c00 * (T(1) - fy) + c10 * fy;
  T c1 = c01 * (T(1) - fy) + c11 * fy;
  *vd = c0 * (T(1) - fz) + c1 * fz;
}

// trilinear(a, ...) (the same arithmetic), taking the corners from `kept` when the
// stencil's base is the same and loading them otherwise.
template <class T>
__device__ inline T trilinear_again(const T* __restrict__ a, int nx, int ny, int nz, T gx, T gy, T gz,
                                    const Corners<T>& kept) {
  int i, j, k;
  T fx, fy, fz;
  clamp_base(gx, nx, i, fx);
  clamp_base(gy, ny, j, fy);
  clamp_base(gz, nz, k, fz);
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const int64_t o = (int64_t)k * sz + (int64_t)j * sy + i;
  T v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = kept.c[q];
  if (o != kept.o) {
    const T* p = a + o;
    v[0] = p[0], v[1] = p[1], v[2] = p[sy], v[3] = p[sy + 1];
    v[4] = p[sz], v[5] = p[sz + 1], v[6] = p[sz + sy], v[7] = p[sz + sy + 1];
  }
  T c00 = v[0] * (T(1) - fx) + v[1] * fx;
  T c10 = v[2] * (T(1) - fx) + v[3] * fx;
  T c01 = v[4] * (T(1) - fx) + v[5] * fx;
  T c11 = v[6] * (T(1) - fx) + v[7] * fx;
  T c0 = c00 * (T(1) - fy) + c10 * fy;
  T c1 = c01 * (T(1) - fy) + c11 * fy;
  return c0 * (T(1) - fz) + c1 * fz;
}

// Geometry of the MAC grid on the device.
template <class T>
struct GridD {
  int nx, ny, nz;
  T dx;
  T ox, oy, oz;
  __device__ V3<T> local(V3<T> p) const {  // grid.rs:163  (p - origin) / dx
    return mk3((p.x - ox) / dx, (p.y - oy) / dx, (p.z - oz) / dx);
  }
  __device__ V3<T> cell_center(int i, int j, int k) const {  // grid.rs:150-157
    return mk3(ox + (T(i) + T(0.5)) * dx, oy + (T(j) + T(0.5)) * dx, oz + (T(k) + T(0.5)) * dx);
  }
  // solver.rs:897-908: Some(cell) or -1.
  __device__ int64_t cell_of(V3<T> p) const {
    V3<T> l = local(p);
    if (l.x < T(0) || l.y < T(0) || l.z < T(0)) return -1;
    int64_t i = as_index(l.x), j = as_index(l.y), k = as_index(l.z);
    if (i < nx && j < ny && k < nz) return (k * ny + j) * nx + i;
    return -1;
  }
  // The cell whose candidate set covers every face a particle scatters to: cell_of
  // clamped into the grid (out-of-grid particles still scatter through the clamped
  // stencil, grid.rs:62-71).
  __device__ int64_t clamped_cell(V3<T> p) const {
    V3<T> l = local(p);
    int64_t i = as_index(l.x), j = as_index(l.y), k = as_index(l.z);
    i = i < nx ? i : nx - 1;
    j = j < ny ? j : ny - 1;
    k = k < nz ? k : nz - 1;
    return (k * ny + j) * nx + i;
  }
};

template <class T>
struct FacesD {
  T* u;
  T* v;
  T* w;
};
template <class T>
struct CFacesD {
  const T* u;
  const T* v;
  const T* w;
};

// grid.rs:177-193 velocity_at.
template <class T>
__device__ inline V3<T> velocity_at(const GridD<T>& g, CFacesD<T> f, V3<T> p) {
  V3<T> l = g.local(p);
  T a = trilinear(f.u, g.nx + 1, g.ny, g.nz, l.x, l.y - T(0.5), l.z - T(0.5));
  l = g.local(p);
  T b = trilinear(f.v, g.nx, g.ny + 1, g.nz, l.x - T(0.5), l.y, l.z - T(0.5));
  l = g.local(p);
  T c = trilinear(f.w, g.nx, g.ny, g.nz + 1, l.x - T(0.5), l.y - T(0.5), l.z);
  return mk3(a, b, c);
}

// sdf.rs:134-176.
template <class T>
struct SdfD {
  int nx, ny, nz;
  T dx;
  T ox, oy, oz;
  const T* data;
  __device__ T sample(V3<T> p) const {
    V3<T> l = mk3((p.x - ox) / dx, (p.y - oy) / dx, (p.z - oz) / dx);
    return trilinear(data, nx, ny, nz, l.x, l.y, l.z);
  }
  __device__ V3<T> gradient(V3<T> p) const {
    T h = dx;
    T gx = sample(p + mk3(h, T(0), T(0))) - sample(p - mk3(h, T(0), T(0)));
    T gy = sample(p + mk3(T(0), h, T(0))) - sample(p - mk3(T(0), h, T(0)));
    T gz = sample(p + mk3(T(0), T(0), h)) - sample(p - mk3(T(0), T(0), h));
    return mk3(gx, gy, gz) / (T(2) * h);
  }
  // sdf.rs:119-129 bounds, inclusive contains (math.rs:249-256).
  __device__ bool in_bounds(V3<T> p) const {
    T mx = ox + T(nx - 1) * dx, my = oy + T(ny - 1) * dx, mz = oz + T(nz - 1) * dx;
    return p.x >= ox && p.x <= mx && p.y >= oy && p.y <= my && p.z >= oz && p.z <= mz;
  }
};

template <class T>
struct BoxD {
  int on;  // 0 = no box
  T x0, y0, z0, x1, y1, z1;
  __device__ bool contains(V3<T> p) const {
    return on && p.x >= x0 && p.x <= x1 && p.y >= y0 && p.y <= y1 && p.z >= z0 && p.z <= z1;
  }
};

// solver.rs:717-732.
template <class T>
__device__ inline V3<T> project_inside(const SdfD<T>& s, T grid_dx, V3<T> p) {
  T margin = T(0.5) * grid_dx;
  for (int it = 0; it < 3; ++it) {
    T phi = s.sample(p);
    if (phi <= -margin) break;
    V3<T> g = normalize_or_zero(s.gradient(p));
    if (len3(g) == T(0)) break;
    p = p - g * (phi + margin);
  }
  return p;
}
// advect's last step (solver.rs:706-708): project_inside, then keep the projected point
// if the SDF there is <= 0, else the start.  A loop that stops on a sample has that
// sample at the point it returns, so the SDF is sampled there again only after the
// third move.
template <class T>
__device__ inline V3<T> project_inside_or_stay(const SdfD<T>& s, T grid_dx, V3<T> p, V3<T> p0) {
  T margin = T(0.5) * grid_dx;
  for (int it = 0; it < 3; ++it) {
    T phi = s.sample(p);
    if (phi <= -margin) return phi <= T(0) ? p : p0;
    V3<T> g = normalize_or_zero(s.gradient(p));
    if (len3(g) == T(0)) return phi <= T(0) ? p : p0;
    p = p - g * (phi + margin);
  }
  return s.sample(p) <= T(0) ? p : p0;
}

// solver.rs:916-930 (splitmix64 finaliser -> [-0.3, 0.3)).
__device__ inline double jitter_mix(uint64_t k) {
  uint64_t z = k + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((double)(z >> 11) / (double)(1ull << 53) - 0.5) * 0.6;
}

// Sheared pressure-solve layout.  Cell (i,j,k) lives at row J = i + j of plane k:
// index (k*NJ + J)*nx + i with NJ = nx + ny - 1.  The MIC(0) hyperplane i+j+k = s is
// then the set of whole rows with J + k = s, so the exact lexicographic sweeps walk
// contiguous rows (coalesced) while every cell still sees the reference operands.
struct Shear {
  int nx, ny, nz, NJ;
  // Work units of the vector kernels that hold in-grid slots, per plane: entry r is
  // (J << 8) | chunk for the r-th (row J, 64-slot chunk) pair with an in-grid slot, in
  // row order (built by the solver; null: every chunk of every row is a unit).
  const uint32_t* units = nullptr;
  int units_per_plane = 0;
  __host__ __device__ int64_t size() const { return (int64_t)nx * NJ * nz; }
  __host__ __device__ int64_t at(int i, int j, int k) const {
    return ((int64_t)k * NJ + (i + j)) * nx + i;
  }
};

}  // namespace sdr
