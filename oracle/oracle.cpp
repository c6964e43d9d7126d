// oracle.cpp — CPU restatement of the reference FLIP/PIC substep (uselessgoddess/sdr).
//
// TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it; the product library (sdr_amd/libsdr_amd.so) never links, loads or calls it.
//
// Every function restates the Rust source it cites, in f64, in the reference's
// operand order (left-associative Vec3 arithmetic, true division, sequential loops in
// index order), compiled with -O2 -ffp-contract=off -fno-fast-math against glibc libm.
// Rust never contracts a*b+c into an FMA, so this reproduces the reference bits.
// The loops the reference runs on rayon (apply_a, the CG axpys, the dot's chunk
// partials and inf_norm: src/pressure.rs:372-388,402,591-608) run on OpenMP here; all
// of them are order-independent, so the thread count never changes a bit.
//
// Parity pins (tests/test_oracle_*.py): the reference's own known-answer tests
// (grid.rs:215-256, pressure.rs:619-752, solver.rs:1046-1335, sdf.rs:280-295,
// particles.rs:66-75), scipy spsolve / dense-Cholesky reductions of the PCG and MIC(0)
// on tiny grids, and numpy brute-force restatements of MIC(0), extrapolate and
// redistribute.  Functions with no such check are marked "parity unpinned" below.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>

#include "../include/sdr_amd.h"

namespace {

// ---------------------------------------------------------------- math.rs:13-194
struct V3 {
  double x, y, z;
};
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // math.rs:42-44
inline V3 cross(V3 a, V3 o) {                                                   // math.rs:47-53
  return v3(a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x);
}
inline double length(V3 a) { return std::sqrt(dot(a, a)); }  // math.rs:56-63
inline V3 normalize_or_zero(V3 a) {                           // math.rs:73-80
  double len = length(a);
  if (len > 1e-12) return a / len;
  return v3(0, 0, 0);
}
// Rust f64::max / f64::min ignore a NaN operand, as fmax / fmin do.
inline double rmax(double a, double b) { return std::fmax(a, b); }
inline double rmin(double a, double b) { return std::fmin(a, b); }
// Rust `f64 as usize`: saturating, NaN -> 0.
inline uint64_t as_usize(double v) {
  if (!(v > 0.0)) return 0;
  if (v >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)v;
}

struct Aabb {  // math.rs:198-257
  V3 min, max;
  bool contains(V3 p) const {
    return p.x >= min.x && p.x <= max.x && p.y >= min.y && p.y <= max.y && p.z >= min.z &&
           p.z <= max.z;
  }
};

// ---------------------------------------------------------------- grid.rs:18-88
struct Base {
  uint64_t i;
  double f;
};
inline Base clamp_base(double g, uint64_t n) {  // grid.rs:26-35
  if (g <= 0.0) return Base{0, 0.0};
  if (g >= (double)(n - 1)) return Base{n - 2, 1.0};
  double i = std::floor(g);
  return Base{as_usize(i), g - i};
}

struct Dims {
  uint64_t nx, ny, nz;
};

double trilinear_sample(const double* arr, Dims d, double gx, double gy, double gz) {
  Base bi = clamp_base(gx, d.nx), bj = clamp_base(gy, d.ny), bk = clamp_base(gz, d.nz);
  uint64_t i = bi.i, j = bj.i, k = bk.i;
  double fx = bi.f, fy = bj.f, fz = bk.f;
  auto at = [&](uint64_t a, uint64_t b, uint64_t c) { return arr[(c * d.ny + b) * d.nx + a]; };
  double c00 = at(i, j, k) * (1.0 - fx) + at(i + 1, j, k) * fx;
  double c10 = at(i, j + 1, k) * (1.0 - fx) + at(i + 1, j + 1, k) * fx;
  double c01 = at(i, j, k + 1) * (1.0 - fx) + at(i + 1, j, k + 1) * fx;
  double c11 = at(i, j + 1, k + 1) * (1.0 - fx) + at(i + 1, j + 1, k + 1) * fx;
  double c0 = c00 * (1.0 - fy) + c10 * fy;
  double c1 = c01 * (1.0 - fy) + c11 * fy;
  return c0 * (1.0 - fz) + c1 * fz;
}

void trilinear_scatter(double* arr, double* wsum, Dims d, double gx, double gy, double gz,
                       double val) {  // grid.rs:52-88
  Base bi = clamp_base(gx, d.nx), bj = clamp_base(gy, d.ny), bk = clamp_base(gz, d.nz);
  double wx[2] = {1.0 - bi.f, bi.f}, wy[2] = {1.0 - bj.f, bj.f}, wz[2] = {1.0 - bk.f, bk.f};
  for (int dk = 0; dk < 2; ++dk)
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di) {
        double w = wx[di] * wy[dj] * wz[dk];
        uint64_t idx = ((bk.i + dk) * d.ny + (bj.i + dj)) * d.nx + (bi.i + di);
        arr[idx] += w * val;
        wsum[idx] += w;
      }
}

// ---------------------------------------------------------------- grid.rs:92-206
struct MacGrid {
  uint64_t nx = 0, ny = 0, nz = 0;
  double dx = 0;
  V3 origin{0, 0, 0};
  std::vector<double> u, v, w;
  void init(uint64_t x, uint64_t y, uint64_t z, double h, V3 o) {
    nx = x, ny = y, nz = z, dx = h, origin = o;
    u.assign((nx + 1) * ny * nz, 0.0);
    v.assign(nx * (ny + 1) * nz, 0.0);
    w.assign(nx * ny * (nz + 1), 0.0);
  }
  Dims ud() const { return Dims{nx + 1, ny, nz}; }
  Dims vd() const { return Dims{nx, ny + 1, nz}; }
  Dims wd() const { return Dims{nx, ny, nz + 1}; }
  uint64_t u_idx(uint64_t i, uint64_t j, uint64_t k) const { return (k * ny + j) * (nx + 1) + i; }
  uint64_t v_idx(uint64_t i, uint64_t j, uint64_t k) const { return (k * (ny + 1) + j) * nx + i; }
  uint64_t w_idx(uint64_t i, uint64_t j, uint64_t k) const { return (k * ny + j) * nx + i; }
  V3 cell_center(uint64_t i, uint64_t j, uint64_t k) const {  // grid.rs:150-157
    return origin + v3(((double)i + 0.5) * dx, ((double)j + 0.5) * dx, ((double)k + 0.5) * dx);
  }
  V3 local(V3 p) const { return (p - origin) / dx; }  // grid.rs:163
  double sample_u(V3 p) const {
    V3 l = local(p);
    return trilinear_sample(u.data(), ud(), l.x, l.y - 0.5, l.z - 0.5);
  }
  double sample_v(V3 p) const {
    V3 l = local(p);
    return trilinear_sample(v.data(), vd(), l.x - 0.5, l.y, l.z - 0.5);
  }
  double sample_w(V3 p) const {
    V3 l = local(p);
    return trilinear_sample(w.data(), wd(), l.x - 0.5, l.y - 0.5, l.z);
  }
  V3 velocity_at(V3 p) const { return v3(sample_u(p), sample_v(p), sample_w(p)); }
  double max_speed() const {  // grid.rs:202-206
    auto m = [](const std::vector<double>& a) {
      double r = 0.0;
      for (double x : a) r = rmax(r, std::fabs(x));
      return r;
    };
    double a = m(u), b = m(v), c = m(w);
    return std::sqrt(a * a + b * b + c * c);
  }
};

// ---------------------------------------------------------------- sdf.rs:23-182
struct Sdf {
  V3 origin{0, 0, 0};
  double dx = 0;
  uint64_t nx = 0, ny = 0, nz = 0;
  std::vector<double> data;
  Aabb bounds() const {  // sdf.rs:119-129
    return Aabb{origin, origin + v3((double)(nx - 1) * dx, (double)(ny - 1) * dx,
                                    (double)(nz - 1) * dx)};
  }
  double sample(V3 p) const {  // sdf.rs:134-166
    V3 l = (p - origin) / dx;
    Base bi = clamp_base(l.x, nx), bj = clamp_base(l.y, ny), bk = clamp_base(l.z, nz);
    uint64_t i = bi.i, j = bj.i, k = bk.i;
    double fx = bi.f, fy = bj.f, fz = bk.f;
    auto at = [&](uint64_t a, uint64_t b, uint64_t c) { return data[(c * ny + b) * nx + a]; };
    double c000 = at(i, j, k), c100 = at(i + 1, j, k), c010 = at(i, j + 1, k),
           c110 = at(i + 1, j + 1, k), c001 = at(i, j, k + 1), c101 = at(i + 1, j, k + 1),
           c011 = at(i, j + 1, k + 1), c111 = at(i + 1, j + 1, k + 1);
    double c00 = c000 * (1.0 - fx) + c100 * fx;
    double c10 = c010 * (1.0 - fx) + c110 * fx;
    double c01 = c001 * (1.0 - fx) + c101 * fx;
    double c11 = c011 * (1.0 - fx) + c111 * fx;
    double c0 = c00 * (1.0 - fy) + c10 * fy;
    double c1 = c01 * (1.0 - fy) + c11 * fy;
    return c0 * (1.0 - fz) + c1 * fz;
  }
  V3 gradient(V3 p) const {  // sdf.rs:170-176
    double h = dx;
    double gx = sample(p + v3(h, 0, 0)) - sample(p - v3(h, 0, 0));
    double gy = sample(p + v3(0, h, 0)) - sample(p - v3(0, h, 0));
    double gz = sample(p + v3(0, 0, h)) - sample(p - v3(0, 0, h));
    return v3(gx, gy, gz) / (2.0 * h);
  }
};

// ---------------------------------------------------------------- pressure.rs
enum : uint8_t { SOLID = 0, FLUID = 1, AIR = 2 };

struct Laplacian {  // pressure.rs:62-77
  std::vector<double> diag, plus_x, plus_y, plus_z;
  uint64_t nx, ny, nz;
  uint64_t idx(uint64_t i, uint64_t j, uint64_t k) const { return (k * ny + j) * nx + i; }
};

void enforce_solid_faces(MacGrid& g, const uint8_t* cells) {  // pressure.rs:243-281
  const int64_t nx = g.nx, ny = g.ny, nz = g.nz;
  auto solid = [&](int64_t i, int64_t j, int64_t k) {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return true;
    return cells[(k * ny + j) * nx + i] == SOLID;
  };
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i <= nx; ++i)
        if (solid(i - 1, j, k) || solid(i, j, k)) g.u[g.u_idx(i, j, k)] = 0.0;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j <= ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        if (solid(i, j - 1, k) || solid(i, j, k)) g.v[g.v_idx(i, j, k)] = 0.0;
  for (int64_t k = 0; k <= nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        if (solid(i, j, k - 1) || solid(i, j, k)) g.w[g.w_idx(i, j, k)] = 0.0;
}

void apply_pressure_gradient(MacGrid& g, const uint8_t* cells, const double* p,
                             double coef) {  // pressure.rs:285-332
  const uint64_t nx = g.nx, ny = g.ny, nz = g.nz;
  auto at = [&](uint64_t i, uint64_t j, uint64_t k) { return cells[(k * ny + j) * nx + i]; };
  auto pr = [&](uint64_t i, uint64_t j, uint64_t k) { return p[(k * ny + j) * nx + i]; };
  for (uint64_t k = 0; k < nz; ++k)
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 1; i < nx; ++i) {
        uint8_t l = at(i - 1, j, k), r = at(i, j, k);
        if (l == SOLID || r == SOLID) continue;
        if (l == FLUID || r == FLUID) g.u[g.u_idx(i, j, k)] -= coef * (pr(i, j, k) - pr(i - 1, j, k));
      }
  for (uint64_t k = 0; k < nz; ++k)
    for (uint64_t j = 1; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        uint8_t l = at(i, j - 1, k), r = at(i, j, k);
        if (l == SOLID || r == SOLID) continue;
        if (l == FLUID || r == FLUID) g.v[g.v_idx(i, j, k)] -= coef * (pr(i, j, k) - pr(i, j - 1, k));
      }
  for (uint64_t k = 1; k < nz; ++k)
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        uint8_t l = at(i, j, k - 1), r = at(i, j, k);
        if (l == SOLID || r == SOLID) continue;
        if (l == FLUID || r == FLUID) g.w[g.w_idx(i, j, k)] -= coef * (pr(i, j, k) - pr(i, j, k - 1));
      }
}

void apply_a(const Laplacian& a, const uint8_t* cells, const double* x, double* out) {
  // pressure.rs:400-439 (rayon over z-slices)
  const uint64_t nx = a.nx, ny = a.ny, nz = a.nz;
#pragma omp parallel for schedule(static)
  for (int64_t kk = 0; kk < (int64_t)nz; ++kk) {
    uint64_t k = (uint64_t)kk;
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        uint64_t c = (k * ny + j) * nx + i;
        if (cells[c] != FLUID) {
          out[c] = 0.0;
          continue;
        }
        double v = a.diag[c] * x[c];
        if (i + 1 < nx) v += a.plus_x[c] * x[a.idx(i + 1, j, k)];
        if (i >= 1) v += a.plus_x[a.idx(i - 1, j, k)] * x[a.idx(i - 1, j, k)];
        if (j + 1 < ny) v += a.plus_y[c] * x[a.idx(i, j + 1, k)];
        if (j >= 1) v += a.plus_y[a.idx(i, j - 1, k)] * x[a.idx(i, j - 1, k)];
        if (k + 1 < nz) v += a.plus_z[c] * x[a.idx(i, j, k + 1)];
        if (k >= 1) v += a.plus_z[a.idx(i, j, k - 1)] * x[a.idx(i, j, k - 1)];
        out[c] = v;
      }
  }
}

std::vector<double> build_mic0(const Laplacian& a, const uint8_t* cells,
                               double TAU = 0.97) {  // pressure.rs:442-513
  const double SIGMA = 0.25;
  const uint64_t nx = a.nx, ny = a.ny, nz = a.nz;
  std::vector<double> precon(nx * ny * nz, 0.0);
  for (uint64_t k = 0; k < nz; ++k)
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        uint64_t c = (k * ny + j) * nx + i;
        if (cells[c] != FLUID) continue;
        double axi = 0, pxi = 0, ayj = 0, pyj = 0, azk = 0, pzk = 0;
        if (i >= 1) { uint64_t cm = a.idx(i - 1, j, k); axi = a.plus_x[cm]; pxi = precon[cm]; }
        if (j >= 1) { uint64_t cm = a.idx(i, j - 1, k); ayj = a.plus_y[cm]; pyj = precon[cm]; }
        if (k >= 1) { uint64_t cm = a.idx(i, j, k - 1); azk = a.plus_z[cm]; pzk = precon[cm]; }
        double term_x = axi * pxi, term_y = ayj * pyj, term_z = azk * pzk;
        double axi_y = 0, axi_z = 0, ayj_x = 0, ayj_z = 0, azk_x = 0, azk_y = 0;
        if (i >= 1) { uint64_t cm = a.idx(i - 1, j, k); axi_y = a.plus_y[cm]; axi_z = a.plus_z[cm]; }
        if (j >= 1) { uint64_t cm = a.idx(i, j - 1, k); ayj_x = a.plus_x[cm]; ayj_z = a.plus_z[cm]; }
        if (k >= 1) { uint64_t cm = a.idx(i, j, k - 1); azk_x = a.plus_x[cm]; azk_y = a.plus_y[cm]; }
        double e = a.diag[c] - term_x * term_x - term_y * term_y - term_z * term_z -
                   TAU * (axi * (axi_y + axi_z) * pxi * pxi + ayj * (ayj_x + ayj_z) * pyj * pyj +
                          azk * (azk_x + azk_y) * pzk * pzk);
        if (e < SIGMA * a.diag[c]) e = a.diag[c];
        precon[c] = e > 0.0 ? 1.0 / std::sqrt(e) : 0.0;
      }
  return precon;
}

void apply_mic0(const Laplacian& a, const double* precon, const uint8_t* cells, const double* r,
                double* z) {  // pressure.rs:517-570
  const uint64_t nx = a.nx, ny = a.ny, nz = a.nz;
  std::vector<double> q(nx * ny * nz, 0.0);
  for (uint64_t k = 0; k < nz; ++k)
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        uint64_t c = (k * ny + j) * nx + i;
        if (cells[c] != FLUID) continue;
        double t = r[c];
        if (i >= 1) { uint64_t cm = a.idx(i - 1, j, k); t -= a.plus_x[cm] * precon[cm] * q[cm]; }
        if (j >= 1) { uint64_t cm = a.idx(i, j - 1, k); t -= a.plus_y[cm] * precon[cm] * q[cm]; }
        if (k >= 1) { uint64_t cm = a.idx(i, j, k - 1); t -= a.plus_z[cm] * precon[cm] * q[cm]; }
        q[c] = t * precon[c];
      }
  for (uint64_t k = nz; k-- > 0;)
    for (uint64_t j = ny; j-- > 0;)
      for (uint64_t i = nx; i-- > 0;) {
        uint64_t c = (k * ny + j) * nx + i;
        if (cells[c] != FLUID) {
          z[c] = 0.0;
          continue;
        }
        double t = q[c];
        if (i + 1 < nx) t -= a.plus_x[c] * precon[c] * z[a.idx(i + 1, j, k)];
        if (j + 1 < ny) t -= a.plus_y[c] * precon[c] * z[a.idx(i, j + 1, k)];
        if (k + 1 < nz) t -= a.plus_z[c] * precon[c] * z[a.idx(i, j, k + 1)];
        z[c] = t * precon[c];
      }
}

// pressure.rs:588-603: 4096-element chunks summed sequentially, partials summed in
// chunk order by `Iterator::sum` (which folds from -0.0 in current Rust; the start
// value only differs from 0.0 on an all-negative-zero sum).
double dot_chunked(const double* a, const double* b, uint64_t n) {
  const uint64_t CHUNK = 4096;
  uint64_t nch = (n + CHUNK - 1) / CHUNK;
  std::vector<double> partials(nch);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < (int64_t)nch; ++c) {
    uint64_t lo = (uint64_t)c * CHUNK, hi = lo + CHUNK < n ? lo + CHUNK : n;
    double acc = 0.0;
    for (uint64_t t = lo; t < hi; ++t) acc += a[t] * b[t];
    partials[c] = acc;
  }
  double s = -0.0;
  for (uint64_t c = 0; c < nch; ++c) s += partials[c];
  return s;
}

double inf_norm(const double* a, uint64_t n) {  // pressure.rs:605-609 (max is order-free)
  double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
  for (int64_t t = 0; t < (int64_t)n; ++t) m = rmax(m, std::fabs(a[t]));
  return m;
}

struct PcgResult {
  std::vector<double> pressure;
  uint64_t iters;
  double residual;
};

PcgResult pcg(const Laplacian& a, const uint8_t* cells, const std::vector<double>& b,
              uint64_t max_iters, double tol) {  // pressure.rs:341-397
  const uint64_t n = b.size();
  std::vector<double> p(n, 0.0);
  double residual = inf_norm(b.data(), n);
  if (residual <= tol) return PcgResult{p, 0, residual};
  std::vector<double> precon = build_mic0(a, cells);
  std::vector<double> r = b, z(n, 0.0);
  apply_mic0(a, precon.data(), cells, r.data(), z.data());
  std::vector<double> s = z;
  double sigma = dot_chunked(z.data(), r.data(), n);
  std::vector<double> as(n, 0.0);
  uint64_t iters = 0;
  for (uint64_t it = 0; it < max_iters; ++it) {
    iters = it + 1;
    apply_a(a, cells, s.data(), as.data());
    double denom = dot_chunked(s.data(), as.data(), n);
    if (std::fabs(denom) < 1e-30) break;
    double alpha = sigma / denom;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)n; ++t) p[t] += alpha * s[t];
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)n; ++t) r[t] -= alpha * as[t];
    residual = inf_norm(r.data(), n);
    if (residual <= tol) break;
    apply_mic0(a, precon.data(), cells, r.data(), z.data());
    double sigma_new = dot_chunked(z.data(), r.data(), n);
    double beta = sigma_new / sigma;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)n; ++t) s[t] = z[t] + beta * s[t];
    sigma = sigma_new;
  }
  return PcgResult{p, iters, residual};
}

Laplacian assemble(const MacGrid& g, const uint8_t* cells) {  // pressure.rs:171-223
  const uint64_t nx = g.nx, ny = g.ny, nz = g.nz, n = nx * ny * nz;
  Laplacian a{std::vector<double>(n, 0.0), std::vector<double>(n, 0.0),
              std::vector<double>(n, 0.0), std::vector<double>(n, 0.0), nx, ny, nz};
  for (uint64_t k = 0; k < nz; ++k)
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        uint64_t c = (k * ny + j) * nx + i;
        if (cells[c] != FLUID) continue;
        double diag = 0.0;
        auto consider = [&](int64_t ni, int64_t nj, int64_t nk, double* plus) {
          bool inside = ni >= 0 && nj >= 0 && nk >= 0 && (uint64_t)ni < nx && (uint64_t)nj < ny &&
                        (uint64_t)nk < nz;
          uint8_t kind = inside ? cells[((uint64_t)nk * ny + (uint64_t)nj) * nx + (uint64_t)ni] : SOLID;
          if (kind == AIR) diag += 1.0;
          else if (kind == FLUID) {
            diag += 1.0;
            if (plus) *plus = -1.0;
          }
        };
        int64_t I = i, J = j, K = k;
        consider(I - 1, J, K, nullptr);
        consider(I + 1, J, K, &a.plus_x[c]);
        consider(I, J - 1, K, nullptr);
        consider(I, J + 1, K, &a.plus_y[c]);
        consider(I, J, K - 1, nullptr);
        consider(I, J, K + 1, &a.plus_z[c]);
        a.diag[c] = diag;
      }
  return a;
}

// pressure.rs:123-239
sdr_solve_report project_capturing_with_source(MacGrid& g, const uint8_t* cells, double rho,
                                               double dt, uint64_t max_iters, double tol,
                                               const double* source, std::vector<double>& pout) {
  const uint64_t nx = g.nx, ny = g.ny, nz = g.nz, n = nx * ny * nz;
  enforce_solid_faces(g, cells);
  double inv_scale = rho * g.dx / dt;
  std::vector<double> b(n, 0.0);
  for (uint64_t k = 0; k < nz; ++k)
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        uint64_t c = (k * ny + j) * nx + i;
        if (cells[c] != FLUID) continue;
        double div = (g.u[g.u_idx(i + 1, j, k)] - g.u[g.u_idx(i, j, k)]) +
                     (g.v[g.v_idx(i, j + 1, k)] - g.v[g.v_idx(i, j, k)]) +
                     (g.w[g.w_idx(i, j, k + 1)] - g.w[g.w_idx(i, j, k)]);
        double s = source ? source[c] : 0.0;
        b[c] = -inv_scale * (div - s);
      }
  Laplacian a = assemble(g, cells);
  PcgResult rep = pcg(a, cells, b, max_iters, tol);
  double coef = dt / (rho * g.dx);
  apply_pressure_gradient(g, cells, rep.pressure.data(), coef);
  pout = std::move(rep.pressure);
  return sdr_solve_report{rep.iters, rep.residual};
}

// ---------------------------------------------------------------- solver.rs
struct Rng {  // solver.rs:120-140
  uint64_t s;
  uint64_t next_u64() {
    uint64_t x = s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    s = x;
    return x;
  }
  double unit() { return (double)(next_u64() >> 11) / (double)(1ull << 53); }
};

V3 hash_jitter(uint64_t key) {  // solver.rs:916-930
  auto mix = [](uint64_t k) {
    uint64_t z = k + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) / (double)(1ull << 53) - 0.5) * 0.6;
  };
  return v3(mix(key * 3), mix(key * 3 + 1), mix(key * 3 + 2));
}

void orthonormal_basis(V3 n, V3& e1, V3& e2) {  // solver.rs:933-942
  V3 a = std::fabs(n.x) < 0.9 ? v3(1, 0, 0) : v3(0, 1, 0);
  e1 = normalize_or_zero(cross(n, a));
  e2 = normalize_or_zero(cross(n, e1));
}

struct Fluid {
  double density;
  V3 gravity;
  double flip_ratio;
  uint64_t ppc;
  double volume_stiffness;
  bool redistribute;
  double redist_cap_factor;
};
struct Needle {
  V3 tip, axis;
  double radius, flow_rate, inject_until;
  double jet_speed() const {  // solver.rs:107-114
    double area = M_PI * radius * radius;
    return area > 0.0 ? flow_rate / area : 0.0;
  }
};

struct Solver {
  MacGrid grid;
  std::vector<V3> pos, vel;
  Sdf solid;
  Fluid fluid;
  Needle needle;
  bool has_outlet = false;
  Aabb outlet{};
  uint64_t max_iters = 200;
  double tol = 1e-6;
  double time = 0;
  uint64_t steps = 0, last_pcg_iters = 0;
  uint32_t last_substeps = 0;
  double last_residual = 0;
  std::vector<double> pressure;
  uint64_t drained = 0, emitted = 0, culled = 0;
  std::vector<uint8_t> cells, cavity_mask;
  std::vector<double> saved_u, saved_v, saved_w;
  std::vector<double> source;
  Rng rng{0x5deece66dull | 1};
  double emit_accum = 0;
  uint64_t redist_active = 0, redist_moved = 0;  // diagnostics of the last redistribute
  // the last particle surface (orc_solver_surface)
  std::vector<double> surf_phi, surf_v;
  std::vector<uint32_t> surf_t;
  sdr_surface_info surf_info{};
  bool have_surface = false;

  void create(Sdf s, double sim_dx, Fluid f, Needle nd) {  // solver.rs:184-222
    solid = std::move(s);
    fluid = f;
    needle = nd;
    Aabb bb = solid.bounds();
    V3 size = bb.max - bb.min;
    uint64_t nx = as_usize(rmax(std::ceil(size.x / sim_dx), 1.0));
    uint64_t ny = as_usize(rmax(std::ceil(size.y / sim_dx), 1.0));
    uint64_t nz = as_usize(rmax(std::ceil(size.z / sim_dx), 1.0));
    grid.init(nx, ny, nz, sim_dx, bb.min);
    uint64_t n = nx * ny * nz;
    cavity_mask.resize(n);
    for (uint64_t c = 0; c < n; ++c) {
      uint64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
      cavity_mask[c] = solid.sample(grid.cell_center(i, j, k)) < 0.0;
    }
    pressure.assign(n, 0.0);
    cells.assign(n, AIR);
    source.assign(n, 0.0);
  }

  double particle_volume() const {  // solver.rs:225-228
    double cell = grid.dx * grid.dx * grid.dx;
    uint64_t p = fluid.ppc > 1 ? fluid.ppc : 1;
    return cell / (double)p;
  }

  bool cell_of(V3 p, uint64_t& i, uint64_t& j, uint64_t& k) const {  // solver.rs:897-908
    V3 l = (p - grid.origin) / grid.dx;
    if (l.x < 0.0 || l.y < 0.0 || l.z < 0.0) return false;
    i = as_usize(l.x), j = as_usize(l.y), k = as_usize(l.z);
    return i < grid.nx && j < grid.ny && k < grid.nz;
  }
  bool in_outlet(V3 p) const { return has_outlet && outlet.contains(p); }

  double cfl_dt() const {  // solver.rs:262-281
    double grid_speed = grid.max_speed();
    double jet = time < needle.inject_until ? needle.jet_speed() : 0.0;
    double g = std::sqrt(length(fluid.gravity) * grid.dx);
    double speed = rmax(rmax(rmax(grid_speed, jet), g), 1e-6);
    return rmin(grid.dx / speed, 0.01);
  }

  void step(double dt) {  // solver.rs:237-259
    const uint32_t MAX_SUBSTEPS = 512;
    double floor_dt = dt / (double)MAX_SUBSTEPS;
    double remaining = dt;
    uint32_t taken = 0;
    while (remaining > 1e-12) {
      double sub = rmin(rmax(cfl_dt(), floor_dt), remaining);
      substep(sub);
      remaining -= sub;
      taken += 1;
    }
    last_substeps = taken;
  }

  void run_stage(int stage, double dt) {
    switch (stage) {
      case SDR_STAGE_EMIT: emit(dt); break;
      case SDR_STAGE_CLASSIFY: classify(); break;
      case SDR_STAGE_P2G: p2g(); break;
      case SDR_STAGE_SAVE: saved_u = grid.u, saved_v = grid.v, saved_w = grid.w; break;
      case SDR_STAGE_FORCES: add_forces(dt); break;
      case SDR_STAGE_PROJECT: {
        volume_source(dt);
        sdr_solve_report r = project_capturing_with_source(grid, cells.data(), fluid.density, dt,
                                                           max_iters, tol, source.data(), pressure);
        last_pcg_iters = r.iters;
        last_residual = r.residual;
      } break;
      case SDR_STAGE_EXTRAPOLATE: extrapolate_velocities(); break;
      case SDR_STAGE_G2P: g2p(); break;
      case SDR_STAGE_ADVECT: advect(dt); break;
      case SDR_STAGE_BOUNDS: enforce_particle_bounds(); break;
      case SDR_STAGE_REDISTRIBUTE:
        if (fluid.redistribute) redistribute();
        break;
      case SDR_STAGE_FINISH: time += dt; steps += 1; break;
    }
  }

  void substep(double dt) {  // solver.rs:284-310
    for (int s = 0; s < SDR_STAGE_COUNT; ++s) run_stage(s, dt);
  }

  void emit(double dt) {  // solver.rs:313-339
    double pv = particle_volume();
    if (pv <= 0.0 || needle.flow_rate <= 0.0 || time >= needle.inject_until) return;
    emit_accum += needle.flow_rate * dt / pv;
    double count = std::floor(emit_accum);
    uint64_t n = as_usize(count);
    emit_accum -= count;
    if (n == 0) return;
    V3 axis = normalize_or_zero(needle.axis);
    V3 e1, e2;
    orthonormal_basis(axis, e1, e2);
    double speed = needle.jet_speed();
    V3 v = axis * speed;
    for (uint64_t t = 0; t < n; ++t) {
      double r = needle.radius * std::sqrt(rng.unit());
      double theta = 2.0 * M_PI * rng.unit();
      double along = (rng.unit() - 0.5) * grid.dx;
      V3 p = needle.tip + e1 * (r * std::cos(theta)) + e2 * (r * std::sin(theta)) + axis * along;
      pos.push_back(p);
      vel.push_back(v);
    }
    emitted += n;
  }

  void classify() {  // solver.rs:342-371
    const uint64_t nx = grid.nx, ny = grid.ny, nz = grid.nz;
    for (uint64_t k = 0; k < nz; ++k)
      for (uint64_t j = 0; j < ny; ++j)
        for (uint64_t i = 0; i < nx; ++i) {
          uint64_t c = (k * ny + j) * nx + i;
          bool out = has_outlet && outlet.contains(grid.cell_center(i, j, k));
          cells[c] = out ? AIR : (!cavity_mask[c] ? SOLID : AIR);
        }
    for (const V3& p : pos) {
      uint64_t i, j, k;
      if (cell_of(p, i, j, k)) {
        uint64_t c = (k * ny + j) * nx + i;
        if (cells[c] == AIR && !in_outlet(p)) cells[c] = FLUID;
      }
    }
  }

  void volume_source(double dt) {  // solver.rs:392-423
    const uint64_t nx = grid.nx, ny = grid.ny, nz = grid.nz, n = nx * ny * nz;
    source.assign(n, 0.0);
    double k_vol = fluid.volume_stiffness;
    if (k_vol <= 0.0 || dt <= 0.0) return;
    std::vector<uint32_t> count(n, 0);
    for (const V3& p : pos) {
      uint64_t i, j, k;
      if (cell_of(p, i, j, k)) {
        uint64_t c = (k * ny + j) * nx + i;
        if (cells[c] == FLUID) count[c] += 1;
      }
    }
    double rest = (double)(fluid.ppc > 1 ? fluid.ppc : 1);
    double unit = grid.dx / dt;
    double cap = 0.25 * unit;
    for (uint64_t c = 0; c < n; ++c) {
      if (cells[c] != FLUID) continue;
      double over = (double)count[c] / rest - 1.0;
      if (over > 0.0) source[c] = rmin(k_vol * over * unit, cap);
    }
  }

  V3 project_inside(V3 p) const {  // solver.rs:717-732
    double margin = 0.5 * grid.dx;
    for (int it = 0; it < 3; ++it) {
      double phi = solid.sample(p);
      if (phi <= -margin) break;
      V3 g = normalize_or_zero(solid.gradient(p));
      if (length(g) == 0.0) break;
      p = p - g * (phi + margin);
    }
    return p;
  }

  void redistribute() {  // solver.rs:454-557
    const uint64_t nx = grid.nx, ny = grid.ny, nz = grid.nz, ncells = nx * ny * nz;
    uint64_t target = fluid.ppc > 1 ? fluid.ppc : 1;
    uint32_t cap = (uint32_t)as_usize(std::ceil(rmax(fluid.redist_cap_factor, 1.0) * (double)target));
    V3 gdir = normalize_or_zero(fluid.gravity);
    std::vector<uint32_t> count(ncells, 0);
    redist_active = redist_moved = 0;
    for (int sweep = 0; sweep < 4; ++sweep) {
      std::fill(count.begin(), count.end(), 0u);
      for (const V3& p : pos) {
        uint64_t i, j, k;
        if (cell_of(p, i, j, k)) count[(k * ny + j) * nx + i] += 1;
      }
      for (const V3& p : pos) {  // diagnostics: particles in over-cap cells
        uint64_t i, j, k;
        if (cell_of(p, i, j, k) && count[(k * ny + j) * nx + i] > cap) redist_active++;
      }
      for (uint64_t idx = 0; idx < pos.size(); ++idx) {
        uint64_t i, j, k;
        if (!cell_of(pos[idx], i, j, k)) continue;
        uint64_t c = (k * ny + j) * nx + i;
        if (count[c] <= cap) continue;
        uint64_t nc_list[6];
        double down_list[6];
        int nn = 0;
        auto add = [&](uint64_t nc, V3 dir) {
          nc_list[nn] = nc;
          down_list[nn] = dot(dir, gdir);
          nn++;
        };
        if (i > 0) add(c - 1, v3(-1, 0, 0));
        if (i + 1 < nx) add(c + 1, v3(1, 0, 0));
        if (j > 0) add(c - nx, v3(0, -1, 0));
        if (j + 1 < ny) add(c + nx, v3(0, 1, 0));
        if (k > 0) add(c - nx * ny, v3(0, 0, -1));
        if (k + 1 < nz) add(c + nx * ny, v3(0, 0, 1));
        bool have = false;
        uint64_t best = 0;
        double best_down = -std::numeric_limits<double>::infinity();
        uint32_t best_room = 0;
        for (int t = 0; t < nn; ++t) {
          uint64_t nc = nc_list[t];
          double down = down_list[t];
          if (!cavity_mask[nc] || count[nc] >= cap) continue;
          uint32_t room = cap - count[nc];
          bool better = !have || down > best_down || (down == best_down && room > best_room);
          if (better) {
            have = true;
            best = nc;
            best_down = down;
            best_room = room;
          }
        }
        if (have) {
          redist_moved++;
          count[c] -= 1;
          count[best] += 1;
          uint64_t ci = best % nx, cj = (best / nx) % ny, ck = best / (nx * ny);
          V3 jitter = hash_jitter(idx);
          V3 p = grid.cell_center(ci, cj, ck) + jitter * grid.dx;
          p = project_inside(p);
          pos[idx] = p;
          vel[idx] = grid.velocity_at(p);
        }
      }
    }
  }

  void p2g() {  // solver.rs:560-585
    std::fill(grid.u.begin(), grid.u.end(), 0.0);
    std::fill(grid.v.begin(), grid.v.end(), 0.0);
    std::fill(grid.w.begin(), grid.w.end(), 0.0);
    std::vector<double> wu(grid.u.size(), 0.0), wv(grid.v.size(), 0.0), ww(grid.w.size(), 0.0);
    for (uint64_t t = 0; t < pos.size(); ++t) {
      V3 l = grid.local(pos[t]);
      trilinear_scatter(grid.u.data(), wu.data(), grid.ud(), l.x, l.y - 0.5, l.z - 0.5, vel[t].x);
      l = grid.local(pos[t]);
      trilinear_scatter(grid.v.data(), wv.data(), grid.vd(), l.x - 0.5, l.y, l.z - 0.5, vel[t].y);
      l = grid.local(pos[t]);
      trilinear_scatter(grid.w.data(), ww.data(), grid.wd(), l.x - 0.5, l.y - 0.5, l.z, vel[t].z);
    }
    normalize(grid.u, wu);
    normalize(grid.v, wv);
    normalize(grid.w, ww);
  }
  static void normalize(std::vector<double>& f, const std::vector<double>& w) {  // solver.rs:944-952
    for (size_t t = 0; t < f.size(); ++t) {
      if (w[t] > 1e-12) f[t] /= w[t];
      else f[t] = 0.0;
    }
  }

  void add_forces(double dt) {  // solver.rs:593-604
    V3 g = fluid.gravity;
    if (g.x != 0.0) for (double& x : grid.u) x += g.x * dt;
    if (g.y != 0.0) for (double& x : grid.v) x += g.y * dt;
    if (g.z != 0.0) for (double& x : grid.w) x += g.z * dt;
  }

  void extrapolate_velocities() {  // solver.rs:608-647
    const uint64_t nx = grid.nx, ny = grid.ny, nz = grid.nz;
    auto fl = [&](uint64_t i, uint64_t j, uint64_t k) { return cells[(k * ny + j) * nx + i] == FLUID; };
    std::vector<uint8_t> vu(grid.u.size()), vv(grid.v.size()), vw(grid.w.size());
    for (uint64_t k = 0; k < nz; ++k)
      for (uint64_t j = 0; j < ny; ++j)
        for (uint64_t i = 0; i <= nx; ++i)
          vu[grid.u_idx(i, j, k)] = (i > 0 && fl(i - 1, j, k)) || (i < nx && fl(i, j, k));
    for (uint64_t k = 0; k < nz; ++k)
      for (uint64_t j = 0; j <= ny; ++j)
        for (uint64_t i = 0; i < nx; ++i)
          vv[grid.v_idx(i, j, k)] = (j > 0 && fl(i, j - 1, k)) || (j < ny && fl(i, j, k));
    for (uint64_t k = 0; k <= nz; ++k)
      for (uint64_t j = 0; j < ny; ++j)
        for (uint64_t i = 0; i < nx; ++i)
          vw[grid.w_idx(i, j, k)] = (k > 0 && fl(i, j, k - 1)) || (k < nz && fl(i, j, k));
    extrapolate(grid.u, vu, grid.ud(), 4);
    extrapolate(grid.v, vv, grid.vd(), 4);
    extrapolate(grid.w, vw, grid.wd(), 4);
  }

  void g2p() {  // solver.rs:650-677
    double flip = fluid.flip_ratio < 0.0 ? 0.0 : (fluid.flip_ratio > 1.0 ? 1.0 : fluid.flip_ratio);
    std::vector<double> du(grid.u.size()), dv(grid.v.size()), dw(grid.w.size());
    for (size_t t = 0; t < du.size(); ++t) du[t] = grid.u[t] - saved_u[t];
    for (size_t t = 0; t < dv.size(); ++t) dv[t] = grid.v[t] - saved_v[t];
    for (size_t t = 0; t < dw.size(); ++t) dw[t] = grid.w[t] - saved_w[t];
    for (uint64_t t = 0; t < pos.size(); ++t) {
      V3 p = pos[t];
      V3 pic = grid.velocity_at(p);
      V3 l = grid.local(p);
      double dvx = trilinear_sample(du.data(), grid.ud(), l.x, l.y - 0.5, l.z - 0.5);
      double dvy = trilinear_sample(dv.data(), grid.vd(), l.x - 0.5, l.y, l.z - 0.5);
      double dvz = trilinear_sample(dw.data(), grid.wd(), l.x - 0.5, l.y - 0.5, l.z);
      V3 flip_v = vel[t] + v3(dvx, dvy, dvz);
      vel[t] = flip_v * flip + pic * (1.0 - flip);
    }
  }

  void advect(double dt) {  // solver.rs:693-711
    double max_step = grid.dx;
    for (uint64_t t = 0; t < pos.size(); ++t) {
      V3 p0 = pos[t];
      V3 v0 = grid.velocity_at(p0);
      V3 mid = p0 + v0 * (0.5 * dt);
      V3 vmid = grid.velocity_at(mid);
      V3 st = vmid * dt;
      double len = length(st);
      if (len > max_step) st = st * (max_step / len);
      V3 p = project_inside(p0 + st);
      pos[t] = solid.sample(p) <= 0.0 ? p : p0;
    }
  }

  void enforce_particle_bounds() {  // solver.rs:736-755 + particles.rs:38-49
    Aabb bb = solid.bounds();
    uint64_t dr = 0, cu = 0, w = 0;
    for (uint64_t r = 0; r < pos.size(); ++r) {
      V3 p = pos[r];
      bool keep;
      if (!bb.contains(p)) {
        keep = false;
        cu++;
      } else if (has_outlet && outlet.contains(p)) {
        keep = false;
        dr++;
      } else {
        keep = solid.sample(p) < 0.5 * solid.dx;
        if (!keep) cu++;
      }
      if (keep) {
        pos[w] = pos[r];
        vel[w] = vel[r];
        w++;
      }
    }
    pos.resize(w);
    vel.resize(w);
    drained += dr;
    culled += cu;
  }

  // solver.rs:960-1009
  static void extrapolate(std::vector<double>& field, const std::vector<uint8_t>& valid, Dims d,
                          int iters) {
    const uint64_t nx = d.nx, ny = d.ny, nz = d.nz;
    auto idx = [&](uint64_t i, uint64_t j, uint64_t k) { return (k * ny + j) * nx + i; };
    std::vector<uint8_t> known = valid;
    for (int it = 0; it < iters; ++it) {
      std::vector<double> snap = field;
      std::vector<uint8_t> ksnap = known;
      for (uint64_t k = 0; k < nz; ++k)
        for (uint64_t j = 0; j < ny; ++j)
          for (uint64_t i = 0; i < nx; ++i) {
            uint64_t c = idx(i, j, k);
            if (ksnap[c]) continue;
            double sum = 0.0;
            int cnt = 0;
            auto acc = [&](uint64_t a, uint64_t b, uint64_t e) {
              uint64_t nc = idx(a, b, e);
              if (ksnap[nc]) {
                sum += snap[nc];
                cnt += 1;
              }
            };
            if (i > 0) acc(i - 1, j, k);
            if (i + 1 < nx) acc(i + 1, j, k);
            if (j > 0) acc(i, j - 1, k);
            if (j + 1 < ny) acc(i, j + 1, k);
            if (k > 0) acc(i, j, k - 1);
            if (k + 1 < nz) acc(i, j, k + 1);
            if (cnt > 0) {
              field[c] = sum / (double)cnt;
              known[c] = 1;
            }
          }
    }
  }
};

thread_local char g_err[512] = "";
int fail(int code, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

}  // namespace

// ================================================================ C API (orc_*)
// Same argument meaning as the sdr_* entry points of include/sdr_amd.h.
extern "C" {

const char* orc_last_error(void) { return g_err; }

int orc_set_threads(int32_t n) {
  if (n > 0) omp_set_num_threads(n);
  return 0;
}
int orc_get_threads(void) { return omp_get_max_threads(); }

int orc_solver_new(const sdr_sdf* sdf, double sim_dx, const sdr_fluid_params* f,
                   const sdr_needle* nd, const sdr_options* /*opts*/, void** out) {
  if (!sdf || !f || !nd || !out || !sdf->data || sdf->nx < 2 || sdf->ny < 2 || sdf->nz < 2)
    return fail(SDR_E_INVALID, "invalid argument");
  Sdf s;
  s.origin = v3(sdf->origin.x, sdf->origin.y, sdf->origin.z);
  s.dx = sdf->dx;
  s.nx = sdf->nx, s.ny = sdf->ny, s.nz = sdf->nz;
  s.data.assign(sdf->data, sdf->data + sdf->nx * sdf->ny * sdf->nz);
  Fluid fl{f->density, v3(f->gravity.x, f->gravity.y, f->gravity.z), f->flip_ratio,
           f->particles_per_cell, f->volume_stiffness, f->redistribute != 0, f->redist_cap_factor};
  Needle n{v3(nd->tip.x, nd->tip.y, nd->tip.z), v3(nd->axis.x, nd->axis.y, nd->axis.z), nd->radius,
           nd->flow_rate, nd->inject_until};
  Solver* so = new Solver();
  so->create(std::move(s), sim_dx, fl, n);
  *out = so;
  return 0;
}
int orc_solver_free(void* h) {
  delete (Solver*)h;
  return 0;
}
int orc_solver_set_outlet(void* h, const sdr_aabb* o) {
  Solver* s = (Solver*)h;
  s->has_outlet = o != nullptr;
  if (o) s->outlet = Aabb{v3(o->min.x, o->min.y, o->min.z), v3(o->max.x, o->max.y, o->max.z)};
  return 0;
}
int orc_solver_set_solve_params(void* h, const sdr_solve_params* p) {
  Solver* s = (Solver*)h;
  s->max_iters = p->max_iters;
  s->tol = p->tol;
  return 0;
}
int orc_solver_set_needle(void* h, const sdr_needle* nd) {
  Solver* s = (Solver*)h;
  s->needle = Needle{v3(nd->tip.x, nd->tip.y, nd->tip.z), v3(nd->axis.x, nd->axis.y, nd->axis.z),
                     nd->radius, nd->flow_rate, nd->inject_until};
  return 0;
}
int orc_solver_step(void* h, double dt) {
  ((Solver*)h)->step(dt);
  return 0;
}
int orc_solver_substep(void* h, double dt) {
  ((Solver*)h)->substep(dt);
  return 0;
}
int orc_solver_run_stage(void* h, int32_t stage, double dt) {
  if (stage < 0 || stage >= SDR_STAGE_COUNT) return fail(SDR_E_INVALID, "bad stage");
  ((Solver*)h)->run_stage(stage, dt);
  return 0;
}
int orc_solver_redistribute(void* h) {
  ((Solver*)h)->redistribute();
  return 0;
}
int orc_solver_particle_count(const void* h, uint64_t* n) {
  *n = ((const Solver*)h)->pos.size();
  return 0;
}
int orc_solver_get_particles(void* h, double* pos, double* vel, uint64_t n) {
  Solver* s = (Solver*)h;
  if (n != s->pos.size()) return fail(SDR_E_INVALID, "particle count mismatch");
  for (uint64_t t = 0; t < n; ++t) {
    if (pos) pos[3 * t] = s->pos[t].x, pos[3 * t + 1] = s->pos[t].y, pos[3 * t + 2] = s->pos[t].z;
    if (vel) vel[3 * t] = s->vel[t].x, vel[3 * t + 1] = s->vel[t].y, vel[3 * t + 2] = s->vel[t].z;
  }
  return 0;
}
int orc_solver_set_particles(void* h, const double* pos, const double* vel, uint64_t n) {
  Solver* s = (Solver*)h;
  // the replaced particles leave the books, the new ones enter them, so that
  // emitted = resident + drained + culled keeps holding
  s->emitted = s->emitted - s->pos.size() + n;
  s->pos.resize(n);
  s->vel.resize(n);
  for (uint64_t t = 0; t < n; ++t) {
    s->pos[t] = v3(pos[3 * t], pos[3 * t + 1], pos[3 * t + 2]);
    s->vel[t] = vel ? v3(vel[3 * t], vel[3 * t + 1], vel[3 * t + 2]) : v3(0, 0, 0);
  }
  return 0;
}
int orc_solver_get_dims(const void* h, sdr_dims* d) {
  const Solver* s = (const Solver*)h;
  d->nx = s->grid.nx, d->ny = s->grid.ny, d->nz = s->grid.nz, d->dx = s->grid.dx;
  d->origin = sdr_vec3{s->grid.origin.x, s->grid.origin.y, s->grid.origin.z};
  return 0;
}
static void cp_out(const std::vector<double>& a, double* o) {
  if (o) memcpy(o, a.data(), a.size() * sizeof(double));
}
static void cp_in(std::vector<double>& a, const double* i, size_t n) {
  if (i) a.assign(i, i + n);
}
int orc_solver_get_grid(void* h, double* u, double* v, double* w) {
  Solver* s = (Solver*)h;
  cp_out(s->grid.u, u), cp_out(s->grid.v, v), cp_out(s->grid.w, w);
  return 0;
}
int orc_solver_set_grid(void* h, const double* u, const double* v, const double* w) {
  Solver* s = (Solver*)h;
  cp_in(s->grid.u, u, s->grid.u.size()), cp_in(s->grid.v, v, s->grid.v.size()),
      cp_in(s->grid.w, w, s->grid.w.size());
  return 0;
}
int orc_solver_get_saved(void* h, double* u, double* v, double* w) {
  Solver* s = (Solver*)h;
  if (s->saved_u.empty()) return fail(SDR_E_STATE, "no saved grid yet");
  cp_out(s->saved_u, u), cp_out(s->saved_v, v), cp_out(s->saved_w, w);
  return 0;
}
int orc_solver_set_saved(void* h, const double* u, const double* v, const double* w) {
  Solver* s = (Solver*)h;
  cp_in(s->saved_u, u, s->grid.u.size()), cp_in(s->saved_v, v, s->grid.v.size()),
      cp_in(s->saved_w, w, s->grid.w.size());
  return 0;
}
int orc_solver_get_cells(void* h, uint8_t* c) {
  Solver* s = (Solver*)h;
  memcpy(c, s->cells.data(), s->cells.size());
  return 0;
}
int orc_solver_set_cells(void* h, const uint8_t* c) {
  Solver* s = (Solver*)h;
  s->cells.assign(c, c + s->cells.size());
  return 0;
}
int orc_solver_get_pressure(void* h, double* p) {
  cp_out(((Solver*)h)->pressure, p);
  return 0;
}
int orc_solver_get_mask(void* h, uint8_t* m) {
  Solver* s = (Solver*)h;
  memcpy(m, s->cavity_mask.data(), s->cavity_mask.size());
  return 0;
}
int orc_solver_get_state(const void* h, sdr_state* st) {
  const Solver* s = (const Solver*)h;
  memset(st, 0, sizeof *st);
  st->time = s->time, st->steps = s->steps, st->last_pcg_iters = s->last_pcg_iters;
  st->last_substeps = s->last_substeps, st->drained = s->drained, st->rng = s->rng.s;
  st->emit_accum = s->emit_accum, st->emitted = s->emitted, st->culled = s->culled;
  st->last_residual = s->last_residual;
  st->redist_active = s->redist_active, st->redist_moved = s->redist_moved;
  return 0;
}
int orc_solver_set_state(void* h, const sdr_state* st) {
  Solver* s = (Solver*)h;
  s->time = st->time, s->steps = st->steps, s->last_pcg_iters = st->last_pcg_iters;
  s->last_substeps = st->last_substeps, s->drained = st->drained, s->rng.s = st->rng;
  s->emit_accum = st->emit_accum, s->emitted = st->emitted, s->culled = st->culled;
  s->last_residual = st->last_residual;
  return 0;
}

int orc_project_capturing_with_source(const sdr_dims* d, double* u, double* v, double* w,
                                      const uint8_t* cells, double rho, double dt,
                                      const sdr_solve_params* prm, const double* source,
                                      double* pout, sdr_solve_report* rep,
                                      const sdr_options* /*opts*/) {
  MacGrid g;
  g.init(d->nx, d->ny, d->nz, d->dx, v3(d->origin.x, d->origin.y, d->origin.z));
  g.u.assign(u, u + g.u.size()), g.v.assign(v, v + g.v.size()), g.w.assign(w, w + g.w.size());
  std::vector<double> p;
  sdr_solve_params sp = prm ? *prm : sdr_solve_params{200, 1e-6};
  sdr_solve_report r =
      project_capturing_with_source(g, cells, rho, dt, sp.max_iters, sp.tol, source, p);
  memcpy(u, g.u.data(), g.u.size() * 8), memcpy(v, g.v.data(), g.v.size() * 8),
      memcpy(w, g.w.data(), g.w.size() * 8);
  if (pout) memcpy(pout, p.data(), p.size() * 8);
  if (rep) *rep = r;
  return 0;
}

// ---- pieces exposed for pinning the oracle against closed forms / brute force
int orc_trilinear_sample(const double* arr, uint64_t nx, uint64_t ny, uint64_t nz, double gx,
                         double gy, double gz, double* out) {
  *out = trilinear_sample(arr, Dims{nx, ny, nz}, gx, gy, gz);
  return 0;
}
int orc_trilinear_scatter(double* arr, double* wsum, uint64_t nx, uint64_t ny, uint64_t nz,
                          double gx, double gy, double gz, double val) {
  trilinear_scatter(arr, wsum, Dims{nx, ny, nz}, gx, gy, gz, val);
  return 0;
}
int orc_sdf_sample(const sdr_sdf* sd, const double* p, double* phi, double* grad) {
  Sdf s;
  s.origin = v3(sd->origin.x, sd->origin.y, sd->origin.z), s.dx = sd->dx;
  s.nx = sd->nx, s.ny = sd->ny, s.nz = sd->nz;
  s.data.assign(sd->data, sd->data + sd->nx * sd->ny * sd->nz);
  V3 q = v3(p[0], p[1], p[2]);
  if (phi) *phi = s.sample(q);
  if (grad) {
    V3 g = s.gradient(q);
    grad[0] = g.x, grad[1] = g.y, grad[2] = g.z;
  }
  return 0;
}
int orc_hash_jitter(uint64_t key, double* out) {
  V3 j = hash_jitter(key);
  out[0] = j.x, out[1] = j.y, out[2] = j.z;
  return 0;
}
int orc_rng_units(uint64_t seed, uint64_t n, double* out, uint64_t* state_out) {
  Rng r{seed | 1};
  for (uint64_t t = 0; t < n; ++t) out[t] = r.unit();
  if (state_out) *state_out = r.s;
  return 0;
}
// Laplacian pieces on an explicit cell classification (tests compare with numpy/scipy).
int orc_laplacian(uint64_t nx, uint64_t ny, uint64_t nz, const uint8_t* cells, double* diag,
                  double* px, double* py, double* pz) {
  MacGrid g;
  g.init(nx, ny, nz, 1.0, v3(0, 0, 0));
  Laplacian a = assemble(g, cells);
  size_t n = nx * ny * nz * sizeof(double);
  memcpy(diag, a.diag.data(), n), memcpy(px, a.plus_x.data(), n), memcpy(py, a.plus_y.data(), n),
      memcpy(pz, a.plus_z.data(), n);
  return 0;
}
// tau < 0 selects the reference constant 0.97; other values exist only to check the
// IC(0) (tau = 0) and modified (tau = 1) identities of the factor in tests.
int orc_mic0(uint64_t nx, uint64_t ny, uint64_t nz, const uint8_t* cells, const double* r,
             double tau, double* precon_out, double* z_out) {
  MacGrid g;
  g.init(nx, ny, nz, 1.0, v3(0, 0, 0));
  Laplacian a = assemble(g, cells);
  std::vector<double> pc = build_mic0(a, cells, tau < 0 ? 0.97 : tau);
  if (precon_out) memcpy(precon_out, pc.data(), pc.size() * 8);
  if (r && z_out) apply_mic0(a, pc.data(), cells, r, z_out);
  return 0;
}
int orc_apply_a(uint64_t nx, uint64_t ny, uint64_t nz, const uint8_t* cells, const double* x,
                double* out) {
  MacGrid g;
  g.init(nx, ny, nz, 1.0, v3(0, 0, 0));
  Laplacian a = assemble(g, cells);
  apply_a(a, cells, x, out);
  return 0;
}
int orc_dot(const double* a, const double* b, uint64_t n, double* out) {
  *out = dot_chunked(a, b, n);
  return 0;
}
int orc_extrapolate(double* field, const uint8_t* valid, uint64_t nx, uint64_t ny, uint64_t nz,
                    int32_t iters) {
  std::vector<double> f(field, field + nx * ny * nz);
  std::vector<uint8_t> v(valid, valid + nx * ny * nz);
  Solver::extrapolate(f, v, Dims{nx, ny, nz}, iters);
  memcpy(field, f.data(), f.size() * 8);
  return 0;
}

// ---- Solver diagnostics (solver.rs:225-233, 758-891; particles.rs:52-58)
int orc_solver_cfl_dt(void* h, double* out) {  // solver.rs:262-281
  *out = ((Solver*)h)->cfl_dt();
  return 0;
}
int orc_solver_diagnostics(void* h, sdr_diagnostics* d) {
  Solver* s = (Solver*)h;
  const uint64_t nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz, n = nx * ny * nz;
  memset(d, 0, sizeof *d);
  d->particle_volume = s->particle_volume();                              // :225-228
  d->fluid_volume = (double)s->pos.size() * s->particle_volume();         // :231-233
  d->drained_volume = (double)s->drained * s->particle_volume();          // :758-760
  for (uint64_t c = 0; c < n; ++c) d->cells_fluid_count += s->cells[c] == FLUID;  // :763-765
  std::vector<uint32_t> count(n, 0);                                       // :770-782
  uint32_t mx = 0;
  for (const V3& p : s->pos) {
    uint64_t i, j, k;
    if (s->cell_of(p, i, j, k)) {
      uint64_t c = (k * ny + j) * nx + i;
      count[c] += 1;
      mx = count[c] > mx ? count[c] : mx;
    }
  }
  d->max_particles_per_cell = mx;
  double pmax = 0.0, psum = 0.0;                                           // :859-882
  uint64_t pn = 0;
  for (uint64_t c = 0; c < n; ++c) {
    if (s->cells[c] != FLUID) continue;
    pmax = rmax(pmax, s->pressure[c]);
    psum = psum + s->pressure[c];
    pn += 1;
  }
  d->max_pressure = pmax;
  d->mean_pressure = pn > 0 ? psum / (double)pn : 0.0;
  double vmax = 0.0, ke = 0.0;                                             // :884-890
  for (const V3& v : s->vel) {
    vmax = rmax(vmax, length(v));
    ke = ke + dot(v, v);
  }
  d->max_speed = vmax;
  d->kinetic_energy = 0.5 * ke;                                            // particles.rs:52-58
  d->particles = s->pos.size();
  return 0;
}

// ---- MetricsCollector (metrics.rs:116-446)
namespace {
constexpr double kM3ToMl = 1.0e6;              // metrics.rs:23
constexpr double kPaToMmHg = 1.0 / 133.322;    // metrics.rs:24
constexpr double kMembraneCapFactor = 3.0;     // metrics.rs:454

// metrics.rs:460-479
void membrane_stats(const double* v, uint64_t n, double cap, double* mean, double* peak, uint64_t* art) {
  double sum = 0.0, pk = 0.0;
  uint64_t m = 0, a = 0;
  for (uint64_t t = 0; t < n; ++t) {
    double p = v[t];
    if (p > cap) {
      a += 1;
      continue;
    }
    sum += p;
    pk = rmax(pk, p);
    m += 1;
  }
  if (m == 0) *mean = 0.0, *peak = 0.0;
  else *mean = sum / (double)m, *peak = pk;
  *art = a;
}

struct Metrics {
  double cavity_volume = 0;
  uint64_t nx = 0, ny = 0, nz = 0;
  V3 origin{0, 0, 0};
  double dx = 0;
  std::vector<uint8_t> is_wall, washed;
  uint64_t wall_count = 0;
  std::vector<sdr_frame_metrics> history;
  double injected_ml = 0;

  void create(const Solver& s, double cv) {  // metrics.rs:119-182
    cavity_volume = cv;
    nx = s.grid.nx, ny = s.grid.ny, nz = s.grid.nz, origin = s.grid.origin, dx = s.grid.dx;
    const uint64_t n = nx * ny * nz;
    std::vector<uint8_t> inside(n);
    for (uint64_t k = 0; k < nz; ++k)
      for (uint64_t j = 0; j < ny; ++j)
        for (uint64_t i = 0; i < nx; ++i)
          inside[(k * ny + j) * nx + i] = s.solid.sample(s.grid.cell_center(i, j, k)) < 0.0;
    is_wall.assign(n, 0);
    for (uint64_t k = 0; k < nz; ++k)
      for (uint64_t j = 0; j < ny; ++j)
        for (uint64_t i = 0; i < nx; ++i) {
          uint64_t c = (k * ny + j) * nx + i;
          if (!inside[c]) continue;
          auto nb_solid = [&](int64_t ci, int64_t cj, int64_t ck) {
            if (ci < 0 || cj < 0 || ck < 0 || (uint64_t)ci >= nx || (uint64_t)cj >= ny || (uint64_t)ck >= nz)
              return true;
            return !inside[((uint64_t)ck * ny + (uint64_t)cj) * nx + (uint64_t)ci];
          };
          int64_t a = (int64_t)i, b = (int64_t)j, e = (int64_t)k;
          bool touches = nb_solid(a - 1, b, e) || nb_solid(a + 1, b, e) || nb_solid(a, b - 1, e) ||
                         nb_solid(a, b + 1, e) || nb_solid(a, b, e - 1) || nb_solid(a, b, e + 1);
          if (touches) is_wall[c] = 1, wall_count += 1;
        }
    washed.assign(n, 0);
  }
  bool cell_index(V3 p, uint64_t& c) const {  // metrics.rs:434-445
    V3 l = (p - origin) / dx;
    if (l.x < 0.0 || l.y < 0.0 || l.z < 0.0) return false;
    uint64_t i = as_usize(l.x), j = as_usize(l.y), k = as_usize(l.z);
    if (i < nx && j < ny && k < nz) {
      c = (k * ny + j) * nx + i;
      return true;
    }
    return false;
  }
  sdr_frame_metrics record(Solver& s, uint64_t frame) {  // metrics.rs:188-316
    // the wet wall set in ascending cell order (the reference's BTreeSet)
    std::vector<uint8_t> wet(nx * ny * nz, 0);
    for (const V3& p : s.pos) {
      uint64_t c;
      if (cell_index(p, c) && is_wall[c]) washed[c] = 1, wet[c] = 1;
    }
    uint64_t washed_count = 0;
    for (uint8_t w : washed) washed_count += w;
    double coverage = wall_count > 0 ? (double)washed_count / (double)wall_count : 0.0;
    double fluid_volume = (double)s.pos.size() * s.particle_volume();
    double drained_volume = (double)s.drained * s.particle_volume();
    double fill = cavity_volume > 0.0 ? fluid_volume / cavity_volume : 0.0;
    sdr_diagnostics dg;
    orc_solver_diagnostics(&s, &dg);
    double max_p = dg.max_pressure;
    injected_ml = (fluid_volume + drained_volume) * kM3ToMl;
    double v_jet = s.needle.jet_speed();
    double stagnation = 0.5 * s.fluid.density * v_jet * v_jet;
    double hydrostatic = s.fluid.density * length(s.fluid.gravity) * ((double)ny * dx);
    double p_cap = kMembraneCapFactor * stagnation + hydrostatic + 500.0;
    std::vector<double> wall_p;
    for (uint64_t c = 0; c < wet.size(); ++c)
      if (wet[c]) wall_p.push_back(s.pressure[c]);
    double wall_mean, wall_peak;
    uint64_t art;
    membrane_stats(wall_p.data(), wall_p.size(), p_cap, &wall_mean, &wall_peak, &art);
    sdr_frame_metrics m;
    memset(&m, 0, sizeof m);
    m.frame = frame;
    m.time_s = s.time;
    m.particles = s.pos.size();
    m.fluid_volume_ml = fluid_volume * kM3ToMl;
    m.drained_volume_ml = drained_volume * kM3ToMl;
    m.fill_fraction = fill;
    m.wall_coverage = coverage;
    m.mean_pressure_pa = dg.mean_pressure;
    m.max_pressure_pa = max_p;
    m.max_pressure_mmhg = max_p * kPaToMmHg;
    m.wall_pressure_pa = wall_mean;
    m.peak_wall_pressure_pa = wall_peak;
    m.peak_wall_pressure_mmhg = wall_peak * kPaToMmHg;
    m.membrane_artifacts = art;
    m.max_speed_m_s = dg.max_speed;
    m.kinetic_energy = dg.kinetic_energy;
    m.pcg_iters = s.last_pcg_iters;
    m.wet_wall_cells = wall_p.size();
    m.emitted = s.emitted, m.drained = s.drained, m.culled = s.culled;
    history.push_back(m);
    return m;
  }
  sdr_summary summary() const {  // metrics.rs:324-382
    sdr_summary r;
    memset(&r, 0, sizeof r);
    double peak_fill = 0, peak_cov = 0, peak_p = 0, peak_mem = 0, mean_mem_sum = 0;
    uint64_t art = 0;
    for (const auto& m : history) {
      peak_fill = rmax(peak_fill, m.fill_fraction);
      peak_cov = rmax(peak_cov, m.wall_coverage);
      peak_p = rmax(peak_p, m.max_pressure_pa);
      peak_mem = rmax(peak_mem, m.peak_wall_pressure_pa);
      mean_mem_sum = mean_mem_sum + m.wall_pressure_pa;
      art += m.membrane_artifacts;
    }
    r.frames = history.size();
    double mean_mem = history.empty() ? 0.0 : mean_mem_sum / (double)history.size();
    if (!history.empty()) {
      const auto& l = history.back();
      r.final_fill_fraction = l.fill_fraction, r.final_wall_coverage = l.wall_coverage;
      r.final_time_s = l.time_s;
      r.total_drained_ml = l.drained_volume_ml;
      r.injected_ml = l.fluid_volume_ml + l.drained_volume_ml;
      r.emitted = l.emitted, r.resident = l.particles, r.drained = l.drained, r.culled = l.culled;
    }
    r.peak_fill_fraction = peak_fill, r.peak_wall_coverage = peak_cov;
    r.peak_membrane_pressure_pa = peak_mem, r.peak_membrane_pressure_mmhg = peak_mem * kPaToMmHg;
    r.mean_membrane_pressure_pa = mean_mem, r.mean_membrane_pressure_mmhg = mean_mem * kPaToMmHg;
    r.peak_pressure_pa = peak_p, r.peak_pressure_mmhg = peak_p * kPaToMmHg;
    r.membrane_artifacts_total = art;
    // conservation from the integer counts: emitted = resident + drained + culled
    r.mass_balance = r.emitted > 0 ? (double)(r.resident + r.drained + r.culled) / (double)r.emitted : 1.0;
    return r;
  }
};
}  // namespace

int orc_membrane_stats(const double* v, uint64_t n, double cap, double* mean, double* peak, uint64_t* art) {
  membrane_stats(v, n, cap, mean, peak, art);
  return 0;
}
int orc_metrics_new(void* solver, double cavity_volume, void** out) {
  Metrics* m = new Metrics();
  m->create(*(Solver*)solver, cavity_volume);
  *out = m;
  return 0;
}
int orc_metrics_free(void* m) {
  delete (Metrics*)m;
  return 0;
}
int orc_metrics_wall_count(const void* m, uint64_t* n) {
  *n = ((const Metrics*)m)->wall_count;
  return 0;
}
int orc_metrics_record(void* m, void* solver, uint64_t frame, sdr_frame_metrics* out) {
  sdr_frame_metrics r = ((Metrics*)m)->record(*(Solver*)solver, frame);
  if (out) *out = r;
  return 0;
}
int orc_metrics_summary(const void* m, sdr_summary* out) {
  *out = ((const Metrics*)m)->summary();
  return 0;
}
int orc_metrics_get_cells(const void* m, uint8_t* washed, uint8_t* is_wall) {
  const Metrics* mm = (const Metrics*)m;
  if (washed) memcpy(washed, mm->washed.data(), mm->washed.size());
  if (is_wall) memcpy(is_wall, mm->is_wall.data(), mm->is_wall.size());
  return 0;
}

// ---- Sdf::from_mesh / from_mesh_in_bounds (sdf.rs:47-86), signed_distance_to_mesh,
// point_triangle_distance_sq, solid_angle (sdf.rs:187-272)
namespace {
V3 vmin(V3 a, V3 b) { return v3(rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)); }  // math.rs min
V3 vmax(V3 a, V3 b) { return v3(rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)); }
double length_squared(V3 a) { return dot(a, a); }
double point_triangle_distance_sq(V3 p, V3 a, V3 b, V3 c) {  // sdf.rs:201-249
  V3 ab = b - a, ac = c - a, ap = p - a;
  double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return length_squared(ap);
  V3 bp = p - b;
  double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return length_squared(bp);
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    double v = d1 / (d1 - d3);
    return length_squared(p - (a + ab * v));
  }
  V3 cp = p - c;
  double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return length_squared(cp);
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    double w = d2 / (d2 - d6);
    return length_squared(p - (a + ac * w));
  }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return length_squared(p - (b + (c - b) * w));
  }
  double denom = 1.0 / (va + vb + vc);
  double v = vb * denom, w = vc * denom;
  return length_squared(p - (a + ab * v + ac * w));
}
double solid_angle(V3 p, V3 a, V3 b, V3 c) {  // sdf.rs:254-272
  V3 av = a - p, bv = b - p, cv = c - p;
  double la = length(av), lb = length(bv), lc = length(cv);
  if (la < 1e-12 || lb < 1e-12 || lc < 1e-12) return 0.0;
  double numerator = dot(av, cross(bv, cv));
  double denominator = la * lb * lc + dot(av, bv) * lc + dot(bv, cv) * la + dot(cv, av) * lb;
  return 2.0 * std::atan2(numerator, denominator);
}
}  // namespace

int orc_sdf_from_mesh(const double* verts, uint64_t nv, const uint32_t* tris, uint64_t nt, double dx,
                      const sdr_aabb* bounds, double padding, int32_t /*device*/, sdr_sdf* out, double* data,
                      uint64_t cap) {
  if (!verts || !tris || !out || !(dx > 0.0)) return fail(SDR_E_INVALID, "invalid argument");
  for (uint64_t t = 0; t < 3 * nt; ++t)
    if (tris[t] >= nv) return fail(SDR_E_INVALID, "triangle index out of range");
  Aabb bb;
  if (bounds) {
    bb = Aabb{v3(bounds->min.x, bounds->min.y, bounds->min.z), v3(bounds->max.x, bounds->max.y, bounds->max.z)};
  } else {  // mesh.rs:47-49 (Aabb::from_points), then padded (math.rs:241-246)
    const double inf = std::numeric_limits<double>::infinity();
    bb = Aabb{v3(inf, inf, inf), v3(-inf, -inf, -inf)};
    for (uint64_t q = 0; q < nv; ++q) {
      V3 p = v3(verts[3 * q], verts[3 * q + 1], verts[3 * q + 2]);
      bb.min = vmin(bb.min, p), bb.max = vmax(bb.max, p);
    }
    bb.min = bb.min - v3(padding, padding, padding);
    bb.max = bb.max + v3(padding, padding, padding);
  }
  V3 size = bb.max - bb.min;
  uint64_t nx = as_usize(std::ceil(size.x / dx)) + 1, ny = as_usize(std::ceil(size.y / dx)) + 1,
           nz = as_usize(std::ceil(size.z / dx)) + 1;
  out->origin = sdr_vec3{bb.min.x, bb.min.y, bb.min.z};
  out->dx = dx;
  out->nx = nx, out->ny = ny, out->nz = nz;
  out->data = data;
  if (!data) return 0;
  if (cap < nx * ny * nz) return fail(SDR_E_INVALID, "output holds fewer values than the node grid");
  std::vector<V3> tv(3 * nt);
  for (uint64_t t = 0; t < 3 * nt; ++t) {
    const double* q = verts + 3 * (uint64_t)tris[t];
    tv[t] = v3(q[0], q[1], q[2]);
  }
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)nz; ++k)
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        V3 p = bb.min + v3((double)i * dx, (double)j * dx, (double)k * dx);
        double min_d2 = std::numeric_limits<double>::infinity(), winding = 0.0;
        for (uint64_t t = 0; t < nt; ++t) {  // signed_distance_to_mesh (sdf.rs:187-198)
          double d2 = point_triangle_distance_sq(p, tv[3 * t], tv[3 * t + 1], tv[3 * t + 2]);
          if (d2 < min_d2) min_d2 = d2;
          winding += solid_angle(p, tv[3 * t], tv[3 * t + 1], tv[3 * t + 2]);
        }
        winding /= 4.0 * M_PI;
        double dist = std::sqrt(min_d2);
        data[((uint64_t)k * ny + j) * nx + i] = winding > 0.5 ? -dist : dist;
      }
  return 0;
}


// ---- Particle surfacing: naive surface nets (surface_nets.rs:48-221) over sampled node
// values, and the particle distance field the device surfaces (include/sdr_amd.h:
// phi(x) = min(min_j |x - x_j|, 2 r) - r).  Sequential, in the reference's order.
// One rule is the build's, on the grid's max faces: a ring cell outside the grid counts
// as missing (the reference indexes past the row or the array there, surface_nets.rs:
// 150-192); a field that straddles there is an open surface either way.
namespace {
const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};  // surface_nets.rs:16-25
const int kEdge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                          {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};  // surface_nets.rs:28-42

void net_emit_quad(const int64_t ring[4], V3 desired, const std::vector<V3>& verts,
                   std::vector<uint32_t>& idx) {  // surface_nets.rs:197-221
  for (int q = 0; q < 4; ++q)
    if (ring[q] < 0) return;
  const uint32_t a = (uint32_t)ring[0], b = (uint32_t)ring[1], c = (uint32_t)ring[2], d = (uint32_t)ring[3];
  const V3 pa = verts[a], pb = verts[b], pc = verts[c];
  const V3 n = cross(pb - pa, pc - pa);
  if (dot(n, desired) >= 0.0) {
    idx.insert(idx.end(), {a, b, c, a, c, d});
  } else {
    idx.insert(idx.end(), {a, c, b, a, d, c});
  }
}

void surface_nets(const double* values, uint64_t nx, uint64_t ny, uint64_t nz, V3 origin, double dx,
                  std::vector<V3>& vertices, std::vector<uint32_t>& indices) {  // surface_nets.rs:48-195
  auto node = [&](uint64_t i, uint64_t j, uint64_t k) { return values[(k * ny + j) * nx + i]; };
  const uint64_t cx = nx - 1, cy = ny - 1, cz = nz - 1;
  std::vector<int64_t> cell_vertex(cx * cy * cz, -1);
  auto cell_idx = [&](uint64_t i, uint64_t j, uint64_t k) { return (k * cy + j) * cx + i; };
  vertices.clear(), indices.clear();
  for (uint64_t k = 0; k < cz; ++k)
    for (uint64_t j = 0; j < cy; ++j)
      for (uint64_t i = 0; i < cx; ++i) {
        double vals[8];
        for (int c = 0; c < 8; ++c) vals[c] = node(i + kCorner[c][0], j + kCorner[c][1], k + kCorner[c][2]);
        std::vector<V3> crossings;
        for (const auto& e : kEdge) {
          const double va = vals[e[0]], vb = vals[e[1]];
          if ((va < 0.0) != (vb < 0.0)) {
            const double t = va / (va - vb);
            const V3 pa = v3(kCorner[e[0]][0], kCorner[e[0]][1], kCorner[e[0]][2]);
            const V3 pb = v3(kCorner[e[1]][0], kCorner[e[1]][1], kCorner[e[1]][2]);
            crossings.push_back(pa + (pb - pa) * t);
          }
        }
        if (crossings.empty()) continue;
        V3 sum = v3(0.0, 0.0, 0.0);
        for (const V3& c : crossings) sum = sum + c;
        const V3 local = sum / (double)crossings.size();
        const V3 world = origin + v3(((double)i + local.x) * dx, ((double)j + local.y) * dx, ((double)k + local.z) * dx);
        cell_vertex[cell_idx(i, j, k)] = (int64_t)vertices.size();
        vertices.push_back(world);
      }
  auto get = [&](uint64_t i, uint64_t j, uint64_t k) -> int64_t {
    if (i >= cx || j >= cy || k >= cz) return -1;
    return cell_vertex[cell_idx(i, j, k)];
  };
  for (uint64_t k = 0; k < nz; ++k)
    for (uint64_t j = 0; j < ny; ++j)
      for (uint64_t i = 0; i < nx; ++i) {
        const double v0 = node(i, j, k);
        if (i + 1 < nx && j >= 1 && k >= 1) {
          const double v1 = node(i + 1, j, k);
          if ((v0 < 0.0) != (v1 < 0.0)) {
            const int64_t ring[4] = {get(i, j - 1, k - 1), get(i, j, k - 1), get(i, j, k), get(i, j - 1, k)};
            net_emit_quad(ring, v0 < v1 ? v3(1.0, 0.0, 0.0) : v3(-1.0, 0.0, 0.0), vertices, indices);
          }
        }
        if (j + 1 < ny && i >= 1 && k >= 1) {
          const double v1 = node(i, j + 1, k);
          if ((v0 < 0.0) != (v1 < 0.0)) {
            const int64_t ring[4] = {get(i - 1, j, k - 1), get(i, j, k - 1), get(i, j, k), get(i - 1, j, k)};
            net_emit_quad(ring, v0 < v1 ? v3(0.0, 1.0, 0.0) : v3(0.0, -1.0, 0.0), vertices, indices);
          }
        }
        if (k + 1 < nz && i >= 1 && j >= 1) {
          const double v1 = node(i, j, k + 1);
          if ((v0 < 0.0) != (v1 < 0.0)) {
            const int64_t ring[4] = {get(i - 1, j - 1, k), get(i, j - 1, k), get(i, j, k), get(i - 1, j, k)};
            net_emit_quad(ring, v0 < v1 ? v3(0.0, 0.0, 1.0) : v3(0.0, 0.0, -1.0), vertices, indices);
          }
        }
      }
}

void mesh_measure(const std::vector<V3>& v, const std::vector<uint32_t>& t, double* area, double* volume) {
  double a = 0.0, vol = 0.0;
  for (size_t f = 0; 3 * f < t.size(); ++f) {
    const V3 p = v[t[3 * f]], q = v[t[3 * f + 1]], r = v[t[3 * f + 2]];
    a += 0.5 * length(cross(q - p, r - p));  // mesh.rs:52-60
    vol += dot(p, cross(q, r)) / 6.0;          // mesh.rs:63-70
  }
  *area = a, *volume = vol;
}

// floor(v) clamped to [0, n - 1] (NaN -> 0): the cells a node's reach overlaps.
int64_t bin_clamp(double v, uint64_t n) {
  const double f = std::floor(v);
  if (!(f >= 0.0)) return 0;
  if (f > (double)(n - 1)) return (int64_t)n - 1;
  return (int64_t)f;
}
}  // namespace

int orc_surface_nets(const double* values, uint64_t nx, uint64_t ny, uint64_t nz, sdr_vec3 origin, double dx,
                     int32_t /*device*/, sdr_surface_info* info, double* vertices, uint64_t vcap, uint32_t* tris,
                     uint64_t tcap) {
  if (!values || !info) return fail(SDR_E_INVALID, "null argument");
  if (nx < 2 || ny < 2 || nz < 2) return fail(SDR_E_INVALID, "surface_nets needs at least 2 nodes per axis");
  std::vector<V3> v;
  std::vector<uint32_t> t;
  surface_nets(values, nx, ny, nz, v3(origin.x, origin.y, origin.z), dx, v, t);
  memset(info, 0, sizeof *info);
  info->nx = nx, info->ny = ny, info->nz = nz, info->origin = origin, info->spacing = dx;
  info->vertices = v.size(), info->triangles = t.size() / 3;
  mesh_measure(v, t, &info->area, &info->volume);
  if ((vertices && vcap < v.size()) || (tris && tcap < t.size() / 3))
    return fail(SDR_E_INVALID, "output holds fewer vertices or triangles than the surface");
  if (vertices)
    for (size_t q = 0; q < v.size(); ++q) vertices[3 * q] = v[q].x, vertices[3 * q + 1] = v[q].y, vertices[3 * q + 2] = v[q].z;
  if (tris && !t.empty()) memcpy(tris, t.data(), t.size() * 4);
  return 0;
}

int orc_solver_surface(void* h, const sdr_surface_params* prm, sdr_surface_info* info) {
  Solver* s = (Solver*)h;
  if (!info) return fail(SDR_E_INVALID, "null output");
  const MacGrid& g = s->grid;
  const uint64_t ppc = s->fluid.ppc > 1 ? s->fluid.ppc : 1;
  const double r = prm && prm->radius > 0.0 ? prm->radius : g.dx / std::cbrt((double)ppc);
  const double hh = prm && prm->spacing > 0.0 ? prm->spacing : g.dx;
  const uint64_t ncell[3] = {g.nx, g.ny, g.nz};
  uint64_t nn[3];
  for (int a = 0; a < 3; ++a)
    nn[a] = (uint64_t)(hh == g.dx ? (double)ncell[a] : std::ceil((double)ncell[a] * g.dx / hh)) + 1;
  // particles bucketed by their cell_of clamped into the grid
  const uint64_t nc = g.nx * g.ny * g.nz;
  std::vector<uint32_t> start(nc + 1, 0), order(s->pos.size());
  std::vector<uint64_t> cell(s->pos.size());
  for (size_t q = 0; q < s->pos.size(); ++q) {
    const V3 l = (s->pos[q] - g.origin) / g.dx;
    uint64_t i = as_usize(l.x), j = as_usize(l.y), k = as_usize(l.z);
    i = i < g.nx ? i : g.nx - 1, j = j < g.ny ? j : g.ny - 1, k = k < g.nz ? k : g.nz - 1;
    cell[q] = (k * g.ny + j) * g.nx + i;
    start[cell[q] + 1] += 1;
  }
  for (uint64_t c = 0; c < nc; ++c) start[c + 1] += start[c];
  {
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (size_t q = 0; q < s->pos.size(); ++q) order[fill[cell[q]]++] = (uint32_t)q;
  }
  const uint64_t nnodes = nn[0] * nn[1] * nn[2];
  s->surf_phi.assign(nnodes, 0.0);
  const double reach = 2.0 * r;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)nn[2]; ++k)
    for (uint64_t j = 0; j < nn[1]; ++j)
      for (uint64_t i = 0; i < nn[0]; ++i) {
        const double x = g.origin.x + (double)i * hh, y = g.origin.y + (double)j * hh,
                     z = g.origin.z + (double)k * hh;
        const int64_t i0 = bin_clamp(((x - reach) - g.origin.x) / g.dx, g.nx),
                      i1 = bin_clamp(((x + reach) - g.origin.x) / g.dx, g.nx);
        const int64_t j0 = bin_clamp(((y - reach) - g.origin.y) / g.dx, g.ny),
                      j1 = bin_clamp(((y + reach) - g.origin.y) / g.dx, g.ny);
        const int64_t k0 = bin_clamp(((z - reach) - g.origin.z) / g.dx, g.nz),
                      k1 = bin_clamp(((z + reach) - g.origin.z) / g.dx, g.nz);
        double best = std::numeric_limits<double>::infinity();
        for (int64_t kk = k0; kk <= k1; ++kk)
          for (int64_t jj = j0; jj <= j1; ++jj)
            for (int64_t ii = i0; ii <= i1; ++ii) {
              const uint64_t c = ((uint64_t)kk * g.ny + (uint64_t)jj) * g.nx + (uint64_t)ii;
              for (uint32_t q = start[c]; q < start[c + 1]; ++q) {
                const V3 p = s->pos[order[q]];
                const double ex = x - p.x, ey = y - p.y, ez = z - p.z;
                const double d2 = ex * ex + ey * ey + ez * ez;
                best = d2 < best ? d2 : best;
              }
            }
        const double d = std::sqrt(best);
        s->surf_phi[((uint64_t)k * nn[1] + j) * nn[0] + i] = (d < reach ? d : reach) - r;
      }
  std::vector<V3> v;
  surface_nets(s->surf_phi.data(), nn[0], nn[1], nn[2], g.origin, hh, v, s->surf_t);
  s->surf_v.resize(3 * v.size());
  for (size_t q = 0; q < v.size(); ++q) s->surf_v[3 * q] = v[q].x, s->surf_v[3 * q + 1] = v[q].y, s->surf_v[3 * q + 2] = v[q].z;
  sdr_surface_info o{};
  o.nx = nn[0], o.ny = nn[1], o.nz = nn[2];
  o.origin = sdr_vec3{g.origin.x, g.origin.y, g.origin.z};
  o.spacing = hh, o.radius = r;
  o.vertices = v.size(), o.triangles = s->surf_t.size() / 3;
  mesh_measure(v, s->surf_t, &o.area, &o.volume);
  s->surf_info = o;
  s->have_surface = true;
  *info = o;
  return 0;
}

int orc_solver_get_surface(void* h, double* vertices, uint64_t vcap, uint32_t* tris, uint64_t tcap) {
  Solver* s = (Solver*)h;
  if (!s->have_surface) return fail(SDR_E_STATE, "no surface built yet");
  if ((vertices && vcap < s->surf_info.vertices) || (tris && tcap < s->surf_info.triangles))
    return fail(SDR_E_INVALID, "output holds fewer vertices or triangles than the surface");
  if (vertices && !s->surf_v.empty()) memcpy(vertices, s->surf_v.data(), s->surf_v.size() * 8);
  if (tris && !s->surf_t.empty()) memcpy(tris, s->surf_t.data(), s->surf_t.size() * 4);
  return 0;
}

int orc_solver_get_surface_field(void* h, double* values, uint64_t cap) {
  Solver* s = (Solver*)h;
  if (!s->have_surface) return fail(SDR_E_STATE, "no surface built yet");
  if (!values || cap < s->surf_phi.size()) return fail(SDR_E_INVALID, "output holds fewer values than the node grid");
  memcpy(values, s->surf_phi.data(), s->surf_phi.size() * 8);
  return 0;
}

}  // extern "C"
