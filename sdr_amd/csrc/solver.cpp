// solver.cpp — host side of the MI355X substep: the reference's `Solver` and
// `pressure::project_capturing_with_source` surface (uselessgoddess/sdr
// src/solver.rs:143-310, src/pressure.rs:123-239) driving the gfx950 kernels, and the
// C ABI of include/sdr_amd.h on top of it.
//
// Layout on the device (one HIP stream per solver):
//   particles  SoA px,py,pz,vx,vy,vz (T) + rank (u32, the reference index), double
//              buffered; cell-sorted from `classify` until the first moving kernel
//   grid       u,v,w and saved_u,v,w (T), MAC layout of grid.rs:128-142
//   cells      u8 classification (0 Solid, 1 Fluid, 2 Air), static cavity mask
//   solve      b/r, z, s, p, As, q, precon (T) + u16 flags in the sheared layout
//              (device.hpp Shear): row J = i + j of plane k
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sdr_amd.h"
#include "host.hpp"
#include "kernels.hpp"
#include "mesh_sdf.hpp"
#include "metrics.hpp"
#include "mic_pipe.hpp"
#include "sort.hpp"
#include "surface.hpp"

namespace sdr {

// ------------------------------------------------------------------ host math mirror
// Host-side pieces of the substep (emission, CFL) restate the reference in f64 with the
// same libm, so the device path consumes the same particles and timesteps.
struct H3 {
  double x, y, z;
};
static inline H3 operator+(H3 a, H3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline H3 operator*(H3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline H3 operator/(H3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
static inline double hdot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline H3 hcross(H3 a, H3 o) {
  return {a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x};
}
static inline double hlen(H3 a) { return std::sqrt(hdot(a, a)); }
static inline H3 hnorm(H3 a) {  // math.rs:73-80
  double l = hlen(a);
  if (l > 1e-12) return a / l;
  return {0, 0, 0};
}
static inline H3 h3(sdr_vec3 v) { return {v.x, v.y, v.z}; }

static uint64_t xorshift(uint64_t& s) {  // solver.rs:127-134
  uint64_t x = s;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  s = x;
  return x;
}
static double unit(uint64_t& s) { return (double)(xorshift(s) >> 11) / (double)(1ull << 53); }

static int bits_for(uint64_t v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

// Copy a resident surface mesh to host buffers (either may be NULL).
static void download_mesh(const SurfaceNets& m, double* v, uint64_t vcap, uint32_t* t, uint64_t tcap,
                          hipStream_t st) {
  if ((v && vcap < m.nverts) || (t && tcap < m.ntris))
    throw Error(SDR_E_INVALID, "output holds fewer vertices or triangles than the surface");
  if (v && m.nverts) HIPCHK(hipMemcpyAsync(v, m.verts.p, m.nverts * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  if (t && m.ntris) HIPCHK(hipMemcpyAsync(t, m.tris.p, m.ntris * 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------ timing
struct Timer {
  bool on = false;
  // time only the pipelined MIC(0) sweep regions (mic_*), one call in kSweepSample of
  // each: an event pair costs the stream a few microseconds between kernels, so timing
  // every sweep would slow the timed work by ~2%
  bool sweeps_only = false;
  static constexpr uint64_t kSweepSample = 16;
  std::map<std::string, uint64_t> seen;
  hipStream_t st = nullptr;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, uint64_t>> acc;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    // timing only (read after a stream synchronisation): no system-scope fence, which
    // would write back every L2 at each record and stall the next kernel by ~5-10 us
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    return e;
  }
  size_t begin(const char* name) {
    if (!on || (sweeps_only && std::strncmp(name, "mic_", 4) != 0)) return 0;
    // the 16th, 32nd, ... call: the first sweep of a solve runs on cold caches
    if (sweeps_only && ++seen[name] % kSweepSample != 0) return 0;
    Pending p{name, get(), get()};
    HIPCHK(hipEventRecord(p.a, st));
    pending.push_back(p);
    return pending.size();
  }
  void end(size_t id) {
    if (!on || id == 0) return;
    HIPCHK(hipEventRecord(pending[id - 1].b, st));
  }
  void collect() {
    if (pending.empty()) return;
    HIPCHK(hipStreamSynchronize(st));
    for (auto& p : pending) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
      auto& e = acc[p.name];
      e.first += ms;
      e.second += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  ~Timer() {
    for (auto& p : pending) pool.push_back(p.a), pool.push_back(p.b);
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
struct Scope {
  Timer& t;
  size_t id;
  Scope(Timer& t_, const char* n) : t(t_), id(t_.begin(n)) {}
  ~Scope() { t.end(id); }
};

// ------------------------------------------------------------------ pressure projection
// `pressure::project_capturing_with_source` (pressure.rs:123-239) on device buffers.
template <class T>
struct Projector {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0;
  Shear sh{};
  bool exact_dot = true;
  hipStream_t st = nullptr;
  Timer* timer = nullptr;
  Buf b, z, s, s2, p, as, q, precon, flags, partials, scal, units;  // s2: the fused product's second direction buffer
  // the fast mode's partials of the backward sweep's dot(z, r) and of the axpy's max|r|,
  // read by the next kernel while `partials` holds the fused product's dot(s, As)
  Buf part_dot, part_res;
  Buf halo_x, halo_z, pipe_err, sink;
  Buf b64, p64, r64, norms;  // f64 refinement of the f32 solve
  uint32_t sweep_tag = 0;
  PinnedBuf host_scal;
  int mic_variant = 0;   // pipelined sweep tile shape (mic_pipe_shape); -1 = plane-by-plane
  int ncu = 0, dev = 0;  // compute units of the solver's device
  bool trace_refine = false;  // SDR_TRACE_REFINE=1: one stderr line per refinement pass
  // The precision form's solve: reliable updates (default), or restarted refinement
  // passes (SDR_REFINE_RESTART=1); kReplace is the drop of the recursive residual that
  // triggers a replacement (SDR_REPLACE_DROP overrides it, read once per solver).
  bool restart_refine = false;
  double kReplace = 1e-1;
  // All working memory is a view of caller-provided memory: nothing is allocated, and
  // memory is cleared by a fill kernel instead of a memset (graph-capturable).
  bool noalloc = false;
  // PCG iterations replayed from captured graphs, 8 at a time (the polling interval),
  // when SDR_PCG_GRAPH=1 (read once per solver); never while timing or in the no-alloc
  // projection.  The graphs hold exactly the launches of the eager loop, with sweep tags
  // read from tagbuf, so they give the same bits.  Off by default: on 48^3-128^3 grids
  // an iteration is bound by the sweeps' latency, and replays measured no faster.
  bool graph_ok = false;
  Buf tagbuf;
  hipGraphExec_t g_pcg = nullptr, g_rel_first = nullptr, g_rel_block = nullptr;
  double g_pcg_tol = 0.0;
  int64_t ncells() const { return (int64_t)nx * ny * nz; }

  ~Projector() {
    for (hipGraphExec_t g : {g_pcg, g_rel_first, g_rel_block})
      if (g) (void)hipGraphExecDestroy(g);
  }
  bool use_graphs() const { return graph_ok && !(timer && timer->on); }
  // Capture the launches `body` issues on the stream into an executable graph.
  template <class F>
  hipGraphExec_t capture(F&& body) {
    hipGraph_t g = nullptr;
    // relaxed: solvers on other threads keep allocating and copying meanwhile
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    try {
      body();
    } catch (...) {
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(st, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    return ex;
  }

  // Sizes of the working buffers (elements) and the sweep shape for a grid.
  struct Plan {
    size_t n, nb, hx, hz, nunits;
  };
  Plan plan(int x, int y, int zz, double h, hipStream_t stream, Timer* t) {
    nx = x, ny = y, nz = zz, dx = h, st = stream, timer = t;
    sh = Shear{nx, ny, nz, nx + ny - 1};
    Plan pl{};
    // the vector kernels' work units: the (row, 64-slot chunk) pairs of a plane that hold
    // in-grid slots, in row order
    pl.nunits = nx <= 256 * 64 ? (size_t)units_per_plane(nx, ny) : 0;
    const size_t n = (size_t)sh.size();
    pl.n = n;
    int nb = std::max(dot_blocks(sh, true), dot_blocks(sh, false));
    // both precisions: 7-plane tiles with an I/O wave, halo prefetch depth 4 (k_mic_sweep3);
    // SDR_MIC_VARIANT (read once per solver) selects another tested shape
    mic_variant = kMicDefaultVariant;
    if (const char* e = getenv("SDR_MIC_VARIANT")) mic_variant = atoi(e);
    if (mic_variant >= mic_pipe_variants()) mic_variant = kMicDefaultVariant;
    if (const char* e = getenv("SDR_TRACE_REFINE")) trace_refine = atoi(e) != 0;
    if (const char* e = getenv("SDR_REFINE_RESTART")) restart_refine = atoi(e) != 0;
    if (const char* e = getenv("SDR_REPLACE_DROP")) kReplace = atof(e);
    graph_ok = false;
    if (const char* e = getenv("SDR_PCG_GRAPH")) graph_ok = atoi(e) != 0;
    // The pipelined sweeps address with 32-bit byte offsets.
    const size_t lim = (size_t)1 << 31;
    if (n * sizeof(T) >= lim) mic_variant = -1;
    size_t hx = 0, hz = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // halo words of the solve variant and of the factor's variant (both share them)
    for (int v : {mic_variant < 0 ? 0 : mic_variant, sizeof(T) == 8 ? 1 : 0}) {
      MicPipeShape ps = mic_pipe_shape(sh, v, ncu);
      size_t a, c;
      mic_pipe_halo_sizes(sh, ps, &a, &c);
      hx = std::max(hx, a), hz = std::max(hz, c);
    }
    if (hx * sizeof(uint64_t) >= lim || hz * sizeof(uint64_t) >= lim) mic_variant = -1;
    for (int v = 0; v < mic_pipe_variants(); ++v) nb = std::max(nb, mic_pipe_dot_partials(mic_pipe_shape(sh, v, ncu)));
    nb = std::max(nb, spmv_update_blocks(sh));
    if (mic_variant < 0) hx = hz = 1;
    pl.nb = (size_t)nb, pl.hx = hx, pl.hz = hz;
    return pl;
  }

  void init(int x, int y, int zz, double h, hipStream_t stream, Timer* t) {
    const Plan pl = plan(x, y, zz, h, stream, t);
    if (pl.nunits) {
      std::vector<uint32_t> u;
      for (int J = 0; J < sh.NJ; ++J) {
        const int imin = std::max(0, J - ny + 1), imax = std::min(J, nx - 1);
        for (int c = imin / 64; c <= imax / 64; ++c) u.push_back(((uint32_t)J << 8) | (uint32_t)c);
      }
      units.alloc(u.size() * sizeof(uint32_t));
      HIPCHK(hipMemcpy(units.p, u.data(), u.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      sh.units = units.as<uint32_t>();
      sh.units_per_plane = (int)u.size();
    }
    for (Buf* bb : {&b, &z, &s, &p, &as, &q, &precon}) bb->alloc(pl.n * sizeof(T));
    flags.alloc(pl.n * sizeof(uint16_t));
    scal.alloc(sizeof(PcgScalars));
    host_scal.ensure(sizeof(PcgScalars) + 16);
    partials.alloc(pl.nb * sizeof(double));
    part_dot.alloc(pl.nb * sizeof(double));
    part_res.alloc(pl.nb * sizeof(double));
    halo_x.alloc(pl.hx * sizeof(uint64_t));
    halo_z.alloc(pl.hz * sizeof(uint64_t));
    pipe_err.alloc(sizeof(int));
    sink.alloc(1024 * sizeof(T));
    s2.alloc(pl.n * sizeof(T));
    HIPCHK(hipMemsetAsync(halo_x.p, 0, halo_x.bytes, st));
    HIPCHK(hipMemsetAsync(halo_z.p, 0, halo_z.bytes, st));
    HIPCHK(hipMemsetAsync(pipe_err.p, 0, sizeof(int), st));
    // the pipelined sweeps store only in-grid slots: the sheared slots outside the grid
    // of their outputs hold +0 from here on
    HIPCHK(hipMemsetAsync(q.p, 0, q.bytes, st));
    HIPCHK(hipMemsetAsync(z.p, 0, z.bytes, st));
    if (graph_ok) {
      tagbuf.alloc(sizeof(uint32_t));
      const uint32_t t0 = kDeviceTagBase - 1;  // the first launch_tag_next gives the base
      HIPCHK(hipMemcpyAsync(tagbuf.p, &t0, sizeof t0, hipMemcpyHostToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
    }
  }

  // Bytes of caller memory init_noalloc carves for a plan (256-byte aligned pieces).
  static size_t up256(size_t b) { return (b + 255) / 256 * 256; }
  static size_t noalloc_bytes(const Plan& pl) {
    return 8 * up256(pl.n * sizeof(T)) + up256(pl.n * 2) + 3 * up256(pl.nb * 8) + up256(sizeof(PcgScalars)) +
           up256(pl.nunits * 4) + up256(pl.hx * 8) + up256(pl.hz * 8) + up256(sizeof(int)) + up256(1024 * sizeof(T));
  }
  // The same working set as init, as views of [base, base + cap): no allocation, and the
  // memory that must start at zero (halo words, the wait-error flag) cleared by kernels.
  void init_noalloc(int x, int y, int zz, double h, hipStream_t stream, Timer* t, uint8_t* base, size_t cap) {
    const Plan pl = plan(x, y, zz, h, stream, t);
    const size_t need = noalloc_bytes(pl);
    if (need > cap)
      throw Error(SDR_E_INVALID, "grid " + std::to_string(x) + "x" + std::to_string(y) + "x" + std::to_string(zz) +
                                     " needs " + std::to_string(need) + " bytes of working memory; the no-alloc "
                                     "projection holds " + std::to_string(cap));
    noalloc = true;
    graph_ok = false;  // the caller may capture this projection itself
    size_t off = 0;
    auto take = [&](Buf& bf, size_t bytes) {
      bf.view(base + off, up256(bytes));
      off += up256(bytes);
    };
    for (Buf* bb : {&b, &z, &s, &s2, &p, &as, &q, &precon}) take(*bb, pl.n * sizeof(T));
    take(flags, pl.n * 2);
    take(partials, pl.nb * 8);
    take(part_dot, pl.nb * 8);
    take(part_res, pl.nb * 8);
    take(scal, sizeof(PcgScalars));
    take(units, pl.nunits * 4);
    take(halo_x, pl.hx * 8);
    take(halo_z, pl.hz * 8);
    take(pipe_err, sizeof(int));
    take(sink, 1024 * sizeof(T));
    if (pl.nunits) {
      launch_build_units(nx, ny, units.as<uint32_t>(), st);
      sh.units = units.as<uint32_t>();
      sh.units_per_plane = (int)pl.nunits;
    }
    zero(halo_x, halo_x.bytes);
    zero(halo_z, halo_z.bytes);
    zero(pipe_err, sizeof(int));
    // the sweeps store only in-grid slots; an earlier call may have left another grid's
    // values in this memory
    zero(q, pl.n * sizeof(T));
    zero(z, pl.n * sizeof(T));
  }
  void zero(Buf& bf, size_t bytes) {
    if (noalloc)
      launch_zero_u32(bf.as<uint32_t>(), (int64_t)((bytes + 3) / 4), st);
    else
      HIPCHK(hipMemsetAsync(bf.p, 0, bytes, st));
  }

  PcgScalars* sc() const { return scal.as<PcgScalars>(); }

  // The fast reduction mode fuses dot(z, r) into the backward solve (I/O-wave sweeps only).
  bool fused() const { return !exact_dot && mic_variant >= 0 && mic_pipe_shape(sh, mic_variant, ncu).kind == 4; }

  // dot (backward solves of the fused mode): the sweep leaves the partials of
  // dot(q, q) = dot(z, r) in `partials`.
  void mic_sweep(int op, const T* in, T* out, bool dot = false, bool dev_tag = false) {
    if (mic_variant < 0) {
      const int planes = sh.NJ + sh.nz - 1;
      for (int sgm = 0; sgm < planes; ++sgm)
        launch_mic_plane<T>(op, sh, flags.as<uint16_t>(), precon.as<T>(), in, out, sgm, sc(), st);
      return;
    }
    // the factor runs the one-plane-per-wave kernel (kind 0) with 8-plane tiles (256^3 f32:
    // 1.14 ms against 1.27 with 16)
    int variant = mic_variant;
    if (op == kMicBuild && mic_pipe_shape(sh, variant, ncu).kind != 0) variant = 1;
    MicPipeShape ps = mic_pipe_shape(sh, variant, ncu);
    if (dev_tag) {
      launch_tag_next(tagbuf.as<uint32_t>(), st);
    } else if (++sweep_tag >= kDeviceTagBase) {  // host tags wrapped: clear the halo words
      zero(halo_x, halo_x.bytes);
      zero(halo_z, halo_z.bytes);
      sweep_tag = 1;
    }
    MicPipeArgs<T> pa{halo_x.as<uint64_t>(), halo_z.as<uint64_t>(), sink.as<T>(), pipe_err.as<int>(),
                      1L << 22, ps.ntile_i, ps.ntile_k, sweep_tag, (uint32_t)halo_x.bytes,
                      (uint32_t)halo_z.bytes, ncu, dot ? 1 : 0, part_dot.as<double>(), fused() ? 1 : 0,
                      dev_tag ? tagbuf.as<uint32_t>() : nullptr};
    launch_mic_pipe<T>(op, variant, sh, flags.as<uint16_t>(), precon.as<T>(), in, out, sc(), pa, st);
  }

  // A bounded wait of a pipelined sweep expired: the solve result is not valid.
  void check_pipe() {
    int e = 0;
    HIPCHK(hipMemcpyAsync(&e, pipe_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (e) {
      zero(pipe_err, sizeof(int));
      throw Error(SDR_E_NUMERIC, "pipelined MIC(0) sweep: a tile's bounded wait expired");
    }
  }

  void dot(const T* a, const T* bb, int op, double tol) {
    int nb = dot_blocks(sh, exact_dot);
    launch_dot_partials<T>(sh, a, bb, partials.as<double>(), nb, exact_dot, sc(), st);
    launch_dot_finish(partials.as<double>(), nb, exact_dot, op, sc(), tol, st);
  }

  // The fused step head of the fast mode: beta from the last backward sweep's partials,
  // s = z + beta s (s = z on the first step) and As = A s with the partials of
  // dot(s, As), which the axpy combines into denom and alpha.  The new s goes to the
  // other direction buffer (sbuf = {s, s2}, PcgScalars::spar).
  void fused_product() {
    launch_spmv_update<T>(sh, flags.as<uint16_t>(), z.as<T>(), s.as<T>(), s2.as<T>(), as.as<T>(), sc(),
                          partials.as<double>(), part_dot.as<double>(),
                          mic_pipe_dot_partials(mic_pipe_shape(sh, mic_variant, ncu)), st);
  }
  // p += alpha s, r -= alpha As with max|r| after the fused product (denom and alpha
  // combined in the kernel), then the stopping test.
  void fused_axpy(int mb, double tol) {
    launch_axpy_pr_par<T>(sh, p.as<T>(), b.as<T>(), s.as<T>(), s2.as<T>(), as.as<T>(), sc(), part_res.as<double>(),
                          partials.as<double>(), spmv_update_blocks(sh), mb, st);
    launch_dot_finish(part_res.as<double>(), mb, false, kOpResidual, sc(), tol, st);
  }
  // The current direction buffer (fused: by the polled scalars).
  Buf& cur_s(const PcgScalars& h) { return fused() && h.spar ? s2 : s; }

  PcgScalars read_scalars() {
    if (noalloc) {  // no pinned staging: a pageable copy
      PcgScalars r;
      HIPCHK(hipMemcpyAsync(&r, scal.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      return r;
    }
    HIPCHK(hipMemcpyAsync(host_scal.p, scal.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    PcgScalars r;
    memcpy(&r, host_scal.p, sizeof r);
    return r;
  }

  // pcg (pressure.rs:341-397) on b (becomes r); the solution is left in p.  rel: stop at
  // tol * max|b| instead of tol.  build: factor the preconditioner (else reuse precon).
  PcgScalars pcg(uint64_t max_iters, double tol, bool rel = false, bool build = true) {
    const size_t n = (size_t)sh.size();
    {
      Scope sc_(*timer, "pcg_setup");
      for (Buf* bb : {&z, &s, &s2, &p, &as, &q}) zero(*bb, n * sizeof(T));
      if (build) zero(precon, n * sizeof(T));
      zero(scal, sizeof(PcgScalars));
      int mb = maxabs_blocks(sh);
      launch_maxabs_partials<T>(sh, b.as<T>(), partials.as<double>(), mb, sc(), st);
      launch_dot_finish(partials.as<double>(), mb, false, kOpRhsNorm, sc(), tol, st, rel ? 1 : 0);
    }
    if (build) {
      Scope sc_(*timer, "mic_build");
      mic_sweep(kMicBuild, nullptr, nullptr);
    }
    T* r = b.as<T>();
    {
      Scope sc_(*timer, "mic_fwd");
      mic_sweep(kMicFwd, r, q.as<T>());
    }
    {
      Scope sc_(*timer, "mic_bwd");
      mic_sweep(kMicBwd, q.as<T>(), z.as<T>());
    }
    const bool fz = fused();
    // s = z (pressure.rs:367); the fused product takes it on the first step itself
    if (!fz) HIPCHK(hipMemcpyAsync(s.p, z.p, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    dot(z.as<T>(), r, kOpSigma0, tol);
    const int mb = maxabs_blocks(sh);
    const uint64_t check_every = 8;
    // one PCG iteration (pressure.rs:372-396); dev: sweep tags from the device word
    auto iteration = [&](bool dev) {
      if (fz) {
        {
          Scope sc1(*timer, "spmv_update");
          fused_product();
        }
        {
          Scope sc1(*timer, "axpy_norm");
          fused_axpy(mb, tol);
        }
        {
          Scope sc1(*timer, "mic_fwd");
          mic_sweep(kMicFwd, r, q.as<T>(), false, dev);
        }
        {
          Scope sc1(*timer, "mic_bwd_dot");
          mic_sweep(kMicBwd, q.as<T>(), z.as<T>(), true, dev);  // its dot is combined by the next product
        }
        return;
      }
      {
        Scope sc_(*timer, "spmv_dot");
        if (exact_dot) {
          launch_spmv<T>(sh, flags.as<uint16_t>(), s.as<T>(), as.as<T>(), sc(), nullptr, st);
          dot(s.as<T>(), as.as<T>(), kOpDenom, tol);
        } else {  // dot(s, As) fused into the product
          launch_spmv<T>(sh, flags.as<uint16_t>(), s.as<T>(), as.as<T>(), sc(), partials.as<double>(), st);
          launch_dot_finish(partials.as<double>(), cell_blocks(sh), false, kOpDenom, sc(), tol, st);
        }
      }
      {
        Scope sc_(*timer, "axpy_norm");
        launch_axpy_pr<T>(sh, p.as<T>(), r, s.as<T>(), as.as<T>(), sc(), partials.as<double>(), mb, st);
        launch_dot_finish(partials.as<double>(), mb, false, kOpResidual, sc(), tol, st);
      }
      {
        Scope sc_(*timer, "mic_fwd");
        mic_sweep(kMicFwd, r, q.as<T>(), false, dev);
      }
      {
        Scope sc_(*timer, "mic_bwd");
        mic_sweep(kMicBwd, q.as<T>(), z.as<T>(), false, dev);
      }
      {
        Scope sc_(*timer, "dot_update");
        dot(z.as<T>(), r, kOpSigmaNew, tol);
        launch_update_s<T>(sh, s.as<T>(), z.as<T>(), sc(), st);
      }
    };
    // blocks of check_every iterations, the scalars polled between blocks
    uint64_t it = 0;
    while (it < max_iters) {
      const uint64_t nblk = std::min<uint64_t>(check_every, max_iters - it);
      if (nblk == check_every && use_graphs()) {
        if (g_pcg && g_pcg_tol != tol) {  // the stopping tolerance is a launch argument
          HIPCHK(hipGraphExecDestroy(g_pcg));
          g_pcg = nullptr;
        }
        if (!g_pcg) {
          g_pcg = capture([&] {
            for (uint64_t k = 0; k < check_every; ++k) iteration(true);
          });
          g_pcg_tol = tol;
        }
        HIPCHK(hipGraphLaunch(g_pcg, st));
      } else {
        for (uint64_t k = 0; k < nblk; ++k) iteration(false);
      }
      it += nblk;
      if (it < max_iters && read_scalars().done) break;
    }
    return read_scalars();
  }

  // Mixed-precision iterative refinement (f32 storage): r = b - A p in f64, solve
  // A e = r / |r| by the f32 MIC(0)-PCG (its dots and norms accumulate in f64), p += |r| e,
  // until max|r| <= rel_tol * max|b|.  Each correction solve is asked for kInner of the
  // scaled residual (or less in the last pass), which f32 arithmetic still delivers.
  // Returns the total PCG iterations; the f64 pressure is left in p64.
  PcgScalars refine(uint64_t max_iters, double rel_tol, double* final_res) {
    const size_t n = (size_t)sh.size();
    const double kInner = 1e-4;
    const int mb = maxabs_blocks(sh);
    double* hn = (double*)host_scal.p;
    // with p = 0 the residual kernel gives r = b and max|b|
    HIPCHK(hipMemsetAsync(p64.p, 0, n * 8, st));
    launch_residual64(sh, flags.as<uint16_t>(), b64.as<double>(), p64.as<double>(), r64.as<double>(),
                      partials.as<double>(), mb, norms.as<double>(), st);
    HIPCHK(hipMemcpyAsync(hn, norms.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const double bnorm = hn[0];
    double rnorm = bnorm;
    PcgScalars tot{};
    uint64_t used = 0;
    bool first = true;
    for (int outer = 0; outer < 16 && rnorm > rel_tol * bnorm && used < max_iters; ++outer) {
      {
        Scope sc_(*timer, "refine");
        launch_scale_cast<T>(sh, r64.as<double>(), 1.0 / rnorm, b.as<T>(), st);
      }
      const double inner = std::max(kInner, 0.5 * rel_tol * bnorm / rnorm);
      PcgScalars r = pcg(max_iters - used, inner, false, first);
      if (mic_variant >= 0) check_pipe();
      first = false;
      used += (uint64_t)r.iters;
      tot.nan_seen |= r.nan_seen;
      {
        Scope sc_(*timer, "refine");
        launch_accum64<T>(sh, p64.as<double>(), p.as<T>(), rnorm, st);
        launch_residual64(sh, flags.as<uint16_t>(), b64.as<double>(), p64.as<double>(),
                          r64.as<double>(), partials.as<double>(), mb, norms.as<double>(), st);
        HIPCHK(hipMemcpyAsync(hn, norms.p, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
      }
      const double prev = rnorm;
      rnorm = hn[0];
      if (trace_refine)
        fprintf(stderr, "[refine] pass %d: %d iterations (f32 recursive residual %.3e of %.3e), "
                "f64 residual %.3e of max|b| %.3e -> %.3e\n", outer, r.iters, r.residual, inner,
                rnorm, bnorm, rnorm / bnorm);
      if (!(rnorm == rnorm)) {
        tot.nan_seen = 1;
        break;
      }
      if (r.iters == 0 || rnorm >= prev) break;  // no progress left in f32
    }
    tot.iters = (int32_t)used;
    tot.residual = rnorm;
    tot.rhs_norm = bnorm;
    if (final_res) *final_res = bnorm > 0 ? rnorm / bnorm : 0.0;
    return tot;
  }

  // Mixed-precision PCG with reliable updates (residual replacement; van der Vorst & Ye,
  // "Residual replacement strategies for Krylov subspace iterative methods", 2000): one
  // f32 Krylov run on the residual scaled by `scale` (its max norm at the last
  // replacement), with dots and norms in f64.  The f32 iterate is a correction that is
  // folded into the f64 solution p64 whenever the recursive residual has dropped by
  // kReplace (or reached the stopping bound); the f64 true residual r = b - A p64 then
  // replaces the recursive one, and the search direction and sigma are carried into the
  // new units (s * f, sigma * f^2 with f = old scale / new scale), so the Krylov space
  // built so far is kept instead of restarted.  The run stops when the true residual
  // reaches rel_tol * max|b|.  A replacement that has not halved the true residual since
  // the previous one restarts the space (s = z).  Returns the total PCG iterations; the
  // f64 pressure is left in p64.
  PcgScalars reliable(uint64_t max_iters, double rel_tol, double* final_res) {
    const size_t n = (size_t)sh.size();
    const int mb = maxabs_blocks(sh);
    double* hn = (double*)host_scal.p;
    auto true_residual = [&]() {
      launch_residual64(sh, flags.as<uint16_t>(), b64.as<double>(), p64.as<double>(), r64.as<double>(),
                        partials.as<double>(), mb, norms.as<double>(), st);
      HIPCHK(hipMemcpyAsync(hn, norms.p, 8, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      return hn[0];
    };
    HIPCHK(hipMemsetAsync(p64.p, 0, n * 8, st));
    const double bnorm = true_residual();  // p = 0: r = b
    PcgScalars tot{};
    tot.rhs_norm = bnorm;
    if (!(bnorm == bnorm)) {
      tot.nan_seen = 1;
      return tot;
    }
    if (!(bnorm > 0.0)) {  // zero right-hand side: zero pressure, no iterations
      if (final_res) *final_res = 0.0;
      return tot;
    }
    const double target = rel_tol * bnorm;
    double scale = bnorm, rnorm = bnorm, last = bnorm;
    T* r = b.as<T>();
    {
      Scope sc_(*timer, "refine");
      launch_scale_cast<T>(sh, r64.as<double>(), 1.0 / scale, r, st);
    }
    {
      Scope sc_(*timer, "pcg_setup");
      for (Buf* bb : {&z, &s, &s2, &p, &as, &q, &precon}) zero(*bb, n * sizeof(T));
      zero(scal, sizeof(PcgScalars));
      launch_maxabs_partials<T>(sh, r, partials.as<double>(), mb, sc(), st);
      launch_dot_finish(partials.as<double>(), mb, false, kOpRhsNorm, sc(), target / scale, st, 0);
    }
    {
      Scope sc_(*timer, "mic_build");
      mic_sweep(kMicBuild, nullptr, nullptr);
    }
    {
      Scope sc_(*timer, "mic_fwd");
      mic_sweep(kMicFwd, r, q.as<T>());
    }
    {
      Scope sc_(*timer, "mic_bwd");
      mic_sweep(kMicBwd, q.as<T>(), z.as<T>());
    }
    const bool fz = fused();
    if (!fz) HIPCHK(hipMemcpyAsync(s.p, z.p, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    dot(z.as<T>(), r, kOpSigma0, 0.0);
    const uint64_t check_every = 8;
    uint64_t it = 0;
    int replacements = 0, restarts = 0;
    bool converged = false;
    // An iteration is split where the scalars are polled: head = the product, step and
    // residual norm; tail = the preconditioner solve and the direction update.
    auto head = [&] {
      if (fz) {
        {
          Scope sc1(*timer, "spmv_update");
          fused_product();
        }
        Scope sc1(*timer, "axpy_norm");
        fused_axpy(mb, 0.0);
        return;
      }
      {
        Scope sc1(*timer, "spmv_dot");
        launch_spmv<T>(sh, flags.as<uint16_t>(), s.as<T>(), as.as<T>(), sc(), nullptr, st);
        dot(s.as<T>(), as.as<T>(), kOpDenom, 0.0);
      }
      {
        Scope sc1(*timer, "axpy_norm");
        launch_axpy_pr<T>(sh, p.as<T>(), r, s.as<T>(), as.as<T>(), sc(), partials.as<double>(), mb, st);
        launch_dot_finish(partials.as<double>(), mb, false, kOpResidual, sc(), 0.0, st);
      }
    };
    auto tail = [&](bool dev) {
      if (fz) {
        {
          Scope sc1(*timer, "mic_fwd");
          mic_sweep(kMicFwd, r, q.as<T>(), false, dev);
        }
        Scope sc1(*timer, "mic_bwd_dot");
        mic_sweep(kMicBwd, q.as<T>(), z.as<T>(), true, dev);  // its dot is combined by the next product
      } else {
        {
          Scope sc1(*timer, "mic_fwd");
          mic_sweep(kMicFwd, r, q.as<T>(), false, dev);
        }
        {
          Scope sc1(*timer, "mic_bwd");
          mic_sweep(kMicBwd, q.as<T>(), z.as<T>(), false, dev);
        }
        dot(z.as<T>(), r, kOpSigmaNew, 0.0);
      }
      if (!fz) {
        Scope sc1(*timer, "update_s");
        launch_update_s<T>(sh, s.as<T>(), z.as<T>(), sc(), st);
      }
    };
    bool head_done = false;  // the first block starts with a head, every later one with a tail
    while (it < max_iters) {
      // up to the next poll: check_every heads (and the tails between them), as graphs
      // when a whole block remains
      const uint64_t next = std::min<uint64_t>((it / check_every + 1) * check_every, max_iters);
      if (next - it == check_every && it % check_every == 0 && use_graphs()) {
        hipGraphExec_t& g = head_done ? g_rel_block : g_rel_first;
        if (!g) {
          const bool first = !head_done;
          g = capture([&] {
            for (uint64_t k = 0; k < check_every; ++k) {
              if (k > 0 || !first) tail(true);
              head();
            }
          });
        }
        HIPCHK(hipGraphLaunch(g, st));
        it = next;
      } else {
        while (it < next) {
          if (head_done || it > 0) tail(false);
          head();
          head_done = true;
          ++it;
        }
      }
      head_done = true;
      {
        const PcgScalars sc_h = read_scalars();
        if (sc_h.nan_seen) {
          tot.nan_seen = 1;
          break;
        }
        it = (uint64_t)sc_h.iters;  // iterations skipped after the loop stopped did nothing
        const bool reached = sc_h.residual <= sc_h.tol_eff;
        if (sc_h.done || sc_h.residual <= kReplace) {
          Scope sc1(*timer, "refine");
          launch_accum64<T>(sh, p64.as<double>(), p.as<T>(), scale, st);
          zero(p, n * sizeof(T));
          const double rn = true_residual();
          ++replacements;
          if (trace_refine)
            fprintf(stderr, "[reliable] it %llu: recursive %.3e (bound %.3e) -> true %.3e of %.3e (%.3e)%s\n",
                    (unsigned long long)it, sc_h.residual, sc_h.tol_eff, rn, bnorm, rn / bnorm,
                    sc_h.done && !reached ? " breakdown" : "");
          if (!(rn == rn)) {
            tot.nan_seen = 1;
            break;
          }
          rnorm = rn;
          if (rn <= target) {
            converged = true;
            break;
          }
          if (it >= max_iters) break;
          // no progress since the last replacement, or a breakdown: restart the space
          const bool restart = rn > 0.5 * last || (sc_h.done && !reached);
          last = rn;
          const double f = scale / rn;
          scale = rn;
          launch_scale_cast<T>(sh, r64.as<double>(), 1.0 / scale, r, st);
          Buf& sd = cur_s(sc_h);
          if (restart) {
            zero(sd, n * sizeof(T));
            ++restarts;
          } else {
            launch_scale_inplace<T>(sh, sd.as<T>(), f, st);
          }
          launch_pcg_rescale(sc(), f * f, target / scale, restart ? 1 : 0, st);
        } else if (sc_h.done) {
          break;
        }
      }
    }
    if (!converged && !tot.nan_seen) {  // the cap hit: fold in what the f32 run has
      launch_accum64<T>(sh, p64.as<double>(), p.as<T>(), scale, st);
      zero(p, n * sizeof(T));
      rnorm = true_residual();
    }
    if (mic_variant >= 0) check_pipe();
    if (trace_refine)
      fprintf(stderr, "[reliable] %llu iterations, %d replacements, %d restarts, residual %.3e of %.3e\n",
              (unsigned long long)it, replacements, restarts, rnorm, bnorm);
    tot.iters = (int32_t)it;
    tot.residual = rnorm;
    if (final_res) *final_res = rnorm / bnorm;
    return tot;
  }

  // Steps 1-5 of project_capturing_with_source on the device faces (MAC layout).
  // relative: stop at rel_tol * max|b| (f32 storage: by f64 iterative refinement).
  sdr_solve_report project(GridD<T> g, FacesD<T> f, const uint8_t* cells, const uint32_t* count,
                           const double* srcarr, double rho, double dt, uint64_t max_iters,
                           double tol, bool use_count_source, double rest, double k_vol,
                           T* pressure_out, int* nan_flag, bool relative = false,
                           double rel_tol = 0.0) {
    const bool ref = relative && sizeof(T) == 4;
    if (ref && !p64.p) {
      const size_t n = (size_t)sh.size();
      b64.alloc(n * 8), p64.alloc(n * 8), r64.alloc(n * 8), norms.alloc(64);
      // the residual kernel writes in-grid slots only; the others stay +0
      HIPCHK(hipMemsetAsync(r64.p, 0, n * 8, st));
    }
    {
      Scope sc_(*timer, "rhs");
      launch_enforce_solid<T>(g, f, cells, st);
      double inv_scale = rho * g.dx / dt;
      double unit_ = g.dx / dt;
      launch_rhs_flags<T>(g, sh, CFacesD<T>{f.u, f.v, f.w}, cells, count, srcarr, inv_scale,
                          use_count_source ? 1 : 0, rest, k_vol, unit_, 0.25 * unit_, b.as<T>(),
                          flags.as<uint16_t>(), ref ? b64.as<double>() : nullptr, st);
    }
    PcgScalars r;
    if (ref) {
      r = restart_refine ? refine(max_iters, rel_tol, nullptr) : reliable(max_iters, rel_tol, nullptr);
    } else {
      r = pcg(max_iters, relative ? rel_tol : tol, relative);
      if (mic_variant >= 0) check_pipe();
    }
    {
      Scope sc_(*timer, "gradient");
      double coef = dt / (rho * g.dx);
      if (ref) {
        launch_gradient64<T>(g, sh, f, cells, p64.as<double>(), coef, st);
        if (pressure_out) launch_unshear64<T>(sh, p64.as<double>(), pressure_out, st);
      } else {
        launch_gradient<T>(g, sh, f, cells, p.as<T>(), (T)coef, st);
        if (pressure_out) launch_unshear<T>(sh, p.as<T>(), pressure_out, st);
      }
    }
    if (nan_flag) *nan_flag = r.nan_seen;
    return sdr_solve_report{(uint64_t)r.iters, r.residual};
  }
};

// ------------------------------------------------------------------ solver
struct SolverBase {
  virtual ~SolverBase() = default;
};

template <class T>
struct Solver : SolverBase {
  // -- configuration (solver.rs:143-166)
  int nx, ny, nz;
  double dx;
  H3 origin;
  sdr_sdf sdf_desc{};
  sdr_fluid_params fluid{};
  sdr_needle needle{};
  bool has_outlet = false;
  sdr_aabb outlet{};
  sdr_solve_params solve{200, 1e-6};
  sdr_options opts{};
  // -- scalar state
  double time = 0;
  uint64_t steps = 0, last_pcg_iters = 0;
  uint32_t last_substeps = 0;
  double last_residual = 0;
  uint64_t drained = 0, emitted = 0, culled = 0;
  uint64_t rng = 0x5deece66dull | 1;
  double emit_accum = 0;
  uint64_t redist_active = 0, redist_moved = 0;
  // Decision pass: reservation rounds (default) or the single-thread reference loop.
  bool redist_sequential = false;
  uint64_t redist_rounds = 0;
  uint32_t rd_epoch = 0;
  int64_t n = 0, cap = 0;
  bool sorted_valid = false, have_saved = false;
  bool p2g_perface = false;
  // Heavy cells of the fused P2G (fast mode): cells holding at least this many particles
  // are pre-reduced and enter their faces' sums as one term each (SDR_P2G_HEAVY, read
  // once per solver; 0 walks every cell).  The table holds up to one cell in 64.
  uint32_t p2g_heavy = 32;
  Buf p2g_hcell, p2g_hpart;
  int cur = 0;
  // -- device
  hipStream_t st = nullptr;
  Timer timer;
  Buf part[2][7];
  Buf keys, keys_alt, vals, vals_alt, skey_t, sort_big, sort_ctr;
  Buf cstart, cend, count, sort_cursor, mask, static_cls, cells;
  Buf fu, fv, fw, su, sv, sw, ftmp, known0, known1, pressure, sdf_data;
  Buf temp;
  Buf keep, scanpos, cullflag, culled_ranks;
  Buf cellkey, act_rank, act_slot, act_rank_sorted, act_slot_sorted, dest;
  Buf counters, aos_pos, aos_vel, fill_cells, rd_resv, rd_second, rd_ctr;
  Buf rd_sidx, rd_boff, rd_src_cell, rd_src_off, rd_src_head, rd_src_lim, rd_src_fill, rd_pend_a, rd_pend_b,
      rd_big, rd_rec, rd_rmx;
  PinnedBuf pin;
  Projector<T> proj;
  // particle surfacing: node field and the resident mesh of the last sdr_solver_surface
  Buf surf_phi;
  SurfaceNets nets;
  NodeGrid surf_ng{0, 0, 0, 0, 0, 0, 0};
  double surf_r = 0;
  bool have_surface = false;

  GridD<T> grid() const { return GridD<T>{nx, ny, nz, (T)dx, (T)origin.x, (T)origin.y, (T)origin.z}; }
  SdfD<T> sdf() const {
    return SdfD<T>{(int)sdf_desc.nx, (int)sdf_desc.ny, (int)sdf_desc.nz, (T)sdf_desc.dx,
                   (T)sdf_desc.origin.x, (T)sdf_desc.origin.y, (T)sdf_desc.origin.z,
                   sdf_data.as<T>()};
  }
  BoxD<T> outlet_box() const {
    return BoxD<T>{has_outlet ? 1 : 0, (T)outlet.min.x, (T)outlet.min.y, (T)outlet.min.z,
                   (T)outlet.max.x, (T)outlet.max.y, (T)outlet.max.z};
  }
  // The static classification (cavity mask, outlet cells) is computed once, in f64 in
  // both precisions: the precision form then classifies exactly the cells the plain
  // path does, and only its particles' cell_of runs in f32.
  GridD<double> grid64() const { return GridD<double>{nx, ny, nz, dx, origin.x, origin.y, origin.z}; }
  BoxD<double> outlet_box64() const {
    return BoxD<double>{has_outlet ? 1 : 0, outlet.min.x, outlet.min.y, outlet.min.z,
                        outlet.max.x, outlet.max.y, outlet.max.z};
  }
  Parts<T> parts(int which) const {
    return Parts<T>{part[which][0].as<T>(), part[which][1].as<T>(), part[which][2].as<T>(),
                    part[which][3].as<T>(), part[which][4].as<T>(), part[which][5].as<T>(),
                    part[which][6].as<uint32_t>()};
  }
  int64_t ncells() const { return (int64_t)nx * ny * nz; }
  int64_t nu() const { return (int64_t)(nx + 1) * ny * nz; }
  int64_t nv() const { return (int64_t)nx * (ny + 1) * nz; }
  int64_t nw() const { return (int64_t)nx * ny * (nz + 1); }
  FacesD<T> faces() const { return FacesD<T>{fu.as<T>(), fv.as<T>(), fw.as<T>()}; }
  CFacesD<T> cfaces() const { return CFacesD<T>{fu.as<T>(), fv.as<T>(), fw.as<T>()}; }
  CFacesD<T> csaved() const { return CFacesD<T>{su.as<T>(), sv.as<T>(), sw.as<T>()}; }

  ~Solver() override {
    // the buffers go back to their pool (or hipFree) after this: nothing may still use them
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }

  // Solver::new (solver.rs:184-222).
  void create(const sdr_sdf* s, double sim_dx, const sdr_fluid_params* f, const sdr_needle* nd,
              const sdr_options* o) {
    opts = o ? *o : sdr_options{SDR_PRECISION_F64, SDR_DOT_EXACT, 0, 1e-8, SDR_SOLVE_REFERENCE, 0};
    HIPCHK(hipSetDevice(opts.device));
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    timer.st = st;
    sdf_desc = *s;
    sdf_desc.data = nullptr;
    fluid = *f;
    needle = *nd;
    // bounds = origin .. origin + (n-1)*dx (sdf.rs:119-129); dims = ceil(size/dx).max(1)
    double mx = s->origin.x + (double)(s->nx - 1) * s->dx;
    double my = s->origin.y + (double)(s->ny - 1) * s->dx;
    double mz = s->origin.z + (double)(s->nz - 1) * s->dx;
    double szx = mx - s->origin.x, szy = my - s->origin.y, szz = mz - s->origin.z;
    double fx = std::fmax(std::ceil(szx / sim_dx), 1.0), fy = std::fmax(std::ceil(szy / sim_dx), 1.0),
           fz = std::fmax(std::ceil(szz / sim_dx), 1.0);
    if (!(fx * fy * fz < 2.0e9)) throw Error(SDR_E_INVALID, "grid too large");
    nx = (int)fx, ny = (int)fy, nz = (int)fz;
    if (nx < 2 || ny < 2 || nz < 2) throw Error(SDR_E_INVALID, "grid needs at least 2 cells per axis");
    dx = sim_dx;
    origin = {s->origin.x, s->origin.y, s->origin.z};
    const int64_t nc = ncells();
    size_t nsdf = (size_t)(s->nx * s->ny * s->nz);
    sdf_data.alloc(nsdf * sizeof(T));
    {
      std::vector<T> tmp(nsdf);
      for (size_t t = 0; t < nsdf; ++t) tmp[t] = (T)s->data[t];
      HIPCHK(hipMemcpy(sdf_data.p, tmp.data(), nsdf * sizeof(T), hipMemcpyHostToDevice));
    }
    for (Buf* b : {&cstart, &cend, &count, &sort_cursor}) b->alloc((size_t)nc * 4);
    for (Buf* b : {&mask, &static_cls, &cells}) b->alloc((size_t)nc);
    for (Buf* b : {&fu, &su}) b->alloc((size_t)nu() * sizeof(T));
    for (Buf* b : {&fv, &sv}) b->alloc((size_t)nv() * sizeof(T));
    for (Buf* b : {&fw, &sw}) b->alloc((size_t)nw() * sizeof(T));
    int64_t nf = std::max(nu(), std::max(nv(), nw()));
    ftmp.alloc((size_t)nf * sizeof(T));
    known0.alloc((size_t)nf);
    known1.alloc((size_t)nf);
    pressure.alloc((size_t)nc * sizeof(T));
    counters.alloc(4 * sizeof(unsigned long long));
    for (Buf* b : {&fu, &fv, &fw, &su, &sv, &sw, &pressure})
      HIPCHK(hipMemsetAsync(b->p, 0, b->bytes, st));
    HIPCHK(hipMemsetAsync(cells.p, kAir, (size_t)nc, st));
    {
      Buf sdf64;  // the f64 node values for the mask (the device keeps them as T)
      const double* d64 = (const double*)sdf_data.p;
      if (sizeof(T) != sizeof(double)) {
        sdf64.alloc(nsdf * sizeof(double));
        HIPCHK(hipMemcpy(sdf64.p, s->data, nsdf * sizeof(double), hipMemcpyHostToDevice));
        d64 = sdf64.as<double>();
      }
      SdfD<double> sd{(int)sdf_desc.nx, (int)sdf_desc.ny, (int)sdf_desc.nz, sdf_desc.dx, sdf_desc.origin.x,
                      sdf_desc.origin.y, sdf_desc.origin.z, d64};
      launch_cavity_mask<double>(grid64(), sd, mask.as<uint8_t>(), st);
      HIPCHK(hipStreamSynchronize(st));
    }
    launch_static_cls<double>(grid64(), mask.as<uint8_t>(), outlet_box64(), static_cls.as<uint8_t>(), st);
    if (const char* e = getenv("SDR_REDIST_SEQUENTIAL")) redist_sequential = atoi(e) != 0;
    // SDR_P2G_PERFACE=1 (read once per solver) selects the face-by-face gather
    if (const char* e = getenv("SDR_P2G_PERFACE")) p2g_perface = atoi(e) != 0;
    if (const char* e = getenv("SDR_P2G_HEAVY")) p2g_heavy = (uint32_t)std::max(0, atoi(e));
    if (opts.dot_mode != SDR_DOT_EXACT && p2g_heavy) {
      const size_t hcap = (size_t)(nc / 64 + 1);
      p2g_hcell.alloc((hcap + 1) * 4);  // entries, then the counter
      p2g_hpart.alloc(hcap * kP2gHeavyVals * sizeof(T));
    }
    proj.exact_dot = opts.dot_mode == SDR_DOT_EXACT;
    proj.init(nx, ny, nz, dx, st, &timer);
    reserve(1024);
    HIPCHK(hipStreamSynchronize(st));
  }

  // Grow particle storage (keeps the live particles of the current buffer).
  void reserve(int64_t want) {
    if (want <= cap) return;
    int64_t nc = std::max<int64_t>(want, cap + cap / 2);
    if (nc > (int64_t)0x7fffffff) throw Error(SDR_E_INVALID, "too many particles");
    for (int w = 0; w < 2; ++w)
      for (int a = 0; a < 7; ++a) {
        size_t es = a == 6 ? 4 : sizeof(T);
        Buf nb;
        nb.alloc((size_t)nc * es);
        if (w == cur && n > 0) HIPCHK(hipMemcpyAsync(nb.p, part[w][a].p, (size_t)n * es, hipMemcpyDeviceToDevice, st));
        part[w][a].swap(nb);
      }
    for (Buf* b : {&keys, &keys_alt, &vals, &vals_alt, &keep, &scanpos, &cullflag, &culled_ranks,
                   &cellkey, &act_rank, &act_slot, &act_rank_sorted,
                   &act_slot_sorted, &dest})
      b->alloc((size_t)nc * 4);
    if (opts.dot_mode == SDR_DOT_EXACT) skey_t.alloc((size_t)nc * 4);
    // segments longer than a thread's register sort: each holds more than 16 particles
    sort_big.alloc((size_t)(nc / 16 + 1) * 4);
    if (!sort_ctr.p) sort_ctr.alloc(64);
    temp.alloc(scan_u32_temp(std::max<int64_t>(nc, ncells())));
    cap = nc;
  }

  bool exact() const { return opts.dot_mode == SDR_DOT_EXACT; }

  // ---------------------------------------------------------------- stages
  // emit (solver.rs:313-339): host-side (libm cos/sin), appended with rank = index.
  void emit(double dt) {
    uint64_t ppc = fluid.particles_per_cell > 1 ? fluid.particles_per_cell : 1;
    double pv = dx * dx * dx / (double)ppc;
    if (pv <= 0.0 || needle.flow_rate <= 0.0 || time >= needle.inject_until) return;
    emit_accum += needle.flow_rate * dt / pv;
    double cnt = std::floor(emit_accum);
    uint64_t m = cnt > 0 ? (uint64_t)cnt : 0;
    emit_accum -= cnt;
    if (m == 0) return;
    Scope sc_(timer, "emit");
    H3 axis = hnorm(h3(needle.axis));
    H3 a = std::fabs(axis.x) < 0.9 ? H3{1, 0, 0} : H3{0, 1, 0};  // solver.rs:933-942
    H3 e1 = hnorm(hcross(axis, a));
    H3 e2 = hnorm(hcross(axis, e1));
    double area = M_PI * needle.radius * needle.radius;
    double speed = area > 0.0 ? needle.flow_rate / area : 0.0;
    H3 vel = axis * speed;
    reserve(n + (int64_t)m);
    pin.ensure(m * 6 * sizeof(double));
    double* hp = (double*)pin.p;
    double* hv = hp + 3 * m;
    for (uint64_t t = 0; t < m; ++t) {
      double r = needle.radius * std::sqrt(unit(rng));
      double theta = 2.0 * M_PI * unit(rng);
      double along = (unit(rng) - 0.5) * dx;
      H3 p = h3(needle.tip) + e1 * (r * std::cos(theta)) + e2 * (r * std::sin(theta)) + axis * along;
      hp[3 * t] = p.x, hp[3 * t + 1] = p.y, hp[3 * t + 2] = p.z;
      hv[3 * t] = vel.x, hv[3 * t + 1] = vel.y, hv[3 * t + 2] = vel.z;
    }
    aos_pos.ensure(m * 3 * sizeof(double));
    aos_vel.ensure(m * 3 * sizeof(double));
    HIPCHK(hipMemcpyAsync(aos_pos.p, hp, m * 3 * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(aos_vel.p, hv, m * 3 * sizeof(double), hipMemcpyHostToDevice, st));
    launch_from_aos<T>(parts(cur), (int64_t)m, n, aos_pos.as<double>(), aos_vel.as<double>(), st);
    HIPCHK(hipStreamSynchronize(st));  // the pinned staging buffer is reused next substep
    n += (int64_t)m;
    emitted += m;
    sorted_valid = false;
  }

  // Sort of the particles by (clamped) cell; cell segments [cstart, cend).  Inside a
  // segment: reference order in exact mode, the previous order otherwise (stable).
  void sort_particles() {
    Scope sc_(timer, "sort");
    const CellSortBufs b{keys.as<uint32_t>(),   vals.as<uint32_t>(),   keys_alt.as<uint32_t>(),
                         vals_alt.as<uint32_t>(), skey_t.as<uint32_t>(), cstart.as<uint32_t>(),
                         cend.as<uint32_t>(),   sort_cursor.as<uint32_t>(), sort_big.as<uint32_t>(),
                         sort_ctr.as<uint32_t>()};
    cell_sort<T>(parts(cur), n, grid(), exact(), b, temp.p, temp.bytes, st);
    if (n == 0) {
      sorted_valid = true;
      return;
    }
    launch_gather_parts<T>(parts(cur), parts(1 - cur), vals.as<uint32_t>(), n, st);
    cur = 1 - cur;
    sorted_valid = true;
  }

  // classify (solver.rs:342-371) + the per-cell counts of volume_source.
  void classify() {
    sort_particles();
    Scope sc_(timer, "classify");
    launch_classify<T>(static_cls.as<uint8_t>(), parts(cur), n, grid(), outlet_box(), cells.as<uint8_t>(),
                       count.as<uint32_t>(), st);
  }
  void ensure_counts() {
    if (sorted_valid) return;
    // cells stay as set; only the segments and counts are refreshed
    Buf keep_cells;
    keep_cells.alloc((size_t)ncells());
    HIPCHK(hipMemcpyAsync(keep_cells.p, cells.p, (size_t)ncells(), hipMemcpyDeviceToDevice, st));
    classify();
    HIPCHK(hipMemcpyAsync(cells.p, keep_cells.p, (size_t)ncells(), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }

  // p2g (solver.rs:560-585)
  void p2g() {
    ensure_counts();
    Scope sc_(timer, "p2g");
    if (!exact() && !p2g_perface) {
      P2gHeavy h{};
      if (p2g_heavy && p2g_hcell.p) {
        const uint32_t hcap = (uint32_t)(p2g_hcell.bytes / 4 - 1);
        // the sort's per-cell cursor is free once the particles are sorted: entry per cell
        h = P2gHeavy{sort_cursor.as<uint32_t>(), p2g_hcell.as<uint32_t>(), p2g_hcell.as<uint32_t>() + hcap,
                     p2g_hpart.p, hcap, p2g_heavy};
        launch_p2g_heavy<T>(grid(), parts(cur), cstart.as<uint32_t>(), cend.as<uint32_t>(), h, st);
      }
      if (launch_p2g_fused<T>(grid(), parts(cur), cstart.as<uint32_t>(), cend.as<uint32_t>(), faces(), h, st))
        return;
    }
    launch_p2g<T>(0, grid(), parts(cur), cstart.as<uint32_t>(), cend.as<uint32_t>(), fu.as<T>(), exact(), st);
    launch_p2g<T>(1, grid(), parts(cur), cstart.as<uint32_t>(), cend.as<uint32_t>(), fv.as<T>(), exact(), st);
    launch_p2g<T>(2, grid(), parts(cur), cstart.as<uint32_t>(), cend.as<uint32_t>(), fw.as<T>(), exact(), st);
  }
  void save_grid() {  // solver.rs:587-591
    Scope sc_(timer, "save_forces");
    HIPCHK(hipMemcpyAsync(su.p, fu.p, (size_t)nu() * sizeof(T), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(sv.p, fv.p, (size_t)nv() * sizeof(T), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(sw.p, fw.p, (size_t)nw() * sizeof(T), hipMemcpyDeviceToDevice, st));
    have_saved = true;
  }
  void add_forces(double dt) {  // solver.rs:593-604
    if (fluid.gravity.x != 0.0) launch_axpy_const<T>(fu.as<T>(), nu(), (T)(fluid.gravity.x * dt), st);
    if (fluid.gravity.y != 0.0) launch_axpy_const<T>(fv.as<T>(), nv(), (T)(fluid.gravity.y * dt), st);
    if (fluid.gravity.z != 0.0) launch_axpy_const<T>(fw.as<T>(), nw(), (T)(fluid.gravity.z * dt), st);
  }
  // volume_source + project_capturing_with_source (solver.rs:290-299)
  void project(double dt) {
    ensure_counts();
    uint64_t ppc = fluid.particles_per_cell > 1 ? fluid.particles_per_cell : 1;
    bool use_src = fluid.volume_stiffness > 0.0 && dt > 0.0;
    int nanf = 0;
    sdr_solve_report r = proj.project(grid(), faces(), cells.as<uint8_t>(), count.as<uint32_t>(),
                                      nullptr, fluid.density, dt, solve.max_iters, solve.tol, use_src,
                                      (double)ppc, fluid.volume_stiffness, pressure.as<T>(), &nanf,
                                      opts.solve_mode == SDR_SOLVE_RELATIVE, opts.refine_tol);
    last_pcg_iters = r.iters;
    last_residual = r.residual;
    // the reference's NaN-blind inf_norm would report such a solve as converged
    if (nanf) throw Error(SDR_E_NUMERIC, "pressure solve: NaN in the residual (non-finite faces or source?)");
  }
  // extrapolate_velocities (solver.rs:608-647) with 4 passes of extrapolate (960-1009)
  void extrapolate() {
    Scope sc_(timer, "extrapolate");
    T* arr[3] = {fu.as<T>(), fv.as<T>(), fw.as<T>()};
    for (int c = 0; c < 3; ++c) {
      int ax = nx + (c == 0), ay = ny + (c == 1), az = nz + (c == 2);
      launch_extrap_init<T>(grid(), c, cells.as<uint8_t>(), known0.as<uint8_t>(), st);
      T* a = arr[c];
      T* b = ftmp.as<T>();
      uint8_t* k0 = known0.as<uint8_t>();
      uint8_t* k1 = known1.as<uint8_t>();
      for (int it = 0; it < 4; ++it) {
        launch_extrap_iter<T>(ax, ay, az, a, b, k0, k1, st);
        std::swap(a, b);
        std::swap(k0, k1);
      }
    }
  }
  void g2p() {  // solver.rs:650-677
    if (!have_saved) throw Error(SDR_E_STATE, "g2p needs the saved post-P2G grid");
    Scope sc_(timer, "g2p");
    double fr = fluid.flip_ratio < 0.0 ? 0.0 : (fluid.flip_ratio > 1.0 ? 1.0 : fluid.flip_ratio);
    launch_g2p<T>(parts(cur), n, grid(), cfaces(), csaved(), (T)fr, st);
  }
  void advect(double dt) {  // solver.rs:693-711
    Scope sc_(timer, "advect");
    launch_advect<T>(parts(cur), n, grid(), cfaces(), sdf(), (T)dt, st);
    sorted_valid = false;
  }
  // enforce_particle_bounds (solver.rs:736-755) + ParticleSet::retain (particles.rs:38-49)
  void bounds() {
    if (n == 0) return;
    Scope sc_(timer, "bounds");
    HIPCHK(hipMemsetAsync(counters.p, 0, 4 * sizeof(unsigned long long), st));
    launch_bounds<T>(parts(cur), n, sdf(), outlet_box(), keep.as<uint32_t>(), cullflag.as<uint32_t>(),
                     counters.as<unsigned long long>(), st);
    HIPCHK(exclusive_sum_u32(temp.p, temp.bytes, keep.as<uint32_t>(), scanpos.as<uint32_t>(), n, st));
    pin.ensure(64);
    unsigned long long* hc = (unsigned long long*)pin.p;
    HIPCHK(hipMemcpyAsync(hc, counters.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    uint64_t dr = hc[0], cu = hc[1];
    drained += dr;
    culled += cu;
    int64_t removed = (int64_t)(dr + cu);
    if (removed == 0) return;
    int64_t kept = n - removed;
    launch_compact<T>(parts(cur), keep.as<uint32_t>(), scanpos.as<uint32_t>(), n, kept, cullflag.as<uint32_t>(),
                      culled_ranks.as<uint32_t>(), st);
    // ranks of the survivors: minus the removed ranks below them (keep, cullflag and
    // scanpos are free once the compaction has run)
    const int64_t nw = (n + 31) / 32;
    launch_rank_bits(culled_ranks.as<uint32_t>(), removed, n, keep.as<uint32_t>(), cullflag.as<uint32_t>(), st);
    HIPCHK(exclusive_sum_u32(temp.p, temp.bytes, cullflag.as<uint32_t>(), scanpos.as<uint32_t>(), nw, st));
    launch_rank_fix(part[cur][6].as<uint32_t>(), kept, keep.as<uint32_t>(), scanpos.as<uint32_t>(), st);
    n = kept;
    sorted_valid = false;
  }
  // The order-dependent decision pass of redistribute (solver.rs:476-555) by
  // deterministic reservation rounds over cell buckets (redist.hip): sources (cells over
  // the cap) are numbered and given buckets by two scans over the cells, every particle
  // of a source lands in its bucket, the first count - cap entries of each bucket are put
  // in reference order, and rounds run until no source has candidates left.  Chosen
  // destinations land in dest[bucket entry]; returns (bucket entries, moves).
  std::pair<int64_t, uint64_t> decide(uint32_t capc, const RedistDirs& dn) {
    const int64_t nc = ncells();
    if (rd_resv.bytes < (size_t)nc * 8) {
      rd_resv.alloc((size_t)nc * 8);
      rd_second.alloc((size_t)nc * 8);
      HIPCHK(hipMemsetAsync(rd_resv.p, 0, rd_resv.bytes, st));
      HIPCHK(hipMemsetAsync(rd_second.p, 0, rd_second.bytes, st));
      rd_epoch = 0;
      for (Buf* b : {&rd_sidx, &rd_boff, &rd_src_cell, &rd_src_off, &rd_src_head, &rd_src_lim, &rd_src_fill,
                     &rd_pend_a, &rd_pend_b, &rd_big})
        b->alloc((size_t)nc * 4);
      rd_rec.alloc((size_t)nc * 32);  // per entry of either pending list: cell, head, lim, head rank
      rd_rmx.alloc((size_t)nc);       // and its open receivers
    }
    if (!rd_ctr.p) rd_ctr.alloc(64);
    uint32_t* ctr = rd_ctr.as<uint32_t>();
    const uint32_t* cnt = count.as<uint32_t>();
    HIPCHK(exclusive_sum_over_cap(temp.p, temp.bytes, cnt, capc, false, rd_sidx.as<uint32_t>(), nc, st));
    HIPCHK(exclusive_sum_over_cap(temp.p, temp.bytes, cnt, capc, true, rd_boff.as<uint32_t>(), nc, st));
    launch_rd_prepare(cnt, rd_sidx.as<uint32_t>(), rd_boff.as<uint32_t>(), nc, capc,
                      rd_src_cell.as<uint32_t>(), rd_src_off.as<uint32_t>(), rd_src_head.as<uint32_t>(),
                      rd_src_lim.as<uint32_t>(), rd_src_fill.as<uint32_t>(), rd_pend_a.as<uint32_t>(), ctr, st);
    uint32_t* hc = (uint32_t*)pin.p;
    HIPCHK(hipMemcpyAsync(hc, ctr, 24, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int64_t nsrc = hc[0], nact = hc[1];
    const uint32_t max_count = hc[5];  // of the sources too large to sort in one workgroup, else 0
    if (nsrc == 0) return {0, 0};
    uint32_t* brank = act_rank.as<uint32_t>();
    uint32_t* bslot = act_slot.as<uint32_t>();
    launch_rd_buckets(cellkey.as<int32_t>(), cnt, n, capc, rd_sidx.as<uint32_t>(), nsrc,
                      rd_src_cell.as<uint32_t>(), rd_src_off.as<uint32_t>(), rd_src_fill.as<uint32_t>(),
                      part[cur][6].as<uint32_t>(), brank, bslot, act_rank_sorted.as<uint32_t>(),
                      act_slot_sorted.as<uint32_t>(), rd_big.as<uint32_t>(), ctr, max_count, st);
    HIPCHK(hipMemsetAsync(dest.p, 0xff, (size_t)nact * 4, st));
    // Each round commits at least the lowest pending candidate, so the candidates (at
    // most nact) bound the rounds; the pending count is read every few rounds and sizes
    // the launches that follow.
    const int check_every = 8;
    int64_t rounds = 0, pending = nsrc;
    while (pending > 0) {
      for (int r = 0; r < check_every; ++r) {
        if (++rd_epoch == 0) {  // epochs wrapped: no stale reservation may outrank a new one
          HIPCHK(hipMemsetAsync(rd_resv.p, 0, rd_resv.bytes, st));
          HIPCHK(hipMemsetAsync(rd_second.p, 0, rd_second.bytes, st));
          rd_epoch = 1;
        }
        launch_rd_round2(pending, (int)(rounds & 1), rounds == 0, rd_pend_a.as<uint32_t>(),
                         rd_pend_b.as<uint32_t>(), rd_src_cell.as<uint32_t>(), rd_src_head.as<uint32_t>(),
                         rd_src_lim.as<uint32_t>(), rd_rec.as<uint32_t>(), rd_rec.as<uint32_t>() + 4 * nc,
                         rd_rmx.as<uint8_t>(), brank, count.as<uint32_t>(),
                         mask.as<uint8_t>(), nx, ny, nz,
                         capc, dn, rd_resv.as<unsigned long long>(), rd_second.as<unsigned long long>(), rd_epoch,
                         dest.as<int32_t>(), ctr, st);
        ++rounds;
      }
      // the list written by the last round
      HIPCHK(hipMemcpyAsync(hc, ctr + 2 + (rounds & 1), 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      pending = hc[0];
      if (rounds > nact + check_every) throw Error(SDR_E_STATE, "redistribute: rounds did not converge");
    }
    redist_rounds += (uint64_t)rounds;
    HIPCHK(hipMemcpyAsync(hc + 2, ctr + 6, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    uint64_t m;
    memcpy(&m, hc + 2, 8);
    return {nact, m};
  }

  // redistribute (solver.rs:454-557)
  void redistribute() {
    redist_active = redist_moved = redist_rounds = 0;
    if (n == 0) return;
    Scope sc_(timer, "redistribute");
    uint64_t target = fluid.particles_per_cell > 1 ? fluid.particles_per_cell : 1;
    uint32_t capc = (uint32_t)std::ceil(std::fmax(fluid.redist_cap_factor, 1.0) * (double)target);
    H3 gdir = hnorm(h3(fluid.gravity));
    // dir . gdir of the unit directions -x,+x,-y,+y,-z,+z (math.rs dot, left to right)
    RedistDirs dn{{-1.0 * gdir.x + 0.0 * gdir.y + 0.0 * gdir.z, 1.0 * gdir.x + 0.0 * gdir.y + 0.0 * gdir.z,
                   0.0 * gdir.x + -1.0 * gdir.y + 0.0 * gdir.z, 0.0 * gdir.x + 1.0 * gdir.y + 0.0 * gdir.z,
                   0.0 * gdir.x + 0.0 * gdir.y + -1.0 * gdir.z, 0.0 * gdir.x + 0.0 * gdir.y + 1.0 * gdir.z}};
    pin.ensure(64);
    uint32_t* hn = (uint32_t*)pin.p;
    for (int sweep = 0; sweep < 4; ++sweep) {
      // the histogram of the sweep (solver.rs:464-469); after the first, the previous
      // sweep's decision counts corrected by its moves are that histogram
      if (sweep == 0 || redist_sequential) {
        HIPCHK(hipMemsetAsync(count.p, 0, (size_t)ncells() * 4, st));
        launch_hist<T>(parts(cur), n, grid(), cellkey.as<int32_t>(), count.as<uint32_t>(), st);
      }
      if (!redist_sequential) {
        auto [nact, moved] = decide(capc, dn);
        if (nact == 0) break;  // no cell over the cap: later sweeps see the same state
        redist_active += (uint64_t)nact;
        redist_moved += moved;
        launch_redist_apply<T>(parts(cur), act_slot.as<uint32_t>(), dest.as<int32_t>(), nact, grid(),
                               cfaces(), sdf(), cellkey.as<int32_t>(), count.as<uint32_t>(), st);
        continue;
      }
      // The single-thread reference loop over the active particles in reference order.
      launch_active_flags(cellkey.as<int32_t>(), count.as<uint32_t>(), n, capc, keep.as<uint32_t>(), st);
      HIPCHK(exclusive_sum_u32(temp.p, temp.bytes, keep.as<uint32_t>(), scanpos.as<uint32_t>(), n, st));
      HIPCHK(hipMemcpyAsync(hn, scanpos.as<uint32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(hn + 1, keep.as<uint32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      int64_t nact = (int64_t)hn[0] + hn[1];
      if (nact == 0) break;
      launch_active_compact(keep.as<uint32_t>(), scanpos.as<uint32_t>(), n, part[cur][6].as<uint32_t>(),
                            act_rank.as<uint32_t>(), act_slot.as<uint32_t>(), st);
      HIPCHK(sort_pairs_u32(act_rank.as<uint32_t>(), act_rank_sorted.as<uint32_t>(), act_slot.as<uint32_t>(),
                            act_slot_sorted.as<uint32_t>(), keys.as<uint32_t>(), keys_alt.as<uint32_t>(), nact, st));
      redist_active += (uint64_t)nact;
      unsigned long long* mv = counters.as<unsigned long long>() + 3;
      launch_redist_decide(act_slot_sorted.as<uint32_t>(), nact, cellkey.as<int32_t>(), count.as<uint32_t>(),
                           mask.as<uint8_t>(), nx, ny, nz, capc, gdir.x, gdir.y, gdir.z, dest.as<int32_t>(), mv,
                           st);
      HIPCHK(hipMemcpyAsync((unsigned long long*)pin.p + 4, mv, 8, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      redist_moved += ((unsigned long long*)pin.p)[4];
      launch_redist_apply<T>(parts(cur), act_slot_sorted.as<uint32_t>(), dest.as<int32_t>(), nact, grid(),
                             cfaces(), sdf(), nullptr, nullptr, st);
    }
    sorted_valid = false;
  }

  void run_stage(int stage, double dt) {
    switch (stage) {
      case SDR_STAGE_EMIT: emit(dt); break;
      case SDR_STAGE_CLASSIFY: classify(); break;
      case SDR_STAGE_P2G: p2g(); break;
      case SDR_STAGE_SAVE: save_grid(); break;
      case SDR_STAGE_FORCES: add_forces(dt); break;
      case SDR_STAGE_PROJECT: project(dt); break;
      case SDR_STAGE_EXTRAPOLATE: extrapolate(); break;
      case SDR_STAGE_G2P: g2p(); break;
      case SDR_STAGE_ADVECT: advect(dt); break;
      case SDR_STAGE_BOUNDS: bounds(); break;
      case SDR_STAGE_REDISTRIBUTE:
        if (fluid.redistribute) redistribute();
        break;
      case SDR_STAGE_FINISH:
        time += dt;
        steps += 1;
        break;
      default: throw Error(SDR_E_INVALID, "unknown stage");
    }
  }
  // g2p then advect as one pass over the particles (the same values as the two stages)
  void g2p_advect(double dt) {
    if (!have_saved) throw Error(SDR_E_STATE, "g2p needs the saved post-P2G grid");
    Scope sc_(timer, "g2p_advect");
    double fr = fluid.flip_ratio < 0.0 ? 0.0 : (fluid.flip_ratio > 1.0 ? 1.0 : fluid.flip_ratio);
    launch_g2p_advect<T>(parts(cur), n, grid(), cfaces(), csaved(), (T)fr, sdf(), (T)dt, st);
    sorted_valid = false;
  }
  void substep(double dt) {  // solver.rs:284-310
    for (int s = 0; s < SDR_STAGE_COUNT; ++s) {
      if (s == SDR_STAGE_G2P) {
        g2p_advect(dt);
        ++s;  // SDR_STAGE_ADVECT done too
        continue;
      }
      run_stage(s, dt);
    }
    timer.collect();
  }

  // cfl_dt (solver.rs:262-281) with MacGrid::max_speed (grid.rs:202-206) on the device.
  double cfl_dt() {
    HIPCHK(hipMemsetAsync(counters.p, 0, 3 * sizeof(unsigned long long), st));
    unsigned long long* c = counters.as<unsigned long long>();
    launch_maxabs_faces<T>(fu.as<T>(), nu(), c, st);
    launch_maxabs_faces<T>(fv.as<T>(), nv(), c + 1, st);
    launch_maxabs_faces<T>(fw.as<T>(), nw(), c + 2, st);
    pin.ensure(64);
    HIPCHK(hipMemcpyAsync(pin.p, counters.p, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double m[3];
    memcpy(m, pin.p, sizeof m);
    double grid_speed = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    double area = M_PI * needle.radius * needle.radius;
    double jet = time < needle.inject_until ? (area > 0.0 ? needle.flow_rate / area : 0.0) : 0.0;
    double g = std::sqrt(hlen(h3(fluid.gravity)) * dx);
    double speed = std::fmax(std::fmax(std::fmax(grid_speed, jet), g), 1e-6);
    return std::fmin(dx / speed, 0.01);
  }
  void step(double dt) {  // solver.rs:237-259
    const uint32_t MAX_SUBSTEPS = 512;
    double floor_dt = dt / (double)MAX_SUBSTEPS;
    double remaining = dt;
    uint32_t taken = 0;
    while (remaining > 1e-12) {
      double sub = std::fmin(std::fmax(cfl_dt(), floor_dt), remaining);
      substep(sub);
      remaining -= sub;
      taken += 1;
    }
    last_substeps = taken;
  }

  // ---------------------------------------------------------------- reductions for readbacks
  Buf st_cpart, st_ppart, st_out, diag_count;
  // One pass over the particles and one over the cells (metrics.hip); with is_wall set
  // also the metrics' wet-wall marking, with counts the per-cell particle histogram.
  FrameStats frame_stats(const uint8_t* is_wall, uint8_t* washed, uint8_t* wet, double cap, bool counts) {
    if (!st_cpart.p) {
      st_cpart.alloc(kStatBlocks * sizeof(CellStats));
      st_ppart.alloc(kStatBlocks * sizeof(PartStats));
      st_out.alloc(sizeof(FrameStats));
    }
    const uint32_t* cnt = nullptr;
    if (counts) {
      if (!diag_count.p) diag_count.alloc((size_t)ncells() * 4);
      HIPCHK(hipMemsetAsync(diag_count.p, 0, (size_t)ncells() * 4, st));
      if (n > 0) launch_hist<T>(parts(cur), n, grid(), cellkey.as<int32_t>(), diag_count.as<uint32_t>(), st);
      cnt = diag_count.as<uint32_t>();
    }
    launch_frame_stats<T>(parts(cur), n, grid(), cells.as<uint8_t>(), pressure.as<T>(), is_wall, washed, wet, cap,
                          cnt, st_cpart.as<CellStats>(), st_ppart.as<PartStats>(), st_out.as<FrameStats>(), st);
    pin.ensure(sizeof(FrameStats) + 64);
    HIPCHK(hipMemcpyAsync(pin.p, st_out.p, sizeof(FrameStats), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    FrameStats r;
    memcpy(&r, pin.p, sizeof r);
    return r;
  }
  double particle_volume() const {  // solver.rs:225-228
    const double cell = dx * dx * dx;
    const uint64_t p = fluid.particles_per_cell > 1 ? fluid.particles_per_cell : 1;
    return cell / (double)p;
  }
  sdr_diagnostics diagnostics() {  // solver.rs:225-233, 758-891; particles.rs:52-58
    const FrameStats f = frame_stats(nullptr, nullptr, nullptr, 0.0, true);
    sdr_diagnostics d{};
    d.particle_volume = particle_volume();
    d.fluid_volume = (double)n * d.particle_volume;
    d.drained_volume = (double)drained * d.particle_volume;
    d.cells_fluid_count = f.nfluid;
    d.max_particles_per_cell = f.maxcount;
    d.max_pressure = f.pmax;
    d.mean_pressure = f.nfluid > 0 ? f.psum / (double)f.nfluid : 0.0;
    d.max_speed = f.vmax;
    d.kinetic_energy = 0.5 * f.ke;
    d.particles = (uint64_t)n;
    return d;
  }

  // ---------------------------------------------------------------- particle surfacing
  // The particle distance field on a node grid over the grid box, then surface nets.
  // The particles are cell-sorted first if a moving stage ran since the last sort; the
  // next classify re-sorts the same order, so surfacing leaves the run unchanged.
  sdr_surface_info surface(const sdr_surface_params* prm) {
    const uint64_t ppc = fluid.particles_per_cell > 1 ? fluid.particles_per_cell : 1;
    const double r = prm && prm->radius > 0.0 ? prm->radius : dx / std::cbrt((double)ppc);
    const double h = prm && prm->spacing > 0.0 ? prm->spacing : dx;
    if (!(r > 0.0) || !std::isfinite(r) || !(h > 0.0) || !std::isfinite(h))
      throw Error(SDR_E_INVALID, "surface radius and spacing must be positive");
    int nn[3];
    const int ncell[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
      // nodes at origin + i h for i = 0 .. ceil(n dx / h): the grid box and its faces
      const double c = h == dx ? (double)ncell[a] : std::ceil((double)ncell[a] * dx / h);
      if (!(c >= 1.0) || c > 65535.0) throw Error(SDR_E_INVALID, "surface node grid out of range");
      nn[a] = (int)c + 1;
    }
    surf_ng = NodeGrid{nn[0], nn[1], nn[2], origin.x, origin.y, origin.z, h};
    surf_r = r;
    if (!sorted_valid) sort_particles();
    Scope sc_(timer, "surface");
    surf_phi.ensure((size_t)surf_ng.nodes() * sizeof(double));
    const Parts<T> p = parts(cur);
    launch_particle_field<T>(p.px, p.py, p.pz, cstart.as<uint32_t>(), cend.as<uint32_t>(),
                             BinGrid{nx, ny, nz, origin.x, origin.y, origin.z, dx}, surf_ng, r,
                             surf_phi.as<double>(), st);
    nets.run(surf_phi.as<double>(), surf_ng, st);
    have_surface = true;
    return surface_info();
  }
  sdr_surface_info surface_info() const {
    sdr_surface_info o{};
    o.nx = (uint64_t)surf_ng.nx, o.ny = (uint64_t)surf_ng.ny, o.nz = (uint64_t)surf_ng.nz;
    o.origin = sdr_vec3{surf_ng.ox, surf_ng.oy, surf_ng.oz};
    o.spacing = surf_ng.h;
    o.radius = surf_r;
    o.vertices = nets.nverts, o.triangles = nets.ntris;
    o.area = nets.area, o.volume = nets.volume;
    return o;
  }
  void get_surface(double* v, uint64_t vcap, uint32_t* t, uint64_t tcap) {
    if (!have_surface) throw Error(SDR_E_STATE, "no surface built yet");
    download_mesh(nets, v, vcap, t, tcap, st);
  }
  void get_surface_field(double* out, uint64_t cap) {
    if (!have_surface) throw Error(SDR_E_STATE, "no surface built yet");
    const uint64_t m = (uint64_t)surf_ng.nodes();
    if (!out || cap < m) throw Error(SDR_E_INVALID, "output holds fewer values than the node grid");
    HIPCHK(hipMemcpyAsync(out, surf_phi.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  // frame output: positions in reference order, staged into dst on this stream
  void stage_positions(void* dst, bool as_f32, hipEvent_t ev) {
    HIPCHK(hipSetDevice(opts.device));
    if (n > 0) {
      if (as_f32) launch_to_aos_f32<T>(parts(cur), n, (float*)dst, st);
      else launch_to_aos<T>(parts(cur), n, (double*)dst, nullptr, st);
    }
    HIPCHK(hipEventRecord(ev, st));
  }

  // ---------------------------------------------------------------- host I/O
  void get_particles(double* pos, double* vel, uint64_t m) {
    if ((int64_t)m != n) throw Error(SDR_E_INVALID, "particle count mismatch");
    if (n == 0) return;
    aos_pos.ensure((size_t)n * 3 * sizeof(double));
    aos_vel.ensure((size_t)n * 3 * sizeof(double));
    launch_to_aos<T>(parts(cur), n, aos_pos.as<double>(), aos_vel.as<double>(), st);
    if (pos) HIPCHK(hipMemcpyAsync(pos, aos_pos.p, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (vel) HIPCHK(hipMemcpyAsync(vel, aos_vel.p, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  void set_particles(const double* pos, const double* vel, uint64_t m) {
    reserve((int64_t)m);
    // the replaced particles leave the books, the new ones enter them, so that
    // emitted = resident + drained + culled keeps holding
    emitted = emitted - (uint64_t)n + m;
    n = (int64_t)m;
    if (m > 0) {
      aos_pos.ensure(m * 3 * sizeof(double));
      aos_vel.ensure(m * 3 * sizeof(double));
      HIPCHK(hipMemcpyAsync(aos_pos.p, pos, m * 3 * sizeof(double), hipMemcpyHostToDevice, st));
      if (vel) HIPCHK(hipMemcpyAsync(aos_vel.p, vel, m * 3 * sizeof(double), hipMemcpyHostToDevice, st));
      launch_from_aos<T>(parts(cur), n, 0, aos_pos.as<double>(), vel ? aos_vel.as<double>() : nullptr, st);
    }
    HIPCHK(hipStreamSynchronize(st));
    sorted_valid = false;
  }
  void copy_out(const Buf& b, int64_t cnt, double* out) {
    if (!out) return;
    if (sizeof(T) == sizeof(double)) {
      HIPCHK(hipMemcpyAsync(out, b.p, (size_t)cnt * 8, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    } else {
      std::vector<T> tmp((size_t)cnt);
      HIPCHK(hipMemcpyAsync(tmp.data(), b.p, (size_t)cnt * sizeof(T), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      for (int64_t t = 0; t < cnt; ++t) out[t] = (double)tmp[(size_t)t];
    }
  }
  void copy_in(Buf& b, int64_t cnt, const double* in) {
    if (!in) return;
    std::vector<T> tmp((size_t)cnt);
    for (int64_t t = 0; t < cnt; ++t) tmp[(size_t)t] = (T)in[t];
    HIPCHK(hipMemcpyAsync(b.p, tmp.data(), (size_t)cnt * sizeof(T), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  void set_outlet(const sdr_aabb* o) {
    has_outlet = o != nullptr;
    if (o) outlet = *o;
    launch_static_cls<double>(grid64(), mask.as<uint8_t>(), outlet_box64(), static_cls.as<uint8_t>(), st);
    HIPCHK(hipStreamSynchronize(st));
  }
  uint64_t seed_fill(uint32_t per_cell, double y_lo, double y_hi, uint64_t seed, const sdr_vec3* v) {
    std::vector<uint8_t> hm((size_t)ncells());
    HIPCHK(hipMemcpyAsync(hm.data(), mask.p, hm.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<uint32_t> list;
    for (int64_t c = 0; c < ncells(); ++c) {
      int j = (int)((c / nx) % ny);
      double cy = origin.y + ((double)j + 0.5) * dx;
      if (hm[(size_t)c] && cy >= y_lo && cy <= y_hi) list.push_back((uint32_t)c);
    }
    int64_t add = (int64_t)list.size() * per_cell;
    reserve(n + add);
    fill_cells.ensure(list.size() * 4 + 4);
    HIPCHK(hipMemcpyAsync(fill_cells.p, list.data(), list.size() * 4, hipMemcpyHostToDevice, st));
    launch_seed_fill<T>(parts(cur), n, grid(), mask.as<uint8_t>(), fill_cells.as<uint32_t>(),
                        (int64_t)list.size(), per_cell, seed, v ? (T)v->x : T(0), v ? (T)v->y : T(0),
                        v ? (T)v->z : T(0), st);
    HIPCHK(hipStreamSynchronize(st));
    n += add;
    emitted += (uint64_t)add;
    sorted_valid = false;
    return (uint64_t)add;
  }
};

// ------------------------------------------------------------------ clinical metrics
// `MetricsCollector` (metrics.rs:116-446) over a solver's device state: wall cells are
// marked once, each record marks the wet wall cells and reduces the frame on the device.
struct MetricsD {
  const void* owner = nullptr;  // the solver this collector reads
  double cavity_volume = 0;
  int nx = 0, ny = 0, nz = 0;
  double dx = 0;
  hipStream_t st = nullptr;
  Buf is_wall, washed, wet, wc;
  uint64_t wall_count = 0;
  std::vector<sdr_frame_metrics> history;
  double injected_ml = 0;

  template <class T>
  void create(Solver<T>& s, double cv) {  // metrics.rs:119-182
    HIPCHK(hipSetDevice(s.opts.device));
    owner = &s;
    cavity_volume = cv;
    nx = s.nx, ny = s.ny, nz = s.nz, dx = s.dx, st = s.st;
    const size_t nc = (size_t)s.ncells();
    is_wall.alloc(nc), washed.alloc(nc), wet.alloc(nc), wc.alloc(8);
    HIPCHK(hipMemsetAsync(washed.p, 0, nc, st));
    HIPCHK(hipMemsetAsync(wet.p, 0, nc, st));
    HIPCHK(hipMemsetAsync(wc.p, 0, 8, st));
    // the collector's "inside" test is the solver's cavity mask (both sample the SDF at
    // the cell centre, metrics.rs:125-134 and solver.rs:194-199)
    launch_wall_cells(s.mask.template as<uint8_t>(), nx, ny, nz, is_wall.as<uint8_t>(),
                      wc.as<unsigned long long>(), st);
    HIPCHK(hipMemcpyAsync(&wall_count, wc.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }

  template <class T>
  sdr_frame_metrics record(Solver<T>& s, uint64_t frame) {  // metrics.rs:188-316
    if (&s != owner) throw Error(SDR_E_INVALID, "metrics collector used with another solver");
    const double v_jet = [&] {  // solver.rs:107-114
      const double area = M_PI * s.needle.radius * s.needle.radius;
      return area > 0.0 ? s.needle.flow_rate / area : 0.0;
    }();
    const double stagnation = 0.5 * s.fluid.density * v_jet * v_jet;
    const double hydrostatic = s.fluid.density * hlen(h3(s.fluid.gravity)) * ((double)ny * dx);
    const double p_cap = 3.0 * stagnation + hydrostatic + 500.0;  // metrics.rs:236, 454
    const FrameStats f = s.frame_stats(is_wall.as<uint8_t>(), washed.as<uint8_t>(), wet.as<uint8_t>(), p_cap, false);
    const double pv = s.particle_volume();
    const double fluid_volume = (double)s.n * pv, drained_volume = (double)s.drained * pv;
    const double m3_to_ml = 1.0e6, pa_to_mmhg = 1.0 / 133.322;  // metrics.rs:23-24
    injected_ml = (fluid_volume + drained_volume) * m3_to_ml;
    sdr_frame_metrics m{};
    m.frame = frame;
    m.time_s = s.time;
    m.particles = (uint64_t)s.n;
    m.fluid_volume_ml = fluid_volume * m3_to_ml;
    m.drained_volume_ml = drained_volume * m3_to_ml;
    m.fill_fraction = cavity_volume > 0.0 ? fluid_volume / cavity_volume : 0.0;
    m.wall_coverage = wall_count > 0 ? (double)f.nwashed / (double)wall_count : 0.0;
    m.mean_pressure_pa = f.nfluid > 0 ? f.psum / (double)f.nfluid : 0.0;
    m.max_pressure_pa = f.pmax;
    m.max_pressure_mmhg = f.pmax * pa_to_mmhg;
    // membrane_stats (metrics.rs:460-479): (0, 0, n) when every wet cell is an artifact
    m.wall_pressure_pa = f.nacc > 0 ? f.wsum / (double)f.nacc : 0.0;
    m.peak_wall_pressure_pa = f.nacc > 0 ? f.wpeak : 0.0;
    m.peak_wall_pressure_mmhg = m.peak_wall_pressure_pa * pa_to_mmhg;
    m.membrane_artifacts = f.nart;
    m.max_speed_m_s = f.vmax;
    m.kinetic_energy = 0.5 * f.ke;
    m.pcg_iters = s.last_pcg_iters;
    m.wet_wall_cells = f.nwet;
    m.emitted = s.emitted, m.drained = s.drained, m.culled = s.culled;
    history.push_back(m);
    return m;
  }

  sdr_summary summary() const {  // metrics.rs:324-382
    sdr_summary r{};
    double peak_fill = 0, peak_cov = 0, peak_p = 0, peak_mem = 0, mean_mem_sum = 0;
    for (const auto& m : history) {
      peak_fill = std::fmax(peak_fill, m.fill_fraction);
      peak_cov = std::fmax(peak_cov, m.wall_coverage);
      peak_p = std::fmax(peak_p, m.max_pressure_pa);
      peak_mem = std::fmax(peak_mem, m.peak_wall_pressure_pa);
      mean_mem_sum = mean_mem_sum + m.wall_pressure_pa;
      r.membrane_artifacts_total += m.membrane_artifacts;
    }
    const double pa_to_mmhg = 1.0 / 133.322;
    r.frames = history.size();
    const double mean_mem = history.empty() ? 0.0 : mean_mem_sum / (double)history.size();
    if (!history.empty()) {
      const auto& l = history.back();
      r.final_fill_fraction = l.fill_fraction, r.final_wall_coverage = l.wall_coverage;
      r.final_time_s = l.time_s;
      r.total_drained_ml = l.drained_volume_ml;
      r.injected_ml = l.fluid_volume_ml + l.drained_volume_ml;
      r.emitted = l.emitted, r.resident = l.particles, r.drained = l.drained, r.culled = l.culled;
    }
    r.peak_fill_fraction = peak_fill, r.peak_wall_coverage = peak_cov;
    r.peak_membrane_pressure_pa = peak_mem, r.peak_membrane_pressure_mmhg = peak_mem * pa_to_mmhg;
    r.mean_membrane_pressure_pa = mean_mem, r.mean_membrane_pressure_mmhg = mean_mem * pa_to_mmhg;
    r.peak_pressure_pa = peak_p, r.peak_pressure_mmhg = peak_p * pa_to_mmhg;
    // conservation from the integer counts (emitted = resident + drained + culled)
    r.mass_balance = r.emitted > 0 ? (double)(r.resident + r.drained + r.culled) / (double)r.emitted : 1.0;
    return r;
  }
};

// ------------------------------------------------------------------ no-alloc working memory
// The working memory of sdr_project_device_noalloc: one fixed array per device, held
// until the process ends.  In host code `scratch` is only the host-side stand-in of the
// symbol; its device address (per device) comes from hipGetSymbolAddress, taken once per
// device with that device current, and is what kernels and fills are given.
static __device__ __align__(256) unsigned char scratch[SDR_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device `dev`.
void* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw Error(SDR_E_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(SDR_E_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// ------------------------------------------------------------------ C ABI plumbing
thread_local std::string g_err;

}  // namespace sdr

using sdr::Error;
using sdr::guard;

struct sdr_solver {
  int precision = SDR_PRECISION_F64;
  std::unique_ptr<sdr::Solver<double>> d;
  std::unique_ptr<sdr::Solver<float>> f;
};

template <class F>
static int with(sdr_solver* s, F&& fn) {
  if (!s) {
    sdr::g_err = "null solver";
    return SDR_E_INVALID;
  }
  return guard([&] {
    if (s->d) fn(*s->d);
    else fn(*s->f);
  });
}
#define WITH(s, ...) with(s, [&](auto& S) { __VA_ARGS__; })

namespace sdr {
// Frame writer support (output.cpp); called inside its guard.
uint64_t solver_particle_total(sdr_solver* s) {
  if (!s) throw Error(SDR_E_INVALID, "null solver");
  return s->d ? (uint64_t)s->d->n : (uint64_t)s->f->n;
}
void solver_stage_positions(sdr_solver* s, void* dst, bool as_f32, hipEvent_t ev, int* device) {
  if (!s) throw Error(SDR_E_INVALID, "null solver");
  if (s->d) s->d->stage_positions(dst, as_f32, ev), *device = s->d->opts.device;
  else s->f->stage_positions(dst, as_f32, ev), *device = s->f->opts.device;
}
}  // namespace sdr

extern "C" {

const char* sdr_last_error(void) { return sdr::g_err.c_str(); }

int sdr_solver_footprint(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t particles, int32_t precision,
                         uint64_t* bytes) {
  return sdr::guard([&] {
    if (!bytes) throw sdr::Error(SDR_E_INVALID, "null argument");
    const double S = precision == SDR_PRECISION_F32 ? 4.0 : 8.0;
    const double nc = (double)nx * ny * nz;
    const double nsh = (double)nx * (nx + ny - 1) * nz;  // sheared solve layout
    // per particle: two 7-array particle buffers, 14 u32 work arrays, sort scratch (the
    // exact mode's merge keys, the long-segment list)
    const double ppart = 2.0 * (6.0 * S + 4.0) + 56.0 + 5.0;
    // per cell: cell ranges, counts, sort cursor and classes, faces (current, saved,
    // scratch), pressure, redistribution tables, the P2G heavy-cell table (one entry per
    // 64 cells); per sheared slot: 7 solve vectors,
    // flags, f64 refinement vectors (f32 path), halo words of the pipelined sweeps
    const double pcell = 16.0 + 3.0 + 7.0 * S + 2.0 + 89.0 + (sdr::kP2gHeavyVals * S + 4.0) / 64.0;
    const double pslot = 7.0 * S + 2.0 + (precision == SDR_PRECISION_F32 ? 24.0 : 0.0) + 12.0;
    const double est = 1.1 * ((double)particles * ppart + nc * pcell + nsh * pslot) + 64.0 * 1048576.0;
    *bytes = (uint64_t)est;
  });
}

int sdr_device_info(int32_t* count, char* arch, size_t arch_len) {
  return guard([&] {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (count) *count = c;
    if (arch && arch_len) {
      arch[0] = 0;
      if (c > 0) {
        hipDeviceProp_t p;
        HIPCHK(hipGetDeviceProperties(&p, 0));
        snprintf(arch, arch_len, "%s", p.gcnArchName);
      }
    }
  });
}

int sdr_solver_new(const sdr_sdf* solid, double sim_dx, const sdr_fluid_params* fluid,
                   const sdr_needle* needle, const sdr_options* opts, sdr_solver** out) {
  return guard([&] {
    if (!solid || !fluid || !needle || !out || !solid->data)
      throw Error(SDR_E_INVALID, "null argument");
    if (solid->nx < 2 || solid->ny < 2 || solid->nz < 2 || !(sim_dx > 0.0))
      throw Error(SDR_E_INVALID, "invalid SDF dims or sim_dx");
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) throw Error(SDR_E_NODEV, "no HIP device");
    std::unique_ptr<sdr_solver> s(new sdr_solver());
    s->precision = opts ? opts->precision : SDR_PRECISION_F64;
    if (s->precision == SDR_PRECISION_F32) {
      s->f.reset(new sdr::Solver<float>());
      s->f->create(solid, sim_dx, fluid, needle, opts);
    } else if (s->precision == SDR_PRECISION_F64) {
      s->d.reset(new sdr::Solver<double>());
      s->d->create(solid, sim_dx, fluid, needle, opts);
    } else {
      throw Error(SDR_E_INVALID, "unknown precision");
    }
    *out = s.release();
  });
}
int sdr_solver_free(sdr_solver* s) {
  delete s;
  return SDR_OK;
}
int sdr_solver_set_outlet(sdr_solver* s, const sdr_aabb* o) { return WITH(s, S.set_outlet(o)); }
int sdr_solver_set_solve_params(sdr_solver* s, const sdr_solve_params* p) {
  return WITH(s, if (!p) throw Error(SDR_E_INVALID, "null params"); S.solve = *p);
}
int sdr_solver_set_needle(sdr_solver* s, const sdr_needle* nd) {
  return WITH(s, if (!nd) throw Error(SDR_E_INVALID, "null needle"); S.needle = *nd);
}
int sdr_solver_step(sdr_solver* s, double dt) { return WITH(s, S.step(dt)); }
int sdr_solver_substep(sdr_solver* s, double dt) { return WITH(s, S.substep(dt)); }
int sdr_solver_run_stage(sdr_solver* s, int32_t stage, double dt) {
  return WITH(s, S.run_stage(stage, dt); S.timer.collect());
}
int sdr_solver_redistribute(sdr_solver* s) { return WITH(s, S.redistribute()); }
int sdr_solver_particle_count(const sdr_solver* s, uint64_t* n) {
  return WITH(const_cast<sdr_solver*>(s), *n = (uint64_t)S.n);
}
int sdr_solver_get_particles(sdr_solver* s, double* pos, double* vel, uint64_t n) {
  return WITH(s, S.get_particles(pos, vel, n));
}
int sdr_solver_set_particles(sdr_solver* s, const double* pos, const double* vel, uint64_t n) {
  return WITH(s, if (n && !pos) throw Error(SDR_E_INVALID, "null positions"); S.set_particles(pos, vel, n));
}
int sdr_solver_get_dims(const sdr_solver* s, sdr_dims* d) {
  return WITH(const_cast<sdr_solver*>(s), d->nx = S.nx; d->ny = S.ny; d->nz = S.nz; d->dx = S.dx;
              d->origin = sdr_vec3{S.origin.x, S.origin.y, S.origin.z});
}
int sdr_solver_get_grid(sdr_solver* s, double* u, double* v, double* w) {
  return WITH(s, S.copy_out(S.fu, S.nu(), u); S.copy_out(S.fv, S.nv(), v); S.copy_out(S.fw, S.nw(), w));
}
int sdr_solver_set_grid(sdr_solver* s, const double* u, const double* v, const double* w) {
  return WITH(s, S.copy_in(S.fu, S.nu(), u); S.copy_in(S.fv, S.nv(), v); S.copy_in(S.fw, S.nw(), w));
}
int sdr_solver_get_saved(sdr_solver* s, double* u, double* v, double* w) {
  return WITH(s, if (!S.have_saved) throw Error(SDR_E_STATE, "no saved grid yet");
              S.copy_out(S.su, S.nu(), u); S.copy_out(S.sv, S.nv(), v); S.copy_out(S.sw, S.nw(), w));
}
int sdr_solver_set_saved(sdr_solver* s, const double* u, const double* v, const double* w) {
  return WITH(s, S.copy_in(S.su, S.nu(), u); S.copy_in(S.sv, S.nv(), v); S.copy_in(S.sw, S.nw(), w);
              S.have_saved = true);
}
int sdr_solver_get_cells(sdr_solver* s, uint8_t* c) {
  return WITH(s, HIPCHK(hipMemcpyAsync(c, S.cells.p, (size_t)S.ncells(), hipMemcpyDeviceToHost, S.st));
              HIPCHK(hipStreamSynchronize(S.st)));
}
int sdr_solver_set_cells(sdr_solver* s, const uint8_t* c) {
  return WITH(s, HIPCHK(hipMemcpyAsync(S.cells.p, c, (size_t)S.ncells(), hipMemcpyHostToDevice, S.st));
              HIPCHK(hipStreamSynchronize(S.st)));
}
int sdr_solver_get_pressure(sdr_solver* s, double* p) {
  return WITH(s, S.copy_out(S.pressure, S.ncells(), p));
}
int sdr_solver_get_slot_ranks(sdr_solver* s, uint32_t* r, uint64_t m) {
  return WITH(s, if ((int64_t)m != S.n) throw Error(SDR_E_INVALID, "particle count mismatch");
              if (m == 0) return;
              HIPCHK(hipMemcpyAsync(r, S.part[S.cur][6].p, (size_t)m * 4, hipMemcpyDeviceToHost, S.st));
              HIPCHK(hipStreamSynchronize(S.st)));
}
int sdr_solver_get_state(const sdr_solver* s, sdr_state* st) {
  return WITH(const_cast<sdr_solver*>(s), memset(st, 0, sizeof *st); st->time = S.time;
              st->steps = S.steps; st->last_pcg_iters = S.last_pcg_iters;
              st->last_substeps = S.last_substeps; st->drained = S.drained; st->rng = S.rng;
              st->emit_accum = S.emit_accum; st->emitted = S.emitted; st->culled = S.culled;
              st->last_residual = S.last_residual; st->redist_active = S.redist_active;
              st->redist_rounds = (uint32_t)S.redist_rounds;
              st->redist_moved = S.redist_moved);
}
int sdr_solver_set_state(sdr_solver* s, const sdr_state* st) {
  return WITH(s, S.time = st->time; S.steps = st->steps; S.last_pcg_iters = st->last_pcg_iters;
              S.last_substeps = st->last_substeps; S.drained = st->drained; S.rng = st->rng;
              S.emit_accum = st->emit_accum; S.emitted = st->emitted; S.culled = st->culled;
              S.last_residual = st->last_residual);
}
int sdr_solver_synchronize(sdr_solver* s) { return WITH(s, HIPCHK(hipStreamSynchronize(S.st))); }
int sdr_solver_diagnostics(sdr_solver* s, sdr_diagnostics* out) {
  return WITH(s, if (!out) throw Error(SDR_E_INVALID, "null output"); *out = S.diagnostics());
}
int sdr_solver_cfl_dt(sdr_solver* s, double* dt) {
  return WITH(s, if (!dt) throw Error(SDR_E_INVALID, "null output"); *dt = S.cfl_dt());
}

int sdr_solver_surface(sdr_solver* s, const sdr_surface_params* params, sdr_surface_info* out) {
  return WITH(s, if (!out) throw Error(SDR_E_INVALID, "null output"); *out = S.surface(params));
}
int sdr_solver_get_surface(sdr_solver* s, double* v, uint64_t vcap, uint32_t* t, uint64_t tcap) {
  return WITH(s, S.get_surface(v, vcap, t, tcap));
}
int sdr_solver_get_surface_field(sdr_solver* s, double* values, uint64_t cap) {
  return WITH(s, S.get_surface_field(values, cap));
}

int sdr_surface_nets(const double* values, uint64_t nx, uint64_t ny, uint64_t nz, sdr_vec3 origin, double dx,
                     int32_t device, sdr_surface_info* info, double* vertices, uint64_t vcap, uint32_t* tris,
                     uint64_t tcap) {
  return guard([&] {
    if (!values || !info) throw Error(SDR_E_INVALID, "null argument");
    if (nx < 2 || ny < 2 || nz < 2) throw Error(SDR_E_INVALID, "surface_nets needs at least 2 nodes per axis");
    if (nx > 65536 || ny > 65536 || nz > 65536) throw Error(SDR_E_INVALID, "node grid too large");
    if (!(dx > 0.0)) throw Error(SDR_E_INVALID, "node spacing must be positive");
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) throw Error(SDR_E_NODEV, "no HIP device");
    HIPCHK(hipSetDevice(device));
    const sdr::NodeGrid ng{(int)nx, (int)ny, (int)nz, origin.x, origin.y, origin.z, dx};
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    sdr::Buf phi;
    phi.alloc((size_t)ng.nodes() * sizeof(double));
    HIPCHK(hipMemcpyAsync(phi.p, values, (size_t)ng.nodes() * sizeof(double), hipMemcpyHostToDevice, st));
    sdr::SurfaceNets nets;
    nets.run(phi.as<double>(), ng, st);
    sdr_surface_info o{};
    o.nx = nx, o.ny = ny, o.nz = nz, o.origin = origin, o.spacing = dx, o.radius = 0.0;
    o.vertices = nets.nverts, o.triangles = nets.ntris, o.area = nets.area, o.volume = nets.volume;
    *info = o;
    sdr::download_mesh(nets, vertices, vcap, tris, tcap, st);
  });
}

struct sdr_metrics {
  sdr::MetricsD m;
};
int sdr_metrics_new(sdr_solver* s, double cavity_volume, sdr_metrics** out) {
  return WITH(s, if (!out) throw Error(SDR_E_INVALID, "null output");
              std::unique_ptr<sdr_metrics> m(new sdr_metrics()); m->m.create(S, cavity_volume);
              *out = m.release());
}
int sdr_metrics_free(sdr_metrics* m) {
  delete m;
  return SDR_OK;
}
int sdr_metrics_wall_count(const sdr_metrics* m, uint64_t* n) {
  return guard([&] {
    if (!m || !n) throw Error(SDR_E_INVALID, "null argument");
    *n = m->m.wall_count;
  });
}
int sdr_metrics_record(sdr_metrics* m, sdr_solver* s, uint64_t frame, sdr_frame_metrics* out) {
  if (!m) {
    sdr::g_err = "null metrics collector";
    return SDR_E_INVALID;
  }
  return WITH(s, sdr_frame_metrics r = m->m.record(S, frame); if (out) *out = r);
}
int sdr_metrics_summary(const sdr_metrics* m, sdr_summary* out) {
  return guard([&] {
    if (!m || !out) throw Error(SDR_E_INVALID, "null argument");
    *out = m->m.summary();
  });
}
int sdr_metrics_get_cells(const sdr_metrics* m, uint8_t* washed, uint8_t* is_wall) {
  return guard([&] {
    if (!m) throw Error(SDR_E_INVALID, "null argument");
    const size_t nc = (size_t)m->m.nx * m->m.ny * m->m.nz;
    if (washed) HIPCHK(hipMemcpyAsync(washed, m->m.washed.p, nc, hipMemcpyDeviceToHost, m->m.st));
    if (is_wall) HIPCHK(hipMemcpyAsync(is_wall, m->m.is_wall.p, nc, hipMemcpyDeviceToHost, m->m.st));
    HIPCHK(hipStreamSynchronize(m->m.st));
  });
}

int sdr_solver_enable_timing(sdr_solver* s, int32_t on) {
  return WITH(s, S.timer.collect(); S.timer.on = on != 0; S.timer.sweeps_only = on == 2;
              if (on) S.timer.acc.clear(), S.timer.seen.clear());
}
int sdr_solver_get_timing(sdr_solver* s, char* names, size_t name_len, double* ms, uint64_t* calls,
                          uint64_t cap, uint64_t* count) {
  return WITH(s, S.timer.collect(); uint64_t k = 0; for (auto& e : S.timer.acc) {
    if (k >= cap) break;
    if (names) snprintf(names + k * name_len, name_len, "%s", e.first.c_str());
    if (ms) ms[k] = e.second.first;
    if (calls) calls[k] = e.second.second;
    ++k;
  } if (count) *count = k);
}
int sdr_solver_seed_fill(sdr_solver* s, uint32_t per_cell, double y_lo, double y_hi,
                         uint64_t seed, const sdr_vec3* vel, uint64_t* n_out) {
  return WITH(s, uint64_t a = S.seed_fill(per_cell, y_lo, y_hi, seed, vel); if (n_out) *n_out = a);
}

int sdr_project_capturing_with_source(const sdr_dims* gd, double* u, double* v, double* w,
                                      const uint8_t* cells, double rho, double dt,
                                      const sdr_solve_params* params, const double* source,
                                      double* pressure_out, sdr_solve_report* report,
                                      const sdr_options* opts) {
  return guard([&] {
    if (!gd || !u || !v || !w || !cells) throw Error(SDR_E_INVALID, "null argument");
    if (gd->nx < 2 || gd->ny < 2 || gd->nz < 2) throw Error(SDR_E_INVALID, "grid needs >= 2 cells per axis");
    sdr_options o = opts ? *opts : sdr_options{SDR_PRECISION_F64, SDR_DOT_EXACT, 0, 1e-8, SDR_SOLVE_REFERENCE, 0};
    if (o.precision != SDR_PRECISION_F64) throw Error(SDR_E_INVALID, "standalone projection is f64");
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) throw Error(SDR_E_NODEV, "no HIP device");
    HIPCHK(hipSetDevice(o.device));
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    sdr::Timer timer;
    timer.st = st;
    const int nx = (int)gd->nx, ny = (int)gd->ny, nz = (int)gd->nz;
    const int64_t nc = (int64_t)nx * ny * nz, nu = (int64_t)(nx + 1) * ny * nz,
                  nv = (int64_t)nx * (ny + 1) * nz, nw = (int64_t)nx * ny * (nz + 1);
    sdr::Projector<double> pj;
    pj.exact_dot = o.dot_mode == SDR_DOT_EXACT;
    pj.init(nx, ny, nz, gd->dx, st, &timer);
    sdr::Buf du, dv, dw, dc, ds, dp;
    du.alloc(nu * 8), dv.alloc(nv * 8), dw.alloc(nw * 8), dc.alloc(nc), dp.alloc(nc * 8);
    HIPCHK(hipMemcpyAsync(du.p, u, nu * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dv.p, v, nv * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dw.p, w, nw * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dc.p, cells, nc, hipMemcpyHostToDevice, st));
    if (source) {
      ds.alloc(nc * 8);
      HIPCHK(hipMemcpyAsync(ds.p, source, nc * 8, hipMemcpyHostToDevice, st));
    }
    sdr::GridD<double> g{nx, ny, nz, gd->dx, gd->origin.x, gd->origin.y, gd->origin.z};
    sdr_solve_params prm = params ? *params : sdr_solve_params{200, 1e-6};
    int nanf = 0;
    sdr_solve_report r = pj.project(g, sdr::FacesD<double>{du.as<double>(), dv.as<double>(), dw.as<double>()},
                                    dc.as<uint8_t>(), nullptr, source ? ds.as<double>() : nullptr, rho, dt,
                                    prm.max_iters, prm.tol, false, 1.0, 0.0, dp.as<double>(), &nanf,
                                    o.solve_mode == SDR_SOLVE_RELATIVE, o.refine_tol);
    if (nanf) throw Error(SDR_E_NUMERIC, "pressure solve: NaN in the residual (non-finite faces or source?)");
    HIPCHK(hipMemcpyAsync(u, du.p, nu * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(v, dv.p, nv * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(w, dw.p, nw * 8, hipMemcpyDeviceToHost, st));
    if (pressure_out) HIPCHK(hipMemcpyAsync(pressure_out, dp.p, nc * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (report) *report = r;
  });
}

int sdr_project_device_noalloc(const sdr_dims* gd, double* u, double* v, double* w, const uint8_t* cells,
                               double rho, double dt, const sdr_solve_params* params, const double* source,
                               double* pressure_out, sdr_solve_report* report, const sdr_options* opts,
                               void* stream) {
  return guard([&] {
    if (!gd || !u || !v || !w || !cells) throw Error(SDR_E_INVALID, "null argument");
    if (gd->nx < 2 || gd->ny < 2 || gd->nz < 2) throw Error(SDR_E_INVALID, "grid needs >= 2 cells per axis");
    if (gd->nx > (1u << 20) || gd->ny > (1u << 20) || gd->nz > (1u << 20))
      throw Error(SDR_E_INVALID, "grid too large for the no-alloc projection");
    sdr_options o = opts ? *opts : sdr_options{SDR_PRECISION_F64, SDR_DOT_EXACT, 0, 1e-8, SDR_SOLVE_REFERENCE, 0};
    if (o.precision != SDR_PRECISION_F64) throw Error(SDR_E_INVALID, "standalone projection is f64");
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) throw Error(SDR_E_NODEV, "no HIP device");
    HIPCHK(hipSetDevice(o.device));
    uint8_t* base = (uint8_t*)sdr::scratch_on(o.device);
    hipStream_t st = (hipStream_t)stream;
    sdr::Timer timer;  // off: no events
    timer.st = st;
    const int nx = (int)gd->nx, ny = (int)gd->ny, nz = (int)gd->nz;
    sdr::Projector<double> pj;
    pj.exact_dot = o.dot_mode == SDR_DOT_EXACT;
    pj.init_noalloc(nx, ny, nz, gd->dx, st, &timer, base, SDR_NOALLOC_SCRATCH_BYTES);
    sdr::GridD<double> g{nx, ny, nz, gd->dx, gd->origin.x, gd->origin.y, gd->origin.z};
    sdr_solve_params prm = params ? *params : sdr_solve_params{200, 1e-6};
    int nanf = 0;
    sdr_solve_report r = pj.project(g, sdr::FacesD<double>{u, v, w}, cells, nullptr, source, rho, dt,
                                    prm.max_iters, prm.tol, false, 1.0, 0.0, pressure_out, &nanf,
                                    o.solve_mode == SDR_SOLVE_RELATIVE, o.refine_tol);
    HIPCHK(hipStreamSynchronize(st));
    if (nanf) throw Error(SDR_E_NUMERIC, "pressure solve: NaN in the residual (non-finite faces or source?)");
    if (report) *report = r;
  });
}

int sdr_sdf_from_mesh(const double* verts, uint64_t nv, const uint32_t* tris, uint64_t nt, double dx,
                      const sdr_aabb* bounds, double padding, int32_t device, sdr_sdf* out, double* data,
                      uint64_t cap) {
  return guard([&] {
    if (!verts || !tris || !out || !(dx > 0.0)) throw Error(SDR_E_INVALID, "invalid argument");
    for (uint64_t t = 0; t < 3 * nt; ++t)
      if (tris[t] >= nv) throw Error(SDR_E_INVALID, "triangle index out of range");
    double lo[3], hi[3];
    if (bounds) {
      lo[0] = bounds->min.x, lo[1] = bounds->min.y, lo[2] = bounds->min.z;
      hi[0] = bounds->max.x, hi[1] = bounds->max.y, hi[2] = bounds->max.z;
    } else {  // mesh.rs:47-49 (Aabb::from_points), then padded (math.rs:241-246)
      for (int a = 0; a < 3; ++a) lo[a] = INFINITY, hi[a] = -INFINITY;
      for (uint64_t q = 0; q < nv; ++q)
        for (int a = 0; a < 3; ++a) lo[a] = std::fmin(lo[a], verts[3 * q + a]), hi[a] = std::fmax(hi[a], verts[3 * q + a]);
      for (int a = 0; a < 3; ++a) lo[a] = lo[a] - padding, hi[a] = hi[a] + padding;
    }
    uint64_t dims[3];
    for (int a = 0; a < 3; ++a) {  // (size / dx).ceil() as usize + 1 (sdf.rs:54-57)
      const double c = std::ceil((hi[a] - lo[a]) / dx);
      dims[a] = (c > 0.0 ? (uint64_t)c : 0) + 1;
      if (dims[a] > (1u << 16)) throw Error(SDR_E_INVALID, "node grid too large");
    }
    out->origin = sdr_vec3{lo[0], lo[1], lo[2]};
    out->dx = dx;
    out->nx = dims[0], out->ny = dims[1], out->nz = dims[2];
    out->data = data;
    if (!data) return;
    const uint64_t nn = dims[0] * dims[1] * dims[2];
    if (cap < nn) throw Error(SDR_E_INVALID, "output holds fewer values than the node grid");
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) throw Error(SDR_E_NODEV, "no HIP device");
    HIPCHK(hipSetDevice(device));
    std::vector<sdr::MeshTri> tv((size_t)nt);
    for (uint64_t t = 0; t < nt; ++t) {
      const double* a = verts + 3 * (uint64_t)tris[3 * t];
      const double* b = verts + 3 * (uint64_t)tris[3 * t + 1];
      const double* c = verts + 3 * (uint64_t)tris[3 * t + 2];
      tv[(size_t)t] = sdr::MeshTri{a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
    }
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    sdr::Buf dt, dout;
    dt.alloc(tv.size() * sizeof(sdr::MeshTri));
    dout.alloc(nn * sizeof(double));
    if (nt) HIPCHK(hipMemcpyAsync(dt.p, tv.data(), tv.size() * sizeof(sdr::MeshTri), hipMemcpyHostToDevice, st));
    sdr::launch_mesh_sdf(dt.as<sdr::MeshTri>(), (int64_t)nt, (int)dims[0], (int)dims[1], (int)dims[2], lo[0], lo[1],
                         lo[2], dx, dout.as<double>(), st);
    HIPCHK(hipMemcpyAsync(data, dout.p, nn * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  });
}

}  // extern "C"
