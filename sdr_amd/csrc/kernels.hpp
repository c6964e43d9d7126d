// kernels.hpp — host-side launch interface of the substep kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace sdr {

// Particle storage: structure of arrays, kept sorted by cell between the sort in
// `classify` and the first kernel that moves particles.  `rank` is the particle's
// index in the reference's `ParticleSet` order (particles.rs:13-16).
template <class T>
struct Parts {
  T *px, *py, *pz, *vx, *vy, *vz;
  uint32_t* rank;
};

// Scalars of the PCG loop (pressure.rs:341-397), resident on the device.
struct PcgScalars {
  double sigma;
  double denom;
  double alpha;
  double beta;
  double residual;
  double rhs_norm;
  int32_t iters;
  int32_t done;
  int32_t nan_seen;
  int32_t pad;
  double tol_eff;  // the residual bound in use: tol, or tol * max|b| in relative mode
  // The fused product's two direction buffers (launch_spmv_update): spar names the one
  // holding the current s; started is 0 until the first step, whose product takes s = z.
  int32_t spar;
  int32_t started;
  int32_t pad2;
  int32_t pad3;
  // The fast mode's step sigma, set by the fused product (which combines the backward
  // sweep's partials of dot(z, r) itself) and committed to sigma by the axpy (which
  // combines the product's partials of dot(s, As) itself): no finish launches between.
  double sigma_next;
};

enum DotOp : int32_t {
  kOpSigma0 = 0,    // sigma = dot(z, r)
  kOpDenom = 1,     // denom = dot(s, As); break if |denom| < 1e-30; alpha = sigma/denom
  kOpSigmaNew = 2,  // beta = dot(z, r)/sigma; sigma = dot(z, r)
  kOpResidual = 3,  // residual = max|r|; done if residual <= tol   (max partials)
  kOpRhsNorm = 4,   // residual = max|b|; done (0 iterations) if <= tol
};

enum MicOp : int32_t { kMicBuild = 0, kMicFwd = 1, kMicBwd = 2 };

struct SolveCfg {
  double tol;
  double rho, dt, dx;
};

// ---- launches (all asynchronous on `st`)
template <class T>
void launch_gather_parts(const Parts<T>& src, const Parts<T>& dst, const uint32_t* perm, int64_t n,
                         hipStream_t st);
template <class T>
void launch_classify(const uint8_t* static_cls, const Parts<T>& p, int64_t n, GridD<T> g, BoxD<T> outlet,
                     uint8_t* cells, uint32_t* count, hipStream_t st);
template <class T>
void launch_p2g(int comp, GridD<T> g, const Parts<T>& p, const uint32_t* cstart,
                const uint32_t* cend, T* out, bool exact, hipStream_t st);
// Heavy cells of the fused P2G: a cell holding at least `thresh` particles is reduced
// once, by a workgroup of its own (per target face a fixed-order sum over the cell's
// particles), and enters each of its faces' sums as one term after the other particles
// of its slab, so no face walks it particle by particle.  thresh 0: none.
struct P2gHeavy {
  uint32_t* hid;         // per cell: its entry (read only for cells at or over thresh)
  uint32_t* hcell;       // [cap] the cell of each entry
  uint32_t* ctr;         // entries made (may pass cap: the cells past it stay ordinary)
  void* part;            // [cap][kP2gHeavyVals] of T: per component 18 sums of w v, then 18 of w
  uint32_t cap, thresh;
};
constexpr int kP2gHeavyVals = 108;
// List the heavy cells and reduce each (before launch_p2g_fused with the same tables).
template <class T>
void launch_p2g_heavy(GridD<T> g, const Parts<T>& p, const uint32_t* cstart, const uint32_t* cend,
                      P2gHeavy h, hipStream_t st);
// All three P2G components as one LDS-staged gather over the sorted particles
// (p2g.hip), the same sums in the same order as the face-by-face gather (fast mode),
// except that heavy cells (h.thresh > 0) enter as one pre-reduced term each; false when
// the grid is too large for it.
template <class T>
bool launch_p2g_fused(GridD<T> g, const Parts<T>& p, const uint32_t* cstart, const uint32_t* cend, FacesD<T> f,
                      P2gHeavy h, hipStream_t st);
// SDR_P2G_PERFACE=1: the face-by-face gather kernels instead of the fused one.
template <class T>
void launch_axpy_const(T* a, int64_t n, T c, hipStream_t st);
template <class T>
void launch_enforce_solid(GridD<T> g, FacesD<T> f, const uint8_t* cells, hipStream_t st);
template <class T>
void launch_rhs_flags(GridD<T> g, Shear sh, CFacesD<T> f, const uint8_t* cells,
                      const uint32_t* count, const double* srcarr, double inv_scale, int use_source,
                      double rest, double k_vol, double unit, double cap, T* b, uint16_t* flags,
                      double* b64, hipStream_t st);
// As = A x over the in-grid slots; with partials (cell_blocks(sh) of them) it also
// leaves the per-block sums of x . As for the fast reduction mode.
template <class T>
void launch_spmv(Shear sh, const uint16_t* flags, const T* x, T* out, const PcgScalars* sc,
                 double* partials, hipStream_t st);
int cell_blocks(Shear sh);
template <class T>
void launch_dot_partials(Shear sh, const T* a, const T* b, double* partials, int nblocks, bool exact,
                         const PcgScalars* sc, hipStream_t st);
int dot_blocks(Shear sh, bool exact);
// rel = 1: kOpRhsNorm sets the residual bound to tol * max|b| (relative stop)
void launch_dot_finish(double* partials, int nblocks, bool exact, int op, PcgScalars* sc,
                       double tol, hipStream_t st, int rel = 0);
template <class T>
void launch_maxabs_partials(Shear sh, const T* a, double* partials, int nblocks,
                            const PcgScalars* sc, hipStream_t st);
int maxabs_blocks(Shear sh);
template <class T>
void launch_axpy_pr(Shear sh, T* p, T* r, const T* s, const T* as, const PcgScalars* sc,
                    double* partials, int nblocks, hipStream_t st);
template <class T>
void launch_update_s(Shear sh, T* s, const T* z, const PcgScalars* sc, hipStream_t st);
// The direction update s = z + beta s (s = z on a solve's first step) fused with the
// product As = A s and the partials of dot(s, As): s is read from sbuf[spar] and written
// to sbuf[1 - spar] (the residual's finish flips spar).  partials: spmv_update_blocks.
// Each workgroup first combines the backward sweep's ndot partials of dot(z, r) itself
// (kOpSigmaNew: beta = that / sigma, kept in sigma_next; not on a solve's first step).
template <class T>
void launch_spmv_update(Shear sh, const uint16_t* flags, const T* z, T* s0, T* s1, T* out, PcgScalars* sc,
                        double* partials, const double* dot_partials, int ndot, hipStream_t st);
int spmv_update_blocks(Shear sh);
// p += alpha s; r -= alpha As with s the direction just written by launch_spmv_update,
// each workgroup first combining the product's nden partials of dot(s, As) itself
// (kOpDenom: alpha = sigma_next / denom; sigma = sigma_next).  partials: its own, of
// max|r| (kOpResidual).
template <class T>
void launch_axpy_pr_par(Shear sh, T* p, T* r, const T* s0, const T* s1, const T* as, PcgScalars* sc,
                        double* partials, const double* den_partials, int nden, int nblocks, hipStream_t st);
template <class T>
void launch_mic_plane(int op, Shear sh, const uint16_t* flags, T* precon, const T* in, T* out,
                      int sigma, const PcgScalars* sc, hipStream_t st);
template <class T>
void launch_gradient(GridD<T> g, Shear sh, FacesD<T> f, const uint8_t* cells, const T* p, T coef,
                     hipStream_t st);
template <class T>
void launch_unshear(Shear sh, const T* p, T* out, hipStream_t st);
// f64 refinement of the f32 solve
template <class T>
void launch_gradient64(GridD<T> g, Shear sh, FacesD<T> f, const uint8_t* cells, const double* p,
                       double coef, hipStream_t st);
template <class T>
void launch_unshear64(Shear sh, const double* p, T* out, hipStream_t st);
// r = b - A x and *out = max|r| (nblocks partials)
void launch_residual64(Shear sh, const uint16_t* flags, const double* b, const double* x, double* r,
                       double* partials, int nblocks, double* out, hipStream_t st);
// Over the in-grid slots of the sheared layout (the others are left as they are: +0).
template <class T>
void launch_scale_cast(Shear sh, const double* r, double inv, T* b, hipStream_t st);
template <class T>
void launch_accum64(Shear sh, double* p, const T* e, double scale, hipStream_t st);
template <class T>
void launch_scale_inplace(Shear sh, T* x, double f, hipStream_t st);
void launch_pcg_rescale(PcgScalars* sc, double f2, double tol_eff, int restart, hipStream_t st);
template <class T>
void launch_extrap_init(GridD<T> g, int comp, const uint8_t* cells, uint8_t* known, hipStream_t st);
template <class T>
void launch_extrap_iter(int nx, int ny, int nz, const T* src, T* dst, const uint8_t* ksrc,
                        uint8_t* kdst, hipStream_t st);
template <class T>
void launch_g2p(const Parts<T>& p, int64_t n, GridD<T> g, CFacesD<T> f, CFacesD<T> saved, T flip,
                hipStream_t st);
template <class T>
void launch_advect(const Parts<T>& p, int64_t n, GridD<T> g, CFacesD<T> f, SdfD<T> sdf, T dt,
                   hipStream_t st);
// g2p and advect in one pass (the substep's path; the stages stay separate kernels)
template <class T>
void launch_g2p_advect(const Parts<T>& p, int64_t n, GridD<T> g, CFacesD<T> f, CFacesD<T> saved, T flip,
                       SdfD<T> sdf, T dt, hipStream_t st);
template <class T>
void launch_bounds(const Parts<T>& p, int64_t n, SdfD<T> sdf, BoxD<T> outlet, uint32_t* keep,
                   uint32_t* cull_flag, unsigned long long* counters, hipStream_t st);
// Removes the particles with keep = 0 in place (pos = exclusive scan of keep, kept = the
// survivors): the survivors above kept fill the removed slots below it, in slot order;
// the removed particles' ranks go to culled_ranks.  hole_slot: scratch of n entries.
template <class T>
void launch_compact(const Parts<T>& p, const uint32_t* keep, const uint32_t* pos, int64_t n, int64_t kept,
                    uint32_t* hole_slot, uint32_t* culled_ranks, hipStream_t st);
// The reference compacts in place (particles.rs:38-49), so a survivor's rank drops by the
// number of removed ranks below it: bits = a bitmap of the m removed ranks over
// [0, nranks), pop = its words' population counts; with below = the exclusive scan of
// pop, launch_rank_fix subtracts below[r / 32] + the bits under r in its word.
void launch_rank_bits(const uint32_t* culled, int64_t m, int64_t nranks, uint32_t* bits, uint32_t* pop,
                      hipStream_t st);
void launch_rank_fix(uint32_t* rank, int64_t n, const uint32_t* bits, const uint32_t* below, hipStream_t st);
template <class T>
void launch_hist(const Parts<T>& p, int64_t n, GridD<T> g, int32_t* cellkey, uint32_t* count,
                 hipStream_t st);
void launch_active_flags(const int32_t* cellkey, const uint32_t* count, int64_t n, uint32_t cap,
                         uint32_t* flag, hipStream_t st);
void launch_active_compact(const uint32_t* flag, const uint32_t* pos, int64_t n,
                           const uint32_t* rank, uint32_t* act_rank, uint32_t* act_slot,
                           hipStream_t st);
void launch_redist_decide(const uint32_t* act_slot_sorted, int64_t nact, const int32_t* cellkey,
                          uint32_t* count, const uint8_t* mask, int nx, int ny, int nz, uint32_t cap,
                          double gx, double gy, double gz, int32_t* dest, unsigned long long* moved,
                          hipStream_t st);
// dir . gdir of the -x,+x,-y,+y,-z,+z neighbours (solver.rs:492-509)
struct RedistDirs {
  double d[6];
};
// Bucket-based decision pass (redist.hip).  ctr (8 words): [0] sources, [1] bucket
// entries, [2]/[3] pending-list lengths, [4] large buckets, [5] the largest source count
// above what one workgroup sorts (else 0), [6..7] moves (u64).  max_count is ctr[5].
void launch_rd_prepare(const uint32_t* count, const uint32_t* sidx, const uint32_t* boff, int64_t nc,
                       uint32_t cap, uint32_t* src_cell, uint32_t* src_off, uint32_t* src_head,
                       uint32_t* src_lim, uint32_t* src_fill, uint32_t* pend, uint32_t* ctr,
                       hipStream_t st);
void launch_rd_buckets(const int32_t* cellkey, const uint32_t* count, int64_t n, uint32_t cap,
                       const uint32_t* sidx, int64_t nsrc, const uint32_t* src_cell,
                       const uint32_t* src_off, uint32_t* src_fill, const uint32_t* rank,
                       uint32_t* brank, uint32_t* bslot, uint32_t* trank, uint32_t* tslot,
                       uint32_t* big, uint32_t* ctr, uint32_t max_count, hipStream_t st);
// One reservation round; parity selects which pending list is read (0: a, 1: b).
void launch_rd_round2(int64_t max_pending, int parity, bool first, uint32_t* pend_a, uint32_t* pend_b,
                      const uint32_t* src_cell, const uint32_t* src_head, const uint32_t* src_lim,
                      uint32_t* rec_a, uint32_t* rec_b, uint8_t* rmx, const uint32_t* brank, uint32_t* count,
                      const uint8_t* mask,
                      int nx, int ny, int nz, uint32_t cap, RedistDirs dn, unsigned long long* resv,
                      unsigned long long* second, uint32_t epoch, int32_t* dest, uint32_t* ctr,
                      hipStream_t st);
// The moves (solver.rs:546-552).  With count (and cellkey) given, the decision's counts
// become the next sweep's histogram: a moved particle whose projected position lies
// outside its destination cell is counted where it lies, and cellkey follows it.
template <class T>
void launch_redist_apply(const Parts<T>& p, const uint32_t* act_slot_sorted, const int32_t* dest,
                         int64_t nact, GridD<T> g, CFacesD<T> f, SdfD<T> sdf, int32_t* cellkey, uint32_t* count,
                         hipStream_t st);
template <class T>
void launch_maxabs_faces(const T* a, int64_t n, unsigned long long* out_bits, hipStream_t st);
template <class T>
void launch_to_aos(const Parts<T>& p, int64_t n, double* pos, double* vel, hipStream_t st);
template <class T>
void launch_to_aos_f32(const Parts<T>& p, int64_t n, float* pos, hipStream_t st);
template <class T>
void launch_from_aos(const Parts<T>& p, int64_t n, int64_t base, const double* pos,
                     const double* vel, hipStream_t st);
template <class T>
void launch_seed_fill(const Parts<T>& p, int64_t base, GridD<T> g, const uint8_t* mask,
                      const uint32_t* fill_cells, int64_t nfill, uint32_t per_cell, uint64_t seed,
                      T vx, T vy, T vz, hipStream_t st);
template <class T>
void launch_static_cls(GridD<T> g, const uint8_t* mask, BoxD<T> outlet, uint8_t* static_cls,
                       hipStream_t st);
template <class T>
void launch_cavity_mask(GridD<T> g, SdfD<T> sdf, uint8_t* mask, hipStream_t st);

// Zero n 32-bit words with a kernel (no memset node in a captured graph).
void launch_zero_u32(uint32_t* p, int64_t n, hipStream_t st);
// The per-plane work-unit table of the vector kernels on the device (units_per_plane
// entries), the same table Shear::units holds.
int units_per_plane(int nx, int ny);
void launch_build_units(int nx, int ny, uint32_t* out, hipStream_t st);

}  // namespace sdr
