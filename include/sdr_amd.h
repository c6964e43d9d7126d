/*
 * sdr_amd.h — C ABI of the MI355X-native FLIP/PIC substep library (libsdr_amd.so).
 *
 * The boundary mirrors the reference crate's `Solver` / `pressure` surface
 * (uselessgoddess/sdr).  Every entry point below names the Rust item it replaces.
 * A Rust crate binds this header with `extern "C"` blocks;
 * Python binds it with ctypes (sdr_amd/_abi.py).
 *
 * Conventions
 *   - Return value: 0 on success, negative SDR_E_* on failure; the message of the
 *     last failure on the calling thread is available from sdr_last_error().
 *   - The caller owns every host buffer (copied in / out).  The library owns the
 *     device buffers of a solver.  One solver = one HIP stream; a solver is not
 *     thread-safe, distinct solvers may be driven from distinct threads.
 *   - Vectors are passed as `double[3]` triples (AoS, x,y,z), exactly the memory
 *     layout of the reference's `Vec<Vec3>`.
 *   - Cell kinds use the reference enum's discriminant order
 *     (`pressure.rs:24-32`): 0 = Solid, 1 = Fluid, 2 = Air.
 */
#ifndef SDR_AMD_H
#define SDR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDR_OK 0
#define SDR_E_INVALID (-1)  /* bad argument / size mismatch (reference: assert_eq! panics) */
#define SDR_E_OOM (-2)      /* device allocation failed */
#define SDR_E_HIP (-3)      /* a HIP runtime call failed */
#define SDR_E_NODEV (-4)    /* no usable gfx950 device */
#define SDR_E_STATE (-5)    /* call not valid in the current state */
#define SDR_E_NUMERIC (-6)  /* NaN residual or persistent-kernel wait limit hit */
#define SDR_E_IO (-7)       /* a file could not be opened, read or written */

/* Arithmetic of the device path.
 *   SDR_PRECISION_F64: the plain drop-in; every field in f64, reference operand order.
 *   SDR_PRECISION_F32: f32 particle/grid/PCG storage and sweeps, f64 dots, norms and
 *                      mass sums, f64 outer iterative refinement of the pressure solve. */
#define SDR_PRECISION_F64 0
#define SDR_PRECISION_F32 1

/* Dot-product reduction of the f64 path.
 *   SDR_DOT_EXACT: the reference's 4096-element chunk tree (`pressure.rs:588-603`),
 *                  bit-identical to the CPU path.
 *   SDR_DOT_FAST:  wavefront-shuffle tree in f64, deterministic but not the reference tree. */
#define SDR_DOT_EXACT 0
#define SDR_DOT_FAST 1

/* Stopping rule of the pressure solve.
 *   SDR_SOLVE_REFERENCE: the reference's `inf_norm(r) <= tol` with `max_iters`
 *                        (`pressure.rs:341-397`).
 *   SDR_SOLVE_RELATIVE:  `inf_norm(r) <= refine_tol * inf_norm(b)` within `max_iters`
 *                        PCG iterations; with SDR_PRECISION_F32 the residual is formed
 *                        in f64 and the f32 solve is refined until it holds. */
#define SDR_SOLVE_REFERENCE 0
#define SDR_SOLVE_RELATIVE 1

typedef struct sdr_vec3 {
  double x, y, z;
} sdr_vec3;

/* `Aabb` (math.rs:198-201); `contains` is inclusive on both ends. */
typedef struct sdr_aabb {
  sdr_vec3 min, max;
} sdr_aabb;

/* `Sdf` (sdf.rs:23-34): node-grid signed distance, x-fastest, phi < 0 inside the cavity. */
typedef struct sdr_sdf {
  sdr_vec3 origin;
  double dx;
  uint64_t nx, ny, nz;
  const double* data; /* nx*ny*nz values, host memory, copied at solver creation */
} sdr_sdf;

/* `FluidParams` (solver.rs:41-85). */
typedef struct sdr_fluid_params {
  double density;
  sdr_vec3 gravity;
  double flip_ratio;
  uint64_t particles_per_cell;
  double volume_stiffness;
  int32_t redistribute; /* bool */
  double redist_cap_factor;
} sdr_fluid_params;

/* `Needle` (solver.rs:90-104). */
typedef struct sdr_needle {
  sdr_vec3 tip;
  sdr_vec3 axis;
  double radius;
  double flow_rate;
  double inject_until; /* +inf = inject for the whole run */
} sdr_needle;

/* `SolveParams` (pressure.rs:36-50): defaults {200, 1e-6}. */
typedef struct sdr_solve_params {
  uint64_t max_iters;
  double tol;
} sdr_solve_params;

/* `SolveReport` (pressure.rs:54-57). */
typedef struct sdr_solve_report {
  uint64_t iters;
  double residual;
} sdr_solve_report;

/* Device-path options (no reference counterpart: the reference is CPU-only). */
typedef struct sdr_options {
  int32_t precision; /* SDR_PRECISION_* */
  int32_t dot_mode;  /* SDR_DOT_* (f64 path only) */
  int32_t device;    /* HIP device ordinal */
  double refine_tol; /* SDR_SOLVE_RELATIVE: relative residual target (1e-8) */
  int32_t solve_mode; /* SDR_SOLVE_* */
  int32_t _pad;
} sdr_options;

/* Scalar solver state: the reference's pub fields `time, steps, last_pcg_iters,
 * last_substeps, drained` (solver.rs:152-165) plus the private emission state
 * `rng, emit_accum` (solver.rs:177-179) and the particle-conservation counters the
 * build adds (emitted = resident + drained + culled). */
typedef struct sdr_state {
  double time;
  uint64_t steps;
  uint64_t last_pcg_iters;
  uint32_t last_substeps;
  uint32_t redist_rounds; /* reservation rounds of the last substep's redistribution */
  uint64_t drained;
  uint64_t rng;
  double emit_accum;
  uint64_t emitted;
  uint64_t culled;
  double last_residual;
  uint64_t redist_active; /* particles visited by the last substep's redistribution passes */
  uint64_t redist_moved;  /* particles those passes moved */
} sdr_state;

typedef struct sdr_dims {
  uint64_t nx, ny, nz;
  double dx;
  sdr_vec3 origin;
} sdr_dims;

/* Stages of `Solver::substep` (solver.rs:284-310), for stage-isolated runs. */
#define SDR_STAGE_EMIT 0           /* solver.rs:313-339 */
#define SDR_STAGE_CLASSIFY 1       /* solver.rs:342-371 */
#define SDR_STAGE_P2G 2            /* solver.rs:560-585 */
#define SDR_STAGE_SAVE 3           /* solver.rs:587-591 */
#define SDR_STAGE_FORCES 4         /* solver.rs:593-604 */
#define SDR_STAGE_PROJECT 5        /* volume_source + project_capturing_with_source, solver.rs:290-299 */
#define SDR_STAGE_EXTRAPOLATE 6    /* solver.rs:608-647 */
#define SDR_STAGE_G2P 7            /* solver.rs:650-677 */
#define SDR_STAGE_ADVECT 8         /* solver.rs:693-711 */
#define SDR_STAGE_BOUNDS 9         /* solver.rs:736-755 */
#define SDR_STAGE_REDISTRIBUTE 10  /* solver.rs:454-557 (runs only if fluid.redistribute) */
#define SDR_STAGE_FINISH 11        /* time += dt; steps += 1 (solver.rs:308-309) */
#define SDR_STAGE_COUNT 12

typedef struct sdr_solver sdr_solver;

/* Message of the last failure on this thread ("" if none). */
const char* sdr_last_error(void);

/* Library / device query: number of visible devices and gfx name of device 0. */
int sdr_device_info(int32_t* count, char* arch, size_t arch_len);

/* `Solver::new(solid, sim_dx, fluid, needle)` (solver.rs:184-222).  `opts` may be
 * NULL (f64, exact dots, device 0). */
int sdr_solver_new(const sdr_sdf* solid, double sim_dx, const sdr_fluid_params* fluid,
                   const sdr_needle* needle, const sdr_options* opts, sdr_solver** out);
int sdr_solver_free(sdr_solver* s);

/* `solver.outlet = Some(aabb)` / `None` (solver.rs:150). */
int sdr_solver_set_outlet(sdr_solver* s, const sdr_aabb* outlet /* nullable */);
/* `solver.solve = params` (solver.rs:151). */
int sdr_solver_set_solve_params(sdr_solver* s, const sdr_solve_params* params);
/* `solver.needle = needle` (solver.rs:148). */
int sdr_solver_set_needle(sdr_solver* s, const sdr_needle* needle);

/* `Solver::step(dt)` (solver.rs:237-259): CFL sub-stepping of one frame. */
int sdr_solver_step(sdr_solver* s, double dt);
/* `Solver::substep(dt)` (solver.rs:284-310). */
int sdr_solver_substep(sdr_solver* s, double dt);
/* One stage of `Solver::substep` (SDR_STAGE_*), for stage-isolated parity runs. */
int sdr_solver_run_stage(sdr_solver* s, int32_t stage, double dt);
/* `Solver::redistribute()` (solver.rs:454-557), independent of fluid.redistribute. */
int sdr_solver_redistribute(sdr_solver* s);

/* `solver.particles` (particles.rs:13-16): count, read (reference order), replace. */
int sdr_solver_particle_count(const sdr_solver* s, uint64_t* n);
int sdr_solver_get_particles(sdr_solver* s, double* pos_xyz, double* vel_xyz, uint64_t n);
int sdr_solver_set_particles(sdr_solver* s, const double* pos_xyz, const double* vel_xyz,
                             uint64_t n);

int sdr_solver_get_dims(const sdr_solver* s, sdr_dims* out);
/* `solver.grid.{u,v,w}` (grid.rs:92-101), face arrays of (nx+1)ny nz, nx(ny+1)nz, nx ny(nz+1). */
int sdr_solver_get_grid(sdr_solver* s, double* u, double* v, double* w);
int sdr_solver_set_grid(sdr_solver* s, const double* u, const double* v, const double* w);
/* private `saved_{u,v,w}` (solver.rs:174-176): the post-P2G grid FLIP differences against. */
int sdr_solver_get_saved(sdr_solver* s, double* u, double* v, double* w);
int sdr_solver_set_saved(sdr_solver* s, const double* u, const double* v, const double* w);
/* private `cells` (solver.rs:167): 0 Solid, 1 Fluid, 2 Air. */
int sdr_solver_get_cells(sdr_solver* s, uint8_t* cells);
int sdr_solver_set_cells(sdr_solver* s, const uint8_t* cells);
/* `solver.pressure` (solver.rs:163). */
int sdr_solver_get_pressure(sdr_solver* s, double* p);
/* The device's own particle order, for parity tests of its cell sort: the reference
 * index (`rank`) of each particle slot, in slot order (after `classify`: by clamped cell,
 * then reference index in exact mode, previous slot otherwise).  No counterpart in the
 * reference, which keeps its particles in reference order. */
int sdr_solver_get_slot_ranks(sdr_solver* s, uint32_t* ranks, uint64_t n);
int sdr_solver_get_state(const sdr_solver* s, sdr_state* st);
int sdr_solver_set_state(sdr_solver* s, const sdr_state* st);

/* The reference's pub diagnostics of `Solver`, computed on the device in one call
 * (no particle or cell download):
 *   particle_volume         solver.rs:225-228   dx^3 / max(ppc, 1)
 *   fluid_volume            solver.rs:231-233   particles * particle_volume
 *   drained_volume          solver.rs:758-760   drained * particle_volume
 *   cells_fluid_count       solver.rs:763-765   cells classified Fluid
 *   max_particles_per_cell  solver.rs:770-782   largest per-cell count (cell_of, all cells)
 *   max_pressure            solver.rs:859-867   max(0, pressure over Fluid cells)
 *   mean_pressure           solver.rs:869-882   mean pressure over Fluid cells (0 if none)
 *   max_speed               solver.rs:884-890   max particle |v|
 *   kinetic_energy          particles.rs:52-58  0.5 * sum |v|^2
 * Counts and maxima are exact; the two sums are deterministic trees (the reference sums
 * sequentially), equal to it within rounding. */
typedef struct sdr_diagnostics {
  double particle_volume;
  double fluid_volume;
  double drained_volume;
  uint64_t cells_fluid_count;
  uint64_t max_particles_per_cell;
  double max_pressure;
  double mean_pressure;
  double max_speed;
  double kinetic_energy;
  uint64_t particles;
} sdr_diagnostics;
int sdr_solver_diagnostics(sdr_solver* s, sdr_diagnostics* out);

/* The reference's private `Solver::cfl_dt` (solver.rs:262-281, with
 * `MacGrid::max_speed`, grid.rs:202-206) on the current grid: the sub-step `step` would
 * take next before its floor and the frame remainder apply. */
int sdr_solver_cfl_dt(sdr_solver* s, double* dt);

/* `Sdf::from_mesh(mesh, dx, padding)` and `Sdf::from_mesh_in_bounds(mesh, dx, bounds)`
 * (sdf.rs:47-86) with `signed_distance_to_mesh` (sdf.rs:187-272): the signed distance of
 * every node of a grid to a closed triangle mesh, negative inside (generalised winding
 * number > 1/2), built on device `device` (one thread per node over all triangles).
 * vertices: nv xyz triples; tris: nt vertex-index triples (`TriMesh`, mesh.rs:16-45);
 * bounds: NULL = the vertex bounds padded by `padding` (mesh.rs:47-49, math.rs:241-246).
 * Fills out's origin, dx and dims (nx = ceil(size / dx) + 1, ...) and, when data is not
 * NULL, writes the nx*ny*nz node values there (x fastest; cap = values data holds) and
 * sets out->data = data.  Call with data NULL to learn the dims. */
int sdr_sdf_from_mesh(const double* vertices, uint64_t nv, const uint32_t* tris, uint64_t nt, double dx,
                      const sdr_aabb* bounds, double padding, int32_t device, sdr_sdf* out, double* data,
                      uint64_t cap);

/* ---------------------------------------------------------------------------------
 * Frame output (output.rs:23-122, mesh.rs:356-456).  Positions are xyz triples in
 * reference order.  Coordinates in VTK and OBJ text are the shortest decimal that
 * parses back to the same double, in positional notation (Rust's f64 Display), so a
 * frame reads back bit for bit.  File errors return SDR_E_IO. */
#define SDR_FORMAT_VTK 0 /* ASCII legacy VTK, `write_points_vtk` */
#define SDR_FORMAT_XYZ 1 /* raw little-endian f32 triples, `write_points_xyz` */

/* `write_points_vtk(path, positions)` (output.rs:23-47). */
int sdr_write_points_vtk(const char* path, const double* pos_xyz, uint64_t n);
/* `read_points_vtk(path)` (output.rs:56-95): the POINTS block of a legacy VTK file.
 * Sets *n to the point count and, when pos_xyz is not NULL, writes the points there
 * (cap = points pos_xyz holds; fewer than *n is SDR_E_INVALID). */
int sdr_read_points_vtk(const char* path, double* pos_xyz, uint64_t cap, uint64_t* n);
/* `write_points_xyz(path, positions)` (output.rs:99-112). */
int sdr_write_points_xyz(const char* path, const double* pos_xyz, uint64_t n);
/* `frame_path(dir, prefix, frame, ext)` (output.rs:115-122): "dir/prefix_NNNNNN.ext"
 * (frame zero-padded to 6 digits) into buf of len bytes, NUL included. */
int sdr_frame_path(const char* dir, const char* prefix, uint64_t frame, const char* ext, char* buf, size_t len);
/* `TriMesh::save_obj` (mesh.rs:356-379) and `TriMesh::save_stl` (mesh.rs:427-456). */
int sdr_write_mesh_obj(const char* path, const double* vertices, uint64_t nv, const uint32_t* tris, uint64_t nt);
int sdr_write_mesh_stl(const char* path, const double* vertices, uint64_t nv, const uint32_t* tris, uint64_t nt);

/* The reference writes each frame synchronously from host memory (main.rs:255-265).
 * The frame writer instead stages a solver's positions into a device buffer of its own
 * on the solver's stream (one scatter into reference order), copies them to pinned host
 * memory on a copy stream, and writes the file from a host thread: the substeps queued
 * after a submit overlap the copy and the write.  `slots` (1-8) staging buffers bound
 * the frames in flight; submit blocks only while all are busy.  Files are named
 * frame_path(dir, prefix, frame, "vtk" | "xyz"). */
typedef struct sdr_frame_writer sdr_frame_writer;
int sdr_frame_writer_new(const char* dir, const char* prefix, int32_t format, uint32_t slots,
                         sdr_frame_writer** out);
/* Queue frame `frame` of solver s (its particles as of the work queued so far).  An
 * error of an earlier frame's write is returned by the next submit or flush. */
int sdr_frame_writer_submit(sdr_frame_writer* w, sdr_solver* s, uint64_t frame);
/* Wait until every submitted frame is written; *written = frames written so far. */
int sdr_frame_writer_flush(sdr_frame_writer* w, uint64_t* written);
/* Flushes, then frees; a solver submitted to must outlive its pending frames. */
int sdr_frame_writer_free(sdr_frame_writer* w);

/* ---------------------------------------------------------------------------------
 * Particle surfacing.  The reference hands a frame to the external splashsurf tool
 * (recon.rs:60-88); here the surface is built on the device from the resident
 * particles: the particle distance field phi(x) = min(min_j |x - x_j|, 2 r) - r on a
 * node grid (negative inside the union of balls of radius r), then naive surface nets
 * (surface_nets.rs:48-221) over it.  The mesh stays on the device; its counts, area
 * and signed volume come back in sdr_surface_info. */
typedef struct sdr_surface_params {
  double radius;  /* ball radius r (m); 0 = dx / cbrt(particles_per_cell), the seeding spacing */
  double spacing; /* node spacing h (m); 0 = dx (nodes on the MAC cell corners) */
} sdr_surface_params;
typedef struct sdr_surface_info {
  uint64_t nx, ny, nz; /* node grid: node (i,j,k) at origin + (i,j,k) * spacing */
  sdr_vec3 origin;
  double spacing;
  double radius;
  uint64_t vertices;
  uint64_t triangles;
  double area;   /* `TriMesh::surface_area` (mesh.rs:52-60), summed in a fixed tree */
  double volume; /* `TriMesh::signed_volume` (mesh.rs:63-70), summed in a fixed tree */
} sdr_surface_info;
/* Surface the solver's current particles over the grid box [origin, origin + n dx]. */
int sdr_solver_surface(sdr_solver* s, const sdr_surface_params* params, sdr_surface_info* out);
/* The last surface: vertices (xyz, vcap triples) and triangles (tcap index triples);
 * either may be NULL.  Vertices are in cell order, triangles in node order (x, y, z
 * edges per node), as surface_nets.rs numbers them. */
int sdr_solver_get_surface(sdr_solver* s, double* vertices, uint64_t vcap, uint32_t* tris, uint64_t tcap);
/* The last surface's node field phi (nx*ny*nz values, x fastest). */
int sdr_solver_get_surface_field(sdr_solver* s, double* values, uint64_t cap);
/* `surface_nets(origin, dx, dims, field)` (surface_nets.rs:48-195) over sampled node
 * values (nx*ny*nz, x fastest, nodes >= 2 per axis) on device `device`.  Fills info and,
 * for non-NULL buffers, the mesh (vcap / tcap triples held; fewer than needed is
 * SDR_E_INVALID).  Call with NULL buffers to learn the counts. */
int sdr_surface_nets(const double* values, uint64_t nx, uint64_t ny, uint64_t nz, sdr_vec3 origin, double dx,
                     int32_t device, sdr_surface_info* info, double* vertices, uint64_t vcap, uint32_t* tris,
                     uint64_t tcap);

/* ---------------------------------------------------------------------------------
 * Clinical metrics (`MetricsCollector`, metrics.rs:116-446) on the device.  The
 * collector marks wall cells (cavity cells with a non-cavity or out-of-domain face
 * neighbour, metrics.rs:119-182) once; each record marks the wall cells holding
 * particles as washed and reduces the frame's metrics on the device, so a frame costs
 * one pass over the particles and one over the cells and no download.  Sums over wet
 * wall cells and fluid cells run in a fixed tree order (the reference sums wet wall
 * cells sequentially in cell-index order, metrics.rs:189-196): deterministic, equal to
 * the reference within rounding.  mass_balance uses the integer counts
 * emitted = resident + drained + culled instead of the reference's resident+drained
 * over itself (metrics.rs:220,355-362, which is 1 by construction). */
typedef struct sdr_frame_metrics {  /* `FrameMetrics` (metrics.rs:27-67) */
  uint64_t frame;
  double time_s;
  uint64_t particles;
  double fluid_volume_ml;
  double drained_volume_ml;
  double fill_fraction;
  double wall_coverage;
  double mean_pressure_pa;
  double max_pressure_pa;
  double max_pressure_mmhg;
  double wall_pressure_pa;
  double peak_wall_pressure_pa;
  double peak_wall_pressure_mmhg;
  uint64_t membrane_artifacts;
  double max_speed_m_s;
  double kinetic_energy;
  uint64_t pcg_iters;
  uint64_t wet_wall_cells;  /* wall cells holding particles this frame */
  uint64_t emitted, drained, culled;  /* integer conservation counters at this frame */
} sdr_frame_metrics;

typedef struct sdr_summary {  /* `Summary` (metrics.rs:71-107) */
  uint64_t frames;
  double final_time_s;
  double peak_fill_fraction;
  double final_fill_fraction;
  double peak_wall_coverage;
  double final_wall_coverage;
  double peak_membrane_pressure_pa;
  double peak_membrane_pressure_mmhg;
  double mean_membrane_pressure_pa;
  double mean_membrane_pressure_mmhg;
  double peak_pressure_pa;
  double peak_pressure_mmhg;
  uint64_t membrane_artifacts_total;
  double total_drained_ml;
  double injected_ml;
  double mass_balance;  /* (resident + drained + culled) / emitted at the last frame; 1 if none */
  uint64_t emitted, resident, drained, culled;
} sdr_summary;

typedef struct sdr_metrics sdr_metrics;
/* `MetricsCollector::new(solver, cavity_volume)` (metrics.rs:119-182). */
int sdr_metrics_new(sdr_solver* s, double cavity_volume, sdr_metrics** out);
int sdr_metrics_free(sdr_metrics* m);
/* Number of wall cells (`wall_count`, metrics.rs:127). */
int sdr_metrics_wall_count(const sdr_metrics* m, uint64_t* n);
/* `MetricsCollector::record(solver, frame)` (metrics.rs:188-316); s must be the solver
 * the collector was made for. */
int sdr_metrics_record(sdr_metrics* m, sdr_solver* s, uint64_t frame, sdr_frame_metrics* out);
/* `MetricsCollector::summary()` (metrics.rs:324-382) over the recorded frames. */
int sdr_metrics_summary(const sdr_metrics* m, sdr_summary* out);
/* Washed flags per cell (0/1, cell order) and wall flags, for checks. */
int sdr_metrics_get_cells(const sdr_metrics* m, uint8_t* washed, uint8_t* is_wall);

/* Block until every operation queued on the solver's stream has finished. */
int sdr_solver_synchronize(sdr_solver* s);

/* Per-kernel timing of the last substep, filled when enabled (HIP events on the
 * solver's stream): on = 1 times every region, on = 2 only the MIC(0) sweep regions
 * (mic_*), the 16th, 32nd, ... call of each (an event pair delays the next kernel by a
 * few microseconds), 0 stops.  names/ms arrays of length cap; *count
 * receives the number written. */
int sdr_solver_enable_timing(sdr_solver* s, int32_t on);
int sdr_solver_get_timing(sdr_solver* s, char* names, size_t name_len, double* ms,
                          uint64_t* calls, uint64_t cap, uint64_t* count);

/* Append `per_cell` particles to every cavity cell whose centre has y_lo <= y <= y_hi
 * (centre + splitmix64 jitter of (seed, index)), all with velocity `vel` — the synthetic
 * initial fill of the benchmark cavities (no reference counterpart; the reference only
 * emits from the needle).  Runs on the device. */
int sdr_solver_seed_fill(sdr_solver* s, uint32_t per_cell, double y_lo, double y_hi,
                         uint64_t seed, const sdr_vec3* vel, uint64_t* n_out);

/* `pressure::project_capturing_with_source(grid, cells, rho, dt, params, source)`
 * (pressure.rs:123-239) on host buffers.  u,v,w are updated in place; pressure_out
 * (nullable) receives the per-cell pressure; source is nullable (None). */
int sdr_project_capturing_with_source(const sdr_dims* grid, double* u, double* v, double* w,
                                      const uint8_t* cells, double rho, double dt,
                                      const sdr_solve_params* params, const double* source,
                                      double* pressure_out, sdr_solve_report* report,
                                      const sdr_options* opts);

/* Working memory of sdr_project_device_noalloc: a fixed array per device (64 MiB). */
#define SDR_NOALLOC_SCRATCH_BYTES (64u << 20)

/* No counterpart in the reference (which allocates a fresh pressure Vec per call,
 * pressure.rs:123-239): the projection of sdr_project_capturing_with_source, f64, for
 * callers that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  u, v, w (in/out), cells, source
 * (nullable) and pressure_out (nullable) are DEVICE pointers of the same layouts; the
 * work runs on `stream` (a hipStream_t, NULL = the null stream) and the call returns
 * when it has finished.  All working memory (solve vectors, flags, reductions, halo
 * words) comes from one static device array of SDR_NOALLOC_SCRATCH_BYTES per device,
 * cleared by kernels where it must start at zero; nothing is allocated or freed.  A
 * grid whose working set exceeds it (about 80^3 cells) is rejected with SDR_E_INVALID.
 * Every call on a device shares that device's array: calls that overlap on one device
 * (concurrent streams or threads) race and must be serialised by the caller. */
int sdr_project_device_noalloc(const sdr_dims* grid, double* u, double* v, double* w, const uint8_t* cells,
                               double rho, double dt, const sdr_solve_params* params, const double* source,
                               double* pressure_out, sdr_solve_report* report, const sdr_options* opts,
                               void* stream);

/* ---------------------------------------------------------------------------------
 * Serving (no reference counterpart: the reference runs one scenario per process,
 * main.rs:215-386).  Many irrigation scenarios of different sizes share one GPU: each
 * request is admitted when a slot is free and its estimated HBM footprint fits the
 * budget (first come, first served), runs on its own solver and HIP stream from its own
 * host thread, so the work of admitted scenarios overlaps on the device, and is
 * released when it finishes; with resident pools the slot keeps its memory for the
 * next request.  Each scenario's result depends only on its own inputs
 * (no float atomics, no shared state): it is the same alone and in the mix. */
typedef struct sdr_scenario {
  sdr_sdf solid;            /* host SDF, read while the scenario is being set up */
  double sim_dx;
  sdr_fluid_params fluid;
  sdr_needle needle;
  sdr_options opts;
  sdr_solve_params solve;
  int32_t has_outlet;
  sdr_aabb outlet;
  uint32_t fill_per_cell;   /* sdr_solver_seed_fill arguments */
  double fill_y_lo, fill_y_hi;
  uint64_t fill_seed;
  double dt;                /* fixed substep */
  uint32_t substeps;        /* work of the request */
  double arrival_s;         /* arrival time relative to the start of sdr_serve */
} sdr_scenario;

typedef struct sdr_serve_config {
  double hbm_budget_bytes;  /* admission budget over the estimated footprints */
  uint32_t max_slots;       /* scenarios in flight at once */
  uint32_t resident_pools;  /* 1: each slot keeps its device memory in a pool across
                               requests (sdr_pool_*); 0: every request allocates */
} sdr_serve_config;

typedef struct sdr_scenario_result {
  int32_t status;           /* SDR_OK, or the failure (SDR_E_OOM: larger than the budget) */
  uint32_t slot;
  double arrival_s, admit_s, finish_s;  /* seconds from the start of sdr_serve */
  uint64_t footprint_bytes;             /* the admission estimate */
  uint64_t particles;                   /* resident at the end */
  uint64_t particle_substeps;           /* sum over substeps of the particle count */
  uint64_t pcg_iters;                   /* sum over substeps */
  uint64_t checksum;        /* FNV-1a of the final positions and velocities (f64 bits) */
  uint64_t device_allocs;   /* hipMalloc / hipHostMalloc calls of the request (pool misses) */
} sdr_scenario_result;

/* Memory pools (no reference counterpart).  A pool keeps the device and pinned host
 * blocks released by the solvers made from it and hands them to later allocations of a
 * similar size, so a serving slot runs request after request without hipMalloc/hipFree
 * (hipFree waits for the whole device).  sdr_pool_bind(p) makes the calling thread's
 * allocations come from p (NULL: plain allocation again); a buffer returns to the pool
 * it came from, whichever thread frees it.  A pool serves one solver at a time: a
 * solver's blocks return to it when sdr_solver_free has drained the solver's stream.
 * A failed allocation frees the pool's cached blocks and tries once more. */
typedef struct sdr_pool sdr_pool;
typedef struct sdr_pool_stats {
  uint64_t hits;          /* allocations served from cached blocks */
  uint64_t misses;        /* allocations that called hipMalloc / hipHostMalloc */
  uint64_t cached_bytes;  /* held for reuse */
  uint64_t live_bytes;    /* in use by buffers */
  uint64_t cached_blocks;
  uint64_t live_blocks;
} sdr_pool_stats;
int sdr_pool_new(int32_t device, sdr_pool** out);
int sdr_pool_bind(sdr_pool* p /* nullable */);
int sdr_pool_get_stats(const sdr_pool* p, sdr_pool_stats* out);
/* Free the cached blocks (the live ones stay). */
int sdr_pool_trim(sdr_pool* p);
/* SDR_E_STATE while buffers made from it are alive; unbind it from every thread first. */
int sdr_pool_free(sdr_pool* p);

/* Estimated device footprint (bytes) of a solver of the given grid, particle capacity
 * and precision (the admission estimate). */
int sdr_solver_footprint(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t particles, int32_t precision,
                         uint64_t* bytes);

/* Serve `n` scenarios (arrival order need not be sorted); blocks until all finish.
 * out[i] belongs to scenario i.  Returns SDR_OK when every scenario ran (individual
 * failures are in out[i].status). */
int sdr_serve(const sdr_scenario* scenarios, uint32_t n, const sdr_serve_config* cfg,
              sdr_scenario_result* out);

#ifdef __cplusplus
}
#endif

#endif /* SDR_AMD_H */
