//! Safe Rust binding of `libsdr_amd.so`, the MI355X (gfx950) FLIP/PIC substep of `sdr`.
//!
//! The types and methods keep the reference crate's names and meaning
//! (uselessgoddess/sdr `src/solver.rs:41-310`, `src/pressure.rs:36-239`,
//! `src/sdf.rs:47-86`, `src/metrics.rs:116-446`, `src/output.rs:23-122`), so the
//! reference's `Solver` can forward its hot path here behind a cargo feature.  Points are
//! `[f64; 3]`, the memory layout of the reference's `Vec3` (`src/math.rs:13-17`), so a
//! `Vec<Vec3>` is passed as `&[[f64; 3]]` with a pointer cast and no copy.
//!
//! Every call returns `Result<_, Error>`: the library's status code and the message of
//! `sdr_last_error()`.  There is no CPU fallback: without a gfx950 device the
//! constructors fail with `SDR_E_NODEV`.
pub mod sys;

use std::ffi::{CStr, CString};
use std::fmt;
use std::os::raw::c_int;
use std::path::Path;
use std::ptr;

pub type Vec3 = [f64; 3];

/// A failed call: the SDR_E_* code and the library's message.
#[derive(Debug, Clone)]
pub struct Error {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for Error {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "sdr_amd error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for Error {}

pub type Result<T> = std::result::Result<T, Error>;

fn check(rc: c_int) -> Result<()> {
    if rc == sys::SDR_OK {
        return Ok(());
    }
    let message = unsafe {
        let p = sys::sdr_last_error();
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    };
    Err(Error { code: rc, message })
}

fn invalid(message: &str) -> Error {
    Error { code: sys::SDR_E_INVALID, message: message.to_string() }
}

fn v3(v: Vec3) -> sys::sdr_vec3 {
    sys::sdr_vec3 { x: v[0], y: v[1], z: v[2] }
}

fn flat(points: &[Vec3]) -> *const f64 {
    points.as_ptr() as *const f64
}

fn flat_mut(points: &mut [Vec3]) -> *mut f64 {
    points.as_mut_ptr() as *mut f64
}

fn path_c(path: &Path) -> Result<CString> {
    let s = path.to_str().ok_or_else(|| invalid("path is not valid UTF-8"))?;
    CString::new(s).map_err(|_| invalid("path holds a NUL byte"))
}

/// `Aabb` (math.rs:198-201).
#[derive(Debug, Clone, Copy, Default)]
pub struct Aabb {
    pub min: Vec3,
    pub max: Vec3,
}

impl Aabb {
    fn c(&self) -> sys::sdr_aabb {
        sys::sdr_aabb { min: v3(self.min), max: v3(self.max) }
    }
}

/// `Sdf` (sdf.rs:23-34): node-grid signed distance, x fastest, negative inside.
#[derive(Debug, Clone)]
pub struct Sdf {
    pub origin: Vec3,
    pub dx: f64,
    pub nx: usize,
    pub ny: usize,
    pub nz: usize,
    pub data: Vec<f64>,
}

impl Sdf {
    fn c(&self) -> sys::sdr_sdf {
        sys::sdr_sdf {
            origin: v3(self.origin),
            dx: self.dx,
            nx: self.nx as u64,
            ny: self.ny as u64,
            nz: self.nz as u64,
            data: self.data.as_ptr(),
        }
    }

    /// `Sdf::from_mesh(mesh, dx, padding)` / `from_mesh_in_bounds` (sdf.rs:47-86) on
    /// device `device`: winding-number sign, exact point-triangle distance.
    pub fn from_mesh(vertices: &[Vec3], tris: &[[u32; 3]], dx: f64, padding: f64, bounds: Option<Aabb>,
                     device: i32) -> Result<Sdf> {
        let b = bounds.map(|a| a.c());
        let bp = b.as_ref().map_or(ptr::null(), |x| x as *const sys::sdr_aabb);
        let mut d = sys::sdr_sdf { origin: v3([0.0; 3]), dx: 0.0, nx: 0, ny: 0, nz: 0, data: ptr::null() };
        let tp = tris.as_ptr() as *const u32;
        check(unsafe {
            sys::sdr_sdf_from_mesh(flat(vertices), vertices.len() as u64, tp, tris.len() as u64, dx, bp,
                                   padding, device, &mut d, ptr::null_mut(), 0)
        })?;
        let n = (d.nx * d.ny * d.nz) as usize;
        let mut data = vec![0.0f64; n];
        check(unsafe {
            sys::sdr_sdf_from_mesh(flat(vertices), vertices.len() as u64, tp, tris.len() as u64, dx, bp,
                                   padding, device, &mut d, data.as_mut_ptr(), n as u64)
        })?;
        Ok(Sdf {
            origin: [d.origin.x, d.origin.y, d.origin.z],
            dx: d.dx,
            nx: d.nx as usize,
            ny: d.ny as usize,
            nz: d.nz as usize,
            data,
        })
    }
}

/// `FluidParams` (solver.rs:41-85) with its defaults.
#[derive(Debug, Clone, Copy)]
pub struct FluidParams {
    pub density: f64,
    pub gravity: Vec3,
    pub flip_ratio: f64,
    pub particles_per_cell: usize,
    pub volume_stiffness: f64,
    pub redistribute: bool,
    pub redist_cap_factor: f64,
}

impl Default for FluidParams {
    fn default() -> Self {
        FluidParams {
            density: 1000.0,
            gravity: [0.0, -9.81, 0.0],
            flip_ratio: 0.95,
            particles_per_cell: 8,
            volume_stiffness: 1.0,
            redistribute: true,
            redist_cap_factor: 1.0,
        }
    }
}

impl FluidParams {
    fn c(&self) -> sys::sdr_fluid_params {
        sys::sdr_fluid_params {
            density: self.density,
            gravity: v3(self.gravity),
            flip_ratio: self.flip_ratio,
            particles_per_cell: self.particles_per_cell as u64,
            volume_stiffness: self.volume_stiffness,
            redistribute: self.redistribute as i32,
            redist_cap_factor: self.redist_cap_factor,
        }
    }
}

/// `Needle` (solver.rs:90-114).
#[derive(Debug, Clone, Copy)]
pub struct Needle {
    pub tip: Vec3,
    pub axis: Vec3,
    pub radius: f64,
    pub flow_rate: f64,
    pub inject_until: f64,
}

impl Needle {
    /// Mean jet speed from the flow rate and bore (solver.rs:107-114).
    pub fn jet_speed(&self) -> f64 {
        let area = std::f64::consts::PI * self.radius * self.radius;
        if area > 0.0 {
            self.flow_rate / area
        } else {
            0.0
        }
    }

    fn c(&self) -> sys::sdr_needle {
        sys::sdr_needle {
            tip: v3(self.tip),
            axis: v3(self.axis),
            radius: self.radius,
            flow_rate: self.flow_rate,
            inject_until: self.inject_until,
        }
    }
}

/// `SolveParams` (pressure.rs:36-50).
#[derive(Debug, Clone, Copy)]
pub struct SolveParams {
    pub max_iters: usize,
    pub tol: f64,
}

impl Default for SolveParams {
    fn default() -> Self {
        SolveParams { max_iters: 200, tol: 1e-6 }
    }
}

/// `SolveReport` (pressure.rs:54-57).
#[derive(Debug, Clone, Copy, Default)]
pub struct SolveReport {
    pub iters: usize,
    pub residual: f64,
}

/// The device path of a solver: precision, reduction order, device, stopping rule.
#[derive(Debug, Clone, Copy)]
pub struct Options {
    /// `sys::SDR_PRECISION_F64` (the plain drop-in) or `sys::SDR_PRECISION_F32`.
    pub precision: i32,
    /// `sys::SDR_DOT_EXACT` (the reference's dot tree, bit-identical) or `SDR_DOT_FAST`.
    pub dot_mode: i32,
    pub device: i32,
    /// relative residual target of `SDR_SOLVE_RELATIVE`
    pub refine_tol: f64,
    /// `sys::SDR_SOLVE_REFERENCE` (inf_norm(r) <= tol) or `SDR_SOLVE_RELATIVE`.
    pub solve_mode: i32,
}

impl Default for Options {
    fn default() -> Self {
        Options {
            precision: sys::SDR_PRECISION_F64,
            dot_mode: sys::SDR_DOT_EXACT,
            device: 0,
            refine_tol: 1e-8,
            solve_mode: sys::SDR_SOLVE_REFERENCE,
        }
    }
}

impl Options {
    fn c(&self) -> sys::sdr_options {
        sys::sdr_options {
            precision: self.precision,
            dot_mode: self.dot_mode,
            device: self.device,
            refine_tol: self.refine_tol,
            solve_mode: self.solve_mode,
            _pad: 0,
        }
    }
}

/// Scalar state: `time, steps, last_pcg_iters, last_substeps, drained` (solver.rs:152-165)
/// and the emission state.
pub type State = sys::sdr_state;
/// The pub diagnostics of `Solver` (solver.rs:225-233, 758-891), reduced on the device.
pub type Diagnostics = sys::sdr_diagnostics;
/// Counts, area and signed volume of a particle surface.
pub type SurfaceInfo = sys::sdr_surface_info;

/// The reference `Solver`, its particle and grid state resident on the device.
pub struct Solver {
    h: *mut sys::sdr_solver,
    pub nx: usize,
    pub ny: usize,
    pub nz: usize,
    pub dx: f64,
    pub origin: Vec3,
}

// A solver owns one HIP stream; it may move between threads but is not shared.
unsafe impl Send for Solver {}

impl Solver {
    /// `Solver::new(solid, sim_dx, fluid, needle)` (solver.rs:184-222).
    pub fn new(solid: &Sdf, sim_dx: f64, fluid: FluidParams, needle: Needle, opts: Options) -> Result<Solver> {
        if solid.data.len() != solid.nx * solid.ny * solid.nz {
            return Err(invalid("Sdf data does not cover its node grid"));
        }
        let mut h = ptr::null_mut();
        let (sd, fp, nd, op) = (solid.c(), fluid.c(), needle.c(), opts.c());
        check(unsafe { sys::sdr_solver_new(&sd, sim_dx, &fp, &nd, &op, &mut h) })?;
        let mut d = sys::sdr_dims { nx: 0, ny: 0, nz: 0, dx: 0.0, origin: v3([0.0; 3]) };
        check(unsafe { sys::sdr_solver_get_dims(h, &mut d) })?;
        Ok(Solver {
            h,
            nx: d.nx as usize,
            ny: d.ny as usize,
            nz: d.nz as usize,
            dx: d.dx,
            origin: [d.origin.x, d.origin.y, d.origin.z],
        })
    }

    pub fn raw(&self) -> *mut sys::sdr_solver {
        self.h
    }

    /// `solver.outlet = Some(aabb)` / `None` (solver.rs:150).
    pub fn set_outlet(&mut self, outlet: Option<Aabb>) -> Result<()> {
        let o = outlet.map(|a| a.c());
        let p = o.as_ref().map_or(ptr::null(), |x| x as *const sys::sdr_aabb);
        check(unsafe { sys::sdr_solver_set_outlet(self.h, p) })
    }

    /// `solver.solve = params` (solver.rs:151).
    pub fn set_solve(&mut self, p: SolveParams) -> Result<()> {
        let c = sys::sdr_solve_params { max_iters: p.max_iters as u64, tol: p.tol };
        check(unsafe { sys::sdr_solver_set_solve_params(self.h, &c) })
    }

    /// `solver.needle = needle` (solver.rs:148).
    pub fn set_needle(&mut self, n: Needle) -> Result<()> {
        let c = n.c();
        check(unsafe { sys::sdr_solver_set_needle(self.h, &c) })
    }

    /// `Solver::step(dt)` (solver.rs:237-259).
    pub fn step(&mut self, dt: f64) -> Result<()> {
        check(unsafe { sys::sdr_solver_step(self.h, dt) })
    }

    /// `Solver::substep(dt)` (solver.rs:284-310).
    pub fn substep(&mut self, dt: f64) -> Result<()> {
        check(unsafe { sys::sdr_solver_substep(self.h, dt) })
    }

    /// `Solver::redistribute()` (solver.rs:454-557).
    pub fn redistribute(&mut self) -> Result<()> {
        check(unsafe { sys::sdr_solver_redistribute(self.h) })
    }

    pub fn particle_count(&self) -> Result<usize> {
        let mut n = 0u64;
        check(unsafe { sys::sdr_solver_particle_count(self.h, &mut n) })?;
        Ok(n as usize)
    }

    /// `solver.particles` (particles.rs:13-16) in reference order: positions, velocities.
    pub fn particles(&mut self) -> Result<(Vec<Vec3>, Vec<Vec3>)> {
        let n = self.particle_count()?;
        let mut pos = vec![[0.0; 3]; n];
        let mut vel = vec![[0.0; 3]; n];
        check(unsafe { sys::sdr_solver_get_particles(self.h, flat_mut(&mut pos), flat_mut(&mut vel), n as u64) })?;
        Ok((pos, vel))
    }

    pub fn set_particles(&mut self, pos: &[Vec3], vel: &[Vec3]) -> Result<()> {
        if pos.len() != vel.len() {
            return Err(invalid("positions and velocities differ in length"));
        }
        check(unsafe { sys::sdr_solver_set_particles(self.h, flat(pos), flat(vel), pos.len() as u64) })
    }

    fn face_counts(&self) -> (usize, usize, usize) {
        let (x, y, z) = (self.nx, self.ny, self.nz);
        ((x + 1) * y * z, x * (y + 1) * z, x * y * (z + 1))
    }

    /// `solver.grid.{u, v, w}` (grid.rs:92-101).
    pub fn grid(&mut self) -> Result<(Vec<f64>, Vec<f64>, Vec<f64>)> {
        let (a, b, c) = self.face_counts();
        let (mut u, mut v, mut w) = (vec![0.0; a], vec![0.0; b], vec![0.0; c]);
        check(unsafe { sys::sdr_solver_get_grid(self.h, u.as_mut_ptr(), v.as_mut_ptr(), w.as_mut_ptr()) })?;
        Ok((u, v, w))
    }

    pub fn set_grid(&mut self, u: &[f64], v: &[f64], w: &[f64]) -> Result<()> {
        let (a, b, c) = self.face_counts();
        if u.len() != a || v.len() != b || w.len() != c {
            return Err(invalid("face arrays do not match the grid"));
        }
        check(unsafe { sys::sdr_solver_set_grid(self.h, u.as_ptr(), v.as_ptr(), w.as_ptr()) })
    }

    /// `solver.pressure` (solver.rs:163), one value per cell.
    pub fn pressure(&mut self) -> Result<Vec<f64>> {
        let mut p = vec![0.0; self.nx * self.ny * self.nz];
        check(unsafe { sys::sdr_solver_get_pressure(self.h, p.as_mut_ptr()) })?;
        Ok(p)
    }

    /// Cell classes (0 Solid, 1 Fluid, 2 Air; pressure.rs:24-32).
    pub fn cells(&mut self) -> Result<Vec<u8>> {
        let mut c = vec![0u8; self.nx * self.ny * self.nz];
        check(unsafe { sys::sdr_solver_get_cells(self.h, c.as_mut_ptr()) })?;
        Ok(c)
    }

    pub fn state(&self) -> Result<State> {
        let mut s = State::default();
        check(unsafe { sys::sdr_solver_get_state(self.h, &mut s) })?;
        Ok(s)
    }

    pub fn set_state(&mut self, s: &State) -> Result<()> {
        check(unsafe { sys::sdr_solver_set_state(self.h, s) })
    }

    /// `particle_volume`, `fluid_volume`, `drained_volume`, `cells_fluid_count`,
    /// `max_particles_per_cell`, `max_pressure`, `mean_pressure`, `max_speed`
    /// (solver.rs:225-233, 758-891) in one device reduction.
    pub fn diagnostics(&mut self) -> Result<Diagnostics> {
        let mut d = Diagnostics::default();
        check(unsafe { sys::sdr_solver_diagnostics(self.h, &mut d) })?;
        Ok(d)
    }

    /// The private `Solver::cfl_dt` (solver.rs:262-281) on the current grid.
    pub fn cfl_dt(&mut self) -> Result<f64> {
        let mut dt = 0.0;
        check(unsafe { sys::sdr_solver_cfl_dt(self.h, &mut dt) })?;
        Ok(dt)
    }

    /// Surface the resident particles: the particle distance field (balls of `radius`,
    /// 0 = dx / cbrt(particles_per_cell)) on nodes `spacing` apart (0 = dx), polygonised
    /// by naive surface nets (surface_nets.rs:48-221).  The mesh stays on the device.
    pub fn surface(&mut self, radius: f64, spacing: f64) -> Result<SurfaceInfo> {
        let p = sys::sdr_surface_params { radius, spacing };
        let mut info = SurfaceInfo::default();
        check(unsafe { sys::sdr_solver_surface(self.h, &p, &mut info) })?;
        Ok(info)
    }

    /// The last surface: vertices and triangles (surface_nets.rs numbering).
    pub fn surface_mesh(&mut self, info: &SurfaceInfo) -> Result<(Vec<Vec3>, Vec<[u32; 3]>)> {
        let mut v = vec![[0.0; 3]; info.vertices as usize];
        let mut t = vec![[0u32; 3]; info.triangles as usize];
        check(unsafe {
            sys::sdr_solver_get_surface(self.h, flat_mut(&mut v), v.len() as u64, t.as_mut_ptr() as *mut u32,
                                        t.len() as u64)
        })?;
        Ok((v, t))
    }

    /// Block until the solver's queued device work has finished.
    pub fn synchronize(&mut self) -> Result<()> {
        check(unsafe { sys::sdr_solver_synchronize(self.h) })
    }
}

impl Drop for Solver {
    fn drop(&mut self) {
        if !self.h.is_null() {
            unsafe {
                sys::sdr_solver_free(self.h);
            }
            self.h = ptr::null_mut();
        }
    }
}

/// `FrameMetrics` (metrics.rs:27-67) and `Summary` (metrics.rs:71-107).
pub type FrameMetrics = sys::sdr_frame_metrics;
pub type Summary = sys::sdr_summary;

/// `MetricsCollector` (metrics.rs:116-446) reduced on the device.
pub struct MetricsCollector {
    h: *mut sys::sdr_metrics,
}

unsafe impl Send for MetricsCollector {}

impl MetricsCollector {
    /// `MetricsCollector::new(solver, cavity_volume)` (metrics.rs:119-182).
    pub fn new(solver: &mut Solver, cavity_volume: f64) -> Result<MetricsCollector> {
        let mut h = ptr::null_mut();
        check(unsafe { sys::sdr_metrics_new(solver.h, cavity_volume, &mut h) })?;
        Ok(MetricsCollector { h })
    }

    /// `MetricsCollector::record(solver, frame)` (metrics.rs:188-316).
    pub fn record(&mut self, solver: &mut Solver, frame: usize) -> Result<FrameMetrics> {
        let mut m = FrameMetrics::default();
        check(unsafe { sys::sdr_metrics_record(self.h, solver.h, frame as u64, &mut m) })?;
        Ok(m)
    }

    /// `MetricsCollector::summary()` (metrics.rs:324-382).
    pub fn summary(&self) -> Result<Summary> {
        let mut s = Summary::default();
        check(unsafe { sys::sdr_metrics_summary(self.h, &mut s) })?;
        Ok(s)
    }
}

impl Drop for MetricsCollector {
    fn drop(&mut self) {
        if !self.h.is_null() {
            unsafe {
                sys::sdr_metrics_free(self.h);
            }
        }
    }
}

/// `write_points_vtk` (output.rs:23-47).
pub fn write_points_vtk(path: &Path, positions: &[Vec3]) -> Result<()> {
    let p = path_c(path)?;
    check(unsafe { sys::sdr_write_points_vtk(p.as_ptr(), flat(positions), positions.len() as u64) })
}

/// `read_points_vtk` (output.rs:56-95).
pub fn read_points_vtk(path: &Path) -> Result<Vec<Vec3>> {
    let p = path_c(path)?;
    let mut n = 0u64;
    check(unsafe { sys::sdr_read_points_vtk(p.as_ptr(), ptr::null_mut(), 0, &mut n) })?;
    let mut out = vec![[0.0; 3]; n as usize];
    check(unsafe { sys::sdr_read_points_vtk(p.as_ptr(), flat_mut(&mut out), n, &mut n) })?;
    Ok(out)
}

/// `write_points_xyz` (output.rs:99-112).
pub fn write_points_xyz(path: &Path, positions: &[Vec3]) -> Result<()> {
    let p = path_c(path)?;
    check(unsafe { sys::sdr_write_points_xyz(p.as_ptr(), flat(positions), positions.len() as u64) })
}

/// Asynchronous frame files: the positions staged on the device and written from a
/// host thread while the following substeps run.
pub struct FrameWriter {
    h: *mut sys::sdr_frame_writer,
}

unsafe impl Send for FrameWriter {}

impl FrameWriter {
    /// `format`: `sys::SDR_FORMAT_VTK` or `sys::SDR_FORMAT_XYZ`; `slots`: 1-8 frames in flight.
    pub fn new(dir: &Path, prefix: &str, format: i32, slots: u32) -> Result<FrameWriter> {
        let d = path_c(dir)?;
        let p = CString::new(prefix).map_err(|_| invalid("prefix holds a NUL byte"))?;
        let mut h = ptr::null_mut();
        check(unsafe { sys::sdr_frame_writer_new(d.as_ptr(), p.as_ptr(), format, slots, &mut h) })?;
        Ok(FrameWriter { h })
    }

    pub fn submit(&mut self, solver: &mut Solver, frame: usize) -> Result<()> {
        check(unsafe { sys::sdr_frame_writer_submit(self.h, solver.h, frame as u64) })
    }

    /// Wait for every submitted frame; returns the frames written so far.
    pub fn flush(&mut self) -> Result<u64> {
        let mut n = 0u64;
        check(unsafe { sys::sdr_frame_writer_flush(self.h, &mut n) })?;
        Ok(n)
    }
}

impl Drop for FrameWriter {
    fn drop(&mut self) {
        if !self.h.is_null() {
            unsafe {
                sys::sdr_frame_writer_free(self.h);
            }
        }
    }
}

/// Face arrays of a MAC grid (grid.rs:92-101) for the standalone projection.
pub struct Faces<'a> {
    pub u: &'a mut [f64],
    pub v: &'a mut [f64],
    pub w: &'a mut [f64],
}

/// `pressure::project_capturing_with_source(grid, cells, rho, dt, params, source)`
/// (pressure.rs:123-239): projects the faces in place and returns the report and the
/// cell pressures.
pub fn project_capturing_with_source(dims: (usize, usize, usize), dx: f64, origin: Vec3, faces: Faces<'_>,
                                     cells: &[u8], rho: f64, dt: f64, params: SolveParams,
                                     source: Option<&[f64]>, opts: Options) -> Result<(SolveReport, Vec<f64>)> {
    let (nx, ny, nz) = dims;
    let n = nx * ny * nz;
    if cells.len() != n || faces.u.len() != (nx + 1) * ny * nz || faces.v.len() != nx * (ny + 1) * nz
        || faces.w.len() != nx * ny * (nz + 1) || source.map_or(false, |s| s.len() != n)
    {
        return Err(invalid("array sizes do not match the grid"));
    }
    let g = sys::sdr_dims { nx: nx as u64, ny: ny as u64, nz: nz as u64, dx, origin: v3(origin) };
    let sp = sys::sdr_solve_params { max_iters: params.max_iters as u64, tol: params.tol };
    let mut rep = sys::sdr_solve_report { iters: 0, residual: 0.0 };
    let mut p = vec![0.0; n];
    let op = opts.c();
    check(unsafe {
        sys::sdr_project_capturing_with_source(&g, faces.u.as_mut_ptr(), faces.v.as_mut_ptr(), faces.w.as_mut_ptr(),
                                               cells.as_ptr(), rho, dt, &sp,
                                               source.map_or(ptr::null(), |s| s.as_ptr()), p.as_mut_ptr(), &mut rep,
                                               &op)
    })?;
    Ok((SolveReport { iters: rep.iters as usize, residual: rep.residual }, p))
}
