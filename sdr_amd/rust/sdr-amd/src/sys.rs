//! Raw declarations of include/sdr_amd.h (the C ABI of libsdr_amd.so), one for one.
//! Layouts are `#[repr(C)]` mirrors of the header's structs; every function returns
//! SDR_OK (0) or a negative SDR_E_* code, with the message in `sdr_last_error()`.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

pub const SDR_OK: i32 = 0;
pub const SDR_E_INVALID: i32 = -1;
pub const SDR_E_OOM: i32 = -2;
pub const SDR_E_HIP: i32 = -3;
pub const SDR_E_NODEV: i32 = -4;
pub const SDR_E_STATE: i32 = -5;
pub const SDR_E_NUMERIC: i32 = -6;
pub const SDR_E_IO: i32 = -7;
pub const SDR_PRECISION_F64: i32 = 0;
pub const SDR_PRECISION_F32: i32 = 1;
pub const SDR_DOT_EXACT: i32 = 0;
pub const SDR_DOT_FAST: i32 = 1;
pub const SDR_SOLVE_REFERENCE: i32 = 0;
pub const SDR_SOLVE_RELATIVE: i32 = 1;
pub const SDR_STAGE_EMIT: i32 = 0;
pub const SDR_STAGE_CLASSIFY: i32 = 1;
pub const SDR_STAGE_P2G: i32 = 2;
pub const SDR_STAGE_SAVE: i32 = 3;
pub const SDR_STAGE_FORCES: i32 = 4;
pub const SDR_STAGE_PROJECT: i32 = 5;
pub const SDR_STAGE_EXTRAPOLATE: i32 = 6;
pub const SDR_STAGE_G2P: i32 = 7;
pub const SDR_STAGE_ADVECT: i32 = 8;
pub const SDR_STAGE_BOUNDS: i32 = 9;
pub const SDR_STAGE_REDISTRIBUTE: i32 = 10;
pub const SDR_STAGE_FINISH: i32 = 11;
pub const SDR_STAGE_COUNT: i32 = 12;
pub const SDR_FORMAT_VTK: i32 = 0;
pub const SDR_FORMAT_XYZ: i32 = 1;
pub const SDR_NOALLOC_SCRATCH_BYTES: u64 = 64 << 20;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_vec3 {
    pub x: f64,
    pub y: f64,
    pub z: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_aabb {
    pub min: sdr_vec3,
    pub max: sdr_vec3,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct sdr_sdf {
    pub origin: sdr_vec3,
    pub dx: f64,
    pub nx: u64,
    pub ny: u64,
    pub nz: u64,
    pub data: *const f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_fluid_params {
    pub density: f64,
    pub gravity: sdr_vec3,
    pub flip_ratio: f64,
    pub particles_per_cell: u64,
    pub volume_stiffness: f64,
    pub redistribute: i32,
    pub redist_cap_factor: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_needle {
    pub tip: sdr_vec3,
    pub axis: sdr_vec3,
    pub radius: f64,
    pub flow_rate: f64,
    pub inject_until: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_solve_params {
    pub max_iters: u64,
    pub tol: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_solve_report {
    pub iters: u64,
    pub residual: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_options {
    pub precision: i32,
    pub dot_mode: i32,
    pub device: i32,
    pub refine_tol: f64,
    pub solve_mode: i32,
    pub _pad: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_state {
    pub time: f64,
    pub steps: u64,
    pub last_pcg_iters: u64,
    pub last_substeps: u32,
    pub redist_rounds: u32,
    pub drained: u64,
    pub rng: u64,
    pub emit_accum: f64,
    pub emitted: u64,
    pub culled: u64,
    pub last_residual: f64,
    pub redist_active: u64,
    pub redist_moved: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_dims {
    pub nx: u64,
    pub ny: u64,
    pub nz: u64,
    pub dx: f64,
    pub origin: sdr_vec3,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_diagnostics {
    pub particle_volume: f64,
    pub fluid_volume: f64,
    pub drained_volume: f64,
    pub cells_fluid_count: u64,
    pub max_particles_per_cell: u64,
    pub max_pressure: f64,
    pub mean_pressure: f64,
    pub max_speed: f64,
    pub kinetic_energy: f64,
    pub particles: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_surface_params {
    pub radius: f64,
    pub spacing: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_surface_info {
    pub nx: u64,
    pub ny: u64,
    pub nz: u64,
    pub origin: sdr_vec3,
    pub spacing: f64,
    pub radius: f64,
    pub vertices: u64,
    pub triangles: u64,
    pub area: f64,
    pub volume: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_frame_metrics {
    pub frame: u64,
    pub time_s: f64,
    pub particles: u64,
    pub fluid_volume_ml: f64,
    pub drained_volume_ml: f64,
    pub fill_fraction: f64,
    pub wall_coverage: f64,
    pub mean_pressure_pa: f64,
    pub max_pressure_pa: f64,
    pub max_pressure_mmhg: f64,
    pub wall_pressure_pa: f64,
    pub peak_wall_pressure_pa: f64,
    pub peak_wall_pressure_mmhg: f64,
    pub membrane_artifacts: u64,
    pub max_speed_m_s: f64,
    pub kinetic_energy: f64,
    pub pcg_iters: u64,
    pub wet_wall_cells: u64,
    pub emitted: u64,
    pub drained: u64,
    pub culled: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_summary {
    pub frames: u64,
    pub final_time_s: f64,
    pub peak_fill_fraction: f64,
    pub final_fill_fraction: f64,
    pub peak_wall_coverage: f64,
    pub final_wall_coverage: f64,
    pub peak_membrane_pressure_pa: f64,
    pub peak_membrane_pressure_mmhg: f64,
    pub mean_membrane_pressure_pa: f64,
    pub mean_membrane_pressure_mmhg: f64,
    pub peak_pressure_pa: f64,
    pub peak_pressure_mmhg: f64,
    pub membrane_artifacts_total: u64,
    pub total_drained_ml: f64,
    pub injected_ml: f64,
    pub mass_balance: f64,
    pub emitted: u64,
    pub resident: u64,
    pub drained: u64,
    pub culled: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct sdr_scenario {
    pub solid: sdr_sdf,
    pub sim_dx: f64,
    pub fluid: sdr_fluid_params,
    pub needle: sdr_needle,
    pub opts: sdr_options,
    pub solve: sdr_solve_params,
    pub has_outlet: i32,
    pub outlet: sdr_aabb,
    pub fill_per_cell: u32,
    pub fill_y_lo: f64,
    pub fill_y_hi: f64,
    pub fill_seed: u64,
    pub dt: f64,
    pub substeps: u32,
    pub arrival_s: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_serve_config {
    pub hbm_budget_bytes: f64,
    pub max_slots: u32,
    pub resident_pools: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_scenario_result {
    pub status: i32,
    pub slot: u32,
    pub arrival_s: f64,
    pub admit_s: f64,
    pub finish_s: f64,
    pub footprint_bytes: u64,
    pub particles: u64,
    pub particle_substeps: u64,
    pub pcg_iters: u64,
    pub checksum: u64,
    pub device_allocs: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct sdr_pool_stats {
    pub hits: u64,
    pub misses: u64,
    pub cached_bytes: u64,
    pub live_bytes: u64,
    pub cached_blocks: u64,
    pub live_blocks: u64,
}

#[repr(C)]
pub struct sdr_solver {
    _opaque: [u8; 0],
}

#[repr(C)]
pub struct sdr_metrics {
    _opaque: [u8; 0],
}

#[repr(C)]
pub struct sdr_frame_writer {
    _opaque: [u8; 0],
}

#[link(name = "sdr_amd")]
extern "C" {
    pub fn sdr_last_error() -> *const c_char;
    pub fn sdr_device_info(count: *mut i32, arch: *mut c_char, arch_len: usize) -> c_int;
    pub fn sdr_solver_new(solid: *const sdr_sdf, sim_dx: f64, fluid: *const sdr_fluid_params, needle: *const sdr_needle, opts: *const sdr_options, out: *mut *mut sdr_solver) -> c_int;
    pub fn sdr_solver_free(s: *mut sdr_solver) -> c_int;
    pub fn sdr_solver_set_outlet(s: *mut sdr_solver, outlet: *const sdr_aabb) -> c_int;
    pub fn sdr_solver_set_solve_params(s: *mut sdr_solver, params: *const sdr_solve_params) -> c_int;
    pub fn sdr_solver_set_needle(s: *mut sdr_solver, needle: *const sdr_needle) -> c_int;
    pub fn sdr_solver_step(s: *mut sdr_solver, dt: f64) -> c_int;
    pub fn sdr_solver_substep(s: *mut sdr_solver, dt: f64) -> c_int;
    pub fn sdr_solver_run_stage(s: *mut sdr_solver, stage: i32, dt: f64) -> c_int;
    pub fn sdr_solver_redistribute(s: *mut sdr_solver) -> c_int;
    pub fn sdr_solver_particle_count(s: *const sdr_solver, n: *mut u64) -> c_int;
    pub fn sdr_solver_get_particles(s: *mut sdr_solver, pos_xyz: *mut f64, vel_xyz: *mut f64, n: u64) -> c_int;
    pub fn sdr_solver_set_particles(s: *mut sdr_solver, pos_xyz: *const f64, vel_xyz: *const f64, n: u64) -> c_int;
    pub fn sdr_solver_get_dims(s: *const sdr_solver, out: *mut sdr_dims) -> c_int;
    pub fn sdr_solver_get_grid(s: *mut sdr_solver, u: *mut f64, v: *mut f64, w: *mut f64) -> c_int;
    pub fn sdr_solver_set_grid(s: *mut sdr_solver, u: *const f64, v: *const f64, w: *const f64) -> c_int;
    pub fn sdr_solver_get_saved(s: *mut sdr_solver, u: *mut f64, v: *mut f64, w: *mut f64) -> c_int;
    pub fn sdr_solver_set_saved(s: *mut sdr_solver, u: *const f64, v: *const f64, w: *const f64) -> c_int;
    pub fn sdr_solver_get_cells(s: *mut sdr_solver, cells: *mut u8) -> c_int;
    pub fn sdr_solver_set_cells(s: *mut sdr_solver, cells: *const u8) -> c_int;
    pub fn sdr_solver_get_pressure(s: *mut sdr_solver, p: *mut f64) -> c_int;
    pub fn sdr_solver_get_slot_ranks(s: *mut sdr_solver, ranks: *mut u32, n: u64) -> c_int;
    pub fn sdr_solver_get_state(s: *const sdr_solver, st: *mut sdr_state) -> c_int;
    pub fn sdr_solver_set_state(s: *mut sdr_solver, st: *const sdr_state) -> c_int;
    pub fn sdr_solver_diagnostics(s: *mut sdr_solver, out: *mut sdr_diagnostics) -> c_int;
    pub fn sdr_solver_cfl_dt(s: *mut sdr_solver, dt: *mut f64) -> c_int;
    pub fn sdr_sdf_from_mesh(vertices: *const f64, nv: u64, tris: *const u32, nt: u64, dx: f64, bounds: *const sdr_aabb, padding: f64, device: i32, out: *mut sdr_sdf, data: *mut f64, cap: u64) -> c_int;
    pub fn sdr_write_points_vtk(path: *const c_char, pos_xyz: *const f64, n: u64) -> c_int;
    pub fn sdr_read_points_vtk(path: *const c_char, pos_xyz: *mut f64, cap: u64, n: *mut u64) -> c_int;
    pub fn sdr_write_points_xyz(path: *const c_char, pos_xyz: *const f64, n: u64) -> c_int;
    pub fn sdr_frame_path(dir: *const c_char, prefix: *const c_char, frame: u64, ext: *const c_char, buf: *mut c_char, len: usize) -> c_int;
    pub fn sdr_write_mesh_obj(path: *const c_char, vertices: *const f64, nv: u64, tris: *const u32, nt: u64) -> c_int;
    pub fn sdr_write_mesh_stl(path: *const c_char, vertices: *const f64, nv: u64, tris: *const u32, nt: u64) -> c_int;
    pub fn sdr_frame_writer_new(dir: *const c_char, prefix: *const c_char, format: i32, slots: u32, out: *mut *mut sdr_frame_writer) -> c_int;
    pub fn sdr_frame_writer_submit(w: *mut sdr_frame_writer, s: *mut sdr_solver, frame: u64) -> c_int;
    pub fn sdr_frame_writer_flush(w: *mut sdr_frame_writer, written: *mut u64) -> c_int;
    pub fn sdr_frame_writer_free(w: *mut sdr_frame_writer) -> c_int;
    pub fn sdr_solver_surface(s: *mut sdr_solver, params: *const sdr_surface_params, out: *mut sdr_surface_info) -> c_int;
    pub fn sdr_solver_get_surface(s: *mut sdr_solver, vertices: *mut f64, vcap: u64, tris: *mut u32, tcap: u64) -> c_int;
    pub fn sdr_solver_get_surface_field(s: *mut sdr_solver, values: *mut f64, cap: u64) -> c_int;
    pub fn sdr_surface_nets(values: *const f64, nx: u64, ny: u64, nz: u64, origin: sdr_vec3, dx: f64, device: i32, info: *mut sdr_surface_info, vertices: *mut f64, vcap: u64, tris: *mut u32, tcap: u64) -> c_int;
    pub fn sdr_metrics_new(s: *mut sdr_solver, cavity_volume: f64, out: *mut *mut sdr_metrics) -> c_int;
    pub fn sdr_metrics_free(m: *mut sdr_metrics) -> c_int;
    pub fn sdr_metrics_wall_count(m: *const sdr_metrics, n: *mut u64) -> c_int;
    pub fn sdr_metrics_record(m: *mut sdr_metrics, s: *mut sdr_solver, frame: u64, out: *mut sdr_frame_metrics) -> c_int;
    pub fn sdr_metrics_summary(m: *const sdr_metrics, out: *mut sdr_summary) -> c_int;
    pub fn sdr_metrics_get_cells(m: *const sdr_metrics, washed: *mut u8, is_wall: *mut u8) -> c_int;
    pub fn sdr_solver_synchronize(s: *mut sdr_solver) -> c_int;
    pub fn sdr_solver_enable_timing(s: *mut sdr_solver, on: i32) -> c_int;
    pub fn sdr_solver_get_timing(s: *mut sdr_solver, names: *mut c_char, name_len: usize, ms: *mut f64, calls: *mut u64, cap: u64, count: *mut u64) -> c_int;
    pub fn sdr_solver_seed_fill(s: *mut sdr_solver, per_cell: u32, y_lo: f64, y_hi: f64, seed: u64, vel: *const sdr_vec3, n_out: *mut u64) -> c_int;
    pub fn sdr_project_capturing_with_source(grid: *const sdr_dims, u: *mut f64, v: *mut f64, w: *mut f64, cells: *const u8, rho: f64, dt: f64, params: *const sdr_solve_params, source: *const f64, pressure_out: *mut f64, report: *mut sdr_solve_report, opts: *const sdr_options) -> c_int;
    pub fn sdr_project_device_noalloc(grid: *const sdr_dims, u: *mut f64, v: *mut f64, w: *mut f64, cells: *const u8, rho: f64, dt: f64, params: *const sdr_solve_params, source: *const f64, pressure_out: *mut f64, report: *mut sdr_solve_report, opts: *const sdr_options, stream: *mut c_void) -> c_int;
    pub fn sdr_pool_new(device: i32, out: *mut *mut sdr_pool) -> c_int;
    pub fn sdr_pool_bind(p: *mut sdr_pool) -> c_int;
    pub fn sdr_pool_get_stats(p: *const sdr_pool, out: *mut sdr_pool_stats) -> c_int;
    pub fn sdr_pool_trim(p: *mut sdr_pool) -> c_int;
    pub fn sdr_pool_free(p: *mut sdr_pool) -> c_int;
    pub fn sdr_solver_footprint(nx: u64, ny: u64, nz: u64, particles: u64, precision: i32, bytes: *mut u64) -> c_int;
    pub fn sdr_serve(scenarios: *const sdr_scenario, n: u32, cfg: *const sdr_serve_config, out: *mut sdr_scenario_result) -> c_int;
}
