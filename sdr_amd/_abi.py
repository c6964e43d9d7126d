"""ctypes declarations of the C ABI in include/sdr_amd.h.

The same declarations bind any library that exports the ABI under a prefix, so the
Python mirror of the reference `Solver` API (sdr_amd.solver) drives the device
library; test code binds the CPU oracle through the same declarations.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

c_double_p = C.POINTER(C.c_double)
c_u8_p = C.POINTER(C.c_uint8)
c_u64_p = C.POINTER(C.c_uint64)

SDR_OK = 0
SDR_E_INVALID = -1
SDR_E_OOM = -2
SDR_E_HIP = -3
SDR_E_NODEV = -4
SDR_E_STATE = -5
SDR_E_NUMERIC = -6
SDR_E_IO = -7

FORMAT_VTK = 0
FORMAT_XYZ = 1

PRECISION_F64 = 0
PRECISION_F32 = 1
DOT_EXACT = 0
DOT_FAST = 1
SOLVE_REFERENCE = 0
SOLVE_RELATIVE = 1

STAGES = [
    "emit", "classify", "p2g", "save", "forces", "project", "extrapolate", "g2p", "advect",
    "bounds", "redistribute", "finish",
]
STAGE = {name: i for i, name in enumerate(STAGES)}


class Vec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Aabb(C.Structure):
    _fields_ = [("min", Vec3), ("max", Vec3)]


class SdfDesc(C.Structure):
    _fields_ = [("origin", Vec3), ("dx", C.c_double), ("nx", C.c_uint64), ("ny", C.c_uint64),
                ("nz", C.c_uint64), ("data", c_double_p)]


class FluidParams(C.Structure):
    _fields_ = [("density", C.c_double), ("gravity", Vec3), ("flip_ratio", C.c_double),
                ("particles_per_cell", C.c_uint64), ("volume_stiffness", C.c_double),
                ("redistribute", C.c_int32), ("redist_cap_factor", C.c_double)]


class NeedleC(C.Structure):
    _fields_ = [("tip", Vec3), ("axis", Vec3), ("radius", C.c_double), ("flow_rate", C.c_double),
                ("inject_until", C.c_double)]


class SolveParamsC(C.Structure):
    _fields_ = [("max_iters", C.c_uint64), ("tol", C.c_double)]


class SolveReportC(C.Structure):
    _fields_ = [("iters", C.c_uint64), ("residual", C.c_double)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("dot_mode", C.c_int32), ("device", C.c_int32),
                ("refine_tol", C.c_double), ("solve_mode", C.c_int32), ("_pad", C.c_int32)]


class State(C.Structure):
    _fields_ = [("time", C.c_double), ("steps", C.c_uint64), ("last_pcg_iters", C.c_uint64),
                ("last_substeps", C.c_uint32), ("redist_rounds", C.c_uint32), ("drained", C.c_uint64),
                ("rng", C.c_uint64), ("emit_accum", C.c_double), ("emitted", C.c_uint64),
                ("culled", C.c_uint64), ("last_residual", C.c_double),
                ("redist_active", C.c_uint64), ("redist_moved", C.c_uint64)]


class Dims(C.Structure):
    _fields_ = [("nx", C.c_uint64), ("ny", C.c_uint64), ("nz", C.c_uint64), ("dx", C.c_double),
                ("origin", Vec3)]


class Scenario(C.Structure):
    _fields_ = [("solid", SdfDesc), ("sim_dx", C.c_double), ("fluid", FluidParams),
                ("needle", NeedleC), ("opts", Options), ("solve", SolveParamsC),
                ("has_outlet", C.c_int32), ("outlet", Aabb), ("fill_per_cell", C.c_uint32),
                ("fill_y_lo", C.c_double), ("fill_y_hi", C.c_double), ("fill_seed", C.c_uint64),
                ("dt", C.c_double), ("substeps", C.c_uint32), ("arrival_s", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_double), ("max_slots", C.c_uint32), ("resident_pools", C.c_uint32)]


class PoolStats(C.Structure):
    _fields_ = [("hits", C.c_uint64), ("misses", C.c_uint64), ("cached_bytes", C.c_uint64),
                ("live_bytes", C.c_uint64), ("cached_blocks", C.c_uint64), ("live_blocks", C.c_uint64)]


class ScenarioResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("slot", C.c_uint32), ("arrival_s", C.c_double),
                ("admit_s", C.c_double), ("finish_s", C.c_double), ("footprint_bytes", C.c_uint64),
                ("particles", C.c_uint64), ("particle_substeps", C.c_uint64),
                ("pcg_iters", C.c_uint64), ("checksum", C.c_uint64), ("device_allocs", C.c_uint64)]


class Diagnostics(C.Structure):
    _fields_ = [("particle_volume", C.c_double), ("fluid_volume", C.c_double),
                ("drained_volume", C.c_double), ("cells_fluid_count", C.c_uint64),
                ("max_particles_per_cell", C.c_uint64), ("max_pressure", C.c_double),
                ("mean_pressure", C.c_double), ("max_speed", C.c_double),
                ("kinetic_energy", C.c_double), ("particles", C.c_uint64)]


class FrameMetrics(C.Structure):
    _fields_ = [("frame", C.c_uint64), ("time_s", C.c_double), ("particles", C.c_uint64),
                ("fluid_volume_ml", C.c_double), ("drained_volume_ml", C.c_double),
                ("fill_fraction", C.c_double), ("wall_coverage", C.c_double),
                ("mean_pressure_pa", C.c_double), ("max_pressure_pa", C.c_double),
                ("max_pressure_mmhg", C.c_double), ("wall_pressure_pa", C.c_double),
                ("peak_wall_pressure_pa", C.c_double), ("peak_wall_pressure_mmhg", C.c_double),
                ("membrane_artifacts", C.c_uint64), ("max_speed_m_s", C.c_double),
                ("kinetic_energy", C.c_double), ("pcg_iters", C.c_uint64),
                ("wet_wall_cells", C.c_uint64), ("emitted", C.c_uint64), ("drained", C.c_uint64),
                ("culled", C.c_uint64)]


class Summary(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("final_time_s", C.c_double),
                ("peak_fill_fraction", C.c_double), ("final_fill_fraction", C.c_double),
                ("peak_wall_coverage", C.c_double), ("final_wall_coverage", C.c_double),
                ("peak_membrane_pressure_pa", C.c_double), ("peak_membrane_pressure_mmhg", C.c_double),
                ("mean_membrane_pressure_pa", C.c_double), ("mean_membrane_pressure_mmhg", C.c_double),
                ("peak_pressure_pa", C.c_double), ("peak_pressure_mmhg", C.c_double),
                ("membrane_artifacts_total", C.c_uint64), ("total_drained_ml", C.c_double),
                ("injected_ml", C.c_double), ("mass_balance", C.c_double), ("emitted", C.c_uint64),
                ("resident", C.c_uint64), ("drained", C.c_uint64), ("culled", C.c_uint64)]


class SurfaceParams(C.Structure):
    _fields_ = [("radius", C.c_double), ("spacing", C.c_double)]


class SurfaceInfo(C.Structure):
    _fields_ = [("nx", C.c_uint64), ("ny", C.c_uint64), ("nz", C.c_uint64), ("origin", Vec3),
                ("spacing", C.c_double), ("radius", C.c_double), ("vertices", C.c_uint64),
                ("triangles", C.c_uint64), ("area", C.c_double), ("volume", C.c_double)]


def struct_dict(st) -> dict:
    return {name: getattr(st, name) for name, _ in st._fields_}


P = C.c_void_p
# name -> argtypes (restype is always c_int unless listed in _RESTYPE)
SIGNATURES = {
    "last_error": [],
    "solver_new": [C.POINTER(SdfDesc), C.c_double, C.POINTER(FluidParams), C.POINTER(NeedleC),
                   C.POINTER(Options), C.POINTER(P)],
    "solver_free": [P],
    "solver_set_outlet": [P, C.POINTER(Aabb)],
    "solver_set_solve_params": [P, C.POINTER(SolveParamsC)],
    "solver_set_needle": [P, C.POINTER(NeedleC)],
    "solver_step": [P, C.c_double],
    "solver_substep": [P, C.c_double],
    "solver_run_stage": [P, C.c_int32, C.c_double],
    "solver_redistribute": [P],
    "solver_particle_count": [P, c_u64_p],
    "solver_get_particles": [P, c_double_p, c_double_p, C.c_uint64],
    "solver_set_particles": [P, c_double_p, c_double_p, C.c_uint64],
    "solver_get_dims": [P, C.POINTER(Dims)],
    "solver_get_grid": [P, c_double_p, c_double_p, c_double_p],
    "solver_set_grid": [P, c_double_p, c_double_p, c_double_p],
    "solver_get_saved": [P, c_double_p, c_double_p, c_double_p],
    "solver_set_saved": [P, c_double_p, c_double_p, c_double_p],
    "solver_get_cells": [P, c_u8_p],
    "solver_set_cells": [P, c_u8_p],
    "solver_get_pressure": [P, c_double_p],
    "solver_get_state": [P, C.POINTER(State)],
    "solver_set_state": [P, C.POINTER(State)],
    "project_capturing_with_source": [C.POINTER(Dims), c_double_p, c_double_p, c_double_p, c_u8_p,
                                      C.c_double, C.c_double, C.POINTER(SolveParamsC), c_double_p,
                                      c_double_p, C.POINTER(SolveReportC), C.POINTER(Options)],
    "solver_diagnostics": [P, C.POINTER(Diagnostics)],
    "solver_cfl_dt": [P, c_double_p],
    "metrics_new": [P, C.c_double, C.POINTER(P)],
    "metrics_free": [P],
    "metrics_wall_count": [P, c_u64_p],
    "metrics_record": [P, P, C.c_uint64, C.POINTER(FrameMetrics)],
    "metrics_summary": [P, C.POINTER(Summary)],
    "metrics_get_cells": [P, c_u8_p, c_u8_p],
    "sdf_from_mesh": [c_double_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64, C.c_double, C.POINTER(Aabb),
                      C.c_double, C.c_int32, C.POINTER(SdfDesc), c_double_p, C.c_uint64],
    "solver_surface": [P, C.POINTER(SurfaceParams), C.POINTER(SurfaceInfo)],
    "solver_get_surface": [P, c_double_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64],
    "solver_get_surface_field": [P, c_double_p, C.c_uint64],
    "surface_nets": [c_double_p, C.c_uint64, C.c_uint64, C.c_uint64, Vec3, C.c_double, C.c_int32,
                     C.POINTER(SurfaceInfo), c_double_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64],
}
# Entry points only the device library has.
DEVICE_ONLY = {
    "device_info": [C.POINTER(C.c_int32), C.c_char_p, C.c_size_t],
    "solver_synchronize": [P],
    "solver_enable_timing": [P, C.c_int32],
    "solver_get_timing": [P, C.c_char_p, C.c_size_t, c_double_p, c_u64_p, C.c_uint64, c_u64_p],
    "solver_seed_fill": [P, C.c_uint32, C.c_double, C.c_double, C.c_uint64, C.POINTER(Vec3),
                         c_u64_p],
    "solver_get_slot_ranks": [P, C.POINTER(C.c_uint32), C.c_uint64],
    "solver_footprint": [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int32, c_u64_p],
    "serve": [C.POINTER(Scenario), C.c_uint32, C.POINTER(ServeConfig), C.POINTER(ScenarioResult)],
    # device pointers (u, v, w, cells, source, pressure) and a hipStream_t
    "project_device_noalloc": [C.POINTER(Dims), P, P, P, P, C.c_double, C.c_double, C.POINTER(SolveParamsC),
                               P, P, C.POINTER(SolveReportC), C.POINTER(Options), P],
    # frame output (host code of the device library)
    "write_points_vtk": [C.c_char_p, c_double_p, C.c_uint64],
    "read_points_vtk": [C.c_char_p, c_double_p, C.c_uint64, c_u64_p],
    "write_points_xyz": [C.c_char_p, c_double_p, C.c_uint64],
    "frame_path": [C.c_char_p, C.c_char_p, C.c_uint64, C.c_char_p, C.c_char_p, C.c_size_t],
    "write_mesh_obj": [C.c_char_p, c_double_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64],
    "write_mesh_stl": [C.c_char_p, c_double_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64],
    "frame_writer_new": [C.c_char_p, C.c_char_p, C.c_int32, C.c_uint32, C.POINTER(P)],
    "frame_writer_submit": [P, P, C.c_uint64],
    "frame_writer_flush": [P, c_u64_p],
    "frame_writer_free": [P],
    "pool_new": [C.c_int32, C.POINTER(P)],
    "pool_bind": [P],
    "pool_get_stats": [P, C.POINTER(PoolStats)],
    "pool_trim": [P],
    "pool_free": [P],
}


class AbiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Lib:
    """A loaded library exporting the ABI under `prefix` (``sdr_`` or the oracle's)."""

    def __init__(self, path: str | Path, prefix: str, device: bool):
        self.path = str(path)
        self.prefix = prefix
        self.device = device
        self.dll = C.CDLL(self.path)
        sigs = dict(SIGNATURES)
        if device:
            sigs.update(DEVICE_ONLY)
        self._fns = {}
        for name, args in sigs.items():
            fn = getattr(self.dll, prefix + name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name == "last_error" else C.c_int
            self._fns[name] = fn

    def has(self, name: str) -> bool:
        return name in self._fns

    def __getattr__(self, name):
        fns = self.__dict__.get("_fns")
        if fns is None or name not in fns:
            raise AttributeError(name)
        fn = fns[name]
        if name == "last_error":
            return fn

        def call(*args):
            rc = fn(*args)
            if rc != SDR_OK:
                raise AbiError(rc, fns["last_error"]().decode(errors="replace"))
            return rc

        return call


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def u8ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(c_u8_p)


def vec3(v) -> Vec3:
    return Vec3(float(v[0]), float(v[1]), float(v[2]))
