"""Python mirror of the reference's `Solver` / `pressure` API over the C ABI.

Names, argument meaning and error behaviour follow uselessgoddess/sdr:
``Solver.new`` (src/solver.rs:184-222), ``Solver.step`` (:237-259), ``Solver.substep``
(:284-310), the pub fields of ``Solver`` (:143-166), ``FluidParams`` (:41-85),
``Needle`` (:90-114), ``Sdf`` (src/sdf.rs:23-182), ``SolveParams``/``SolveReport``
(src/pressure.rs:36-57) and ``pressure::project_capturing_with_source`` (:123-239).
Size mismatches raise (the reference panics in ``assert_eq!``).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A

SOLID, FLUID, AIR = 0, 1, 2


@dataclass
class Aabb:
    """`Aabb` (src/math.rs:198-257); `contains` is inclusive."""

    min: tuple
    max: tuple

    def contains(self, p) -> bool:
        return all(self.min[a] <= p[a] <= self.max[a] for a in range(3))

    def c(self) -> A.Aabb:
        return A.Aabb(A.vec3(self.min), A.vec3(self.max))


@dataclass
class FluidParams:
    """`FluidParams` with the reference defaults (src/solver.rs:73-85)."""

    density: float = 1000.0
    gravity: tuple = (0.0, -9.81, 0.0)
    flip_ratio: float = 0.95
    particles_per_cell: int = 8
    volume_stiffness: float = 1.0
    redistribute: bool = True
    redist_cap_factor: float = 1.0

    def c(self) -> A.FluidParams:
        return A.FluidParams(self.density, A.vec3(self.gravity), self.flip_ratio,
                             int(self.particles_per_cell), self.volume_stiffness,
                             1 if self.redistribute else 0, self.redist_cap_factor)


@dataclass
class Needle:
    """`Needle` (src/solver.rs:90-114)."""

    tip: tuple
    axis: tuple
    radius: float
    flow_rate: float
    inject_until: float = math.inf

    def jet_speed(self) -> float:
        area = math.pi * self.radius * self.radius
        return self.flow_rate / area if area > 0.0 else 0.0

    def c(self) -> A.NeedleC:
        return A.NeedleC(A.vec3(self.tip), A.vec3(self.axis), self.radius, self.flow_rate,
                         self.inject_until)


@dataclass
class SolveParams:
    """`SolveParams` (src/pressure.rs:36-50)."""

    max_iters: int = 200
    tol: float = 1e-6


@dataclass
class SolveReport:
    """`SolveReport` (src/pressure.rs:54-57)."""

    iters: int
    residual: float


@dataclass
class Sdf:
    """`Sdf` (src/sdf.rs:23-34): node values x-fastest, phi < 0 inside the cavity."""

    origin: tuple
    dx: float
    nx: int
    ny: int
    nz: int
    data: np.ndarray = field(repr=False)

    @staticmethod
    def from_fn(origin, dx, dims, f, slab_nodes: int = 1 << 22) -> "Sdf":
        """`Sdf::from_fn` (src/sdf.rs:90-116) for a numpy-vectorised `f(x, y, z)`.

        Node positions are origin + (i*dx, j*dx, k*dx), as in the reference.  f is
        evaluated on slabs of whole k planes of about `slab_nodes` nodes (it is
        elementwise), so a 10^8-node field needs a slab's temporaries, not the field's."""
        nx, ny, nz = dims
        data = np.empty(nx * ny * nz, dtype=np.float64)
        slab = max(1, slab_nodes // (nx * ny))
        for k0 in range(0, nz, slab):
            k1 = min(nz, k0 + slab)
            k, j, i = np.meshgrid(np.arange(k0, k1, dtype=np.float64), np.arange(ny, dtype=np.float64),
                                  np.arange(nx, dtype=np.float64), indexing="ij")
            x = origin[0] + i * dx
            y = origin[1] + j * dx
            z = origin[2] + k * dx
            data[k0 * nx * ny:k1 * nx * ny] = np.asarray(f(x, y, z), dtype=np.float64).reshape(-1)
        return Sdf(tuple(map(float, origin)), float(dx), nx, ny, nz, data)

    @staticmethod
    def from_mesh(vertices, tris, dx: float, padding: float = 0.0, bounds: "Aabb | None" = None, *,
                  lib: A.Lib | None = None, device: int = 0) -> "Sdf":
        """`Sdf::from_mesh(mesh, dx, padding)` / `Sdf::from_mesh_in_bounds(mesh, dx, bounds)`
        (src/sdf.rs:47-86): signed distance to a closed triangle mesh (vertices (nv, 3),
        tris (nt, 3) vertex indices), negative inside by the generalised winding number."""
        lib = lib if lib is not None else _lib_default()
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        tp = t.ctypes.data_as(C.POINTER(C.c_uint32))
        bb = C.byref(bounds.c()) if bounds is not None else None
        out = A.SdfDesc()
        lib.sdf_from_mesh(A.dptr(v), v.shape[0], tp, t.shape[0], dx, bb, padding, device, C.byref(out), None, 0)
        data = np.empty(out.nx * out.ny * out.nz, np.float64)
        lib.sdf_from_mesh(A.dptr(v), v.shape[0], tp, t.shape[0], dx, bb, padding, device, C.byref(out),
                          A.dptr(data), data.size)
        return Sdf((out.origin.x, out.origin.y, out.origin.z), float(out.dx), int(out.nx), int(out.ny),
                   int(out.nz), data)

    def bounds(self) -> Aabb:
        o = self.origin
        return Aabb(o, (o[0] + (self.nx - 1) * self.dx, o[1] + (self.ny - 1) * self.dx,
                        o[2] + (self.nz - 1) * self.dx))

    def c(self) -> A.SdfDesc:
        assert self.data.dtype == np.float64 and self.data.size == self.nx * self.ny * self.nz
        return A.SdfDesc(A.vec3(self.origin), self.dx, self.nx, self.ny, self.nz, A.dptr(self.data))


def _lib_default():
    from . import load_library

    return load_library()


class Solver:
    """The reference `Solver`, its state resident behind the C ABI.

    `Solver(solid, sim_dx, fluid, needle)` is `Solver::new`; `precision`/`dot_mode` pick
    the device path (f64 plain drop-in by default)."""

    def __init__(self, solid: Sdf, sim_dx: float, fluid: FluidParams, needle: Needle, *,
                 precision: int = A.PRECISION_F64, dot_mode: int = A.DOT_EXACT, device: int = 0,
                 refine_tol: float = 1e-8, solve_mode: int = A.SOLVE_REFERENCE,
                 lib: A.Lib | None = None):
        self.lib = lib if lib is not None else _lib_default()
        self.fluid = fluid
        self._needle = needle
        self.solid = solid
        self._h = C.c_void_p()
        opts = A.Options(precision, dot_mode, device, refine_tol, solve_mode, 0)
        sd = solid.c()
        self.lib.solver_new(C.byref(sd), sim_dx, C.byref(fluid.c()), C.byref(needle.c()),
                            C.byref(opts), C.byref(self._h))
        d = A.Dims()
        self.lib.solver_get_dims(self._h, C.byref(d))
        self.nx, self.ny, self.nz, self.dx = int(d.nx), int(d.ny), int(d.nz), float(d.dx)
        self.origin = (d.origin.x, d.origin.y, d.origin.z)
        self._outlet = None
        self._solve = SolveParams()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.solver_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- pub fields
    @property
    def outlet(self):
        return self._outlet

    @outlet.setter
    def outlet(self, box: Aabb | None):
        self._outlet = box
        self.lib.solver_set_outlet(self._h, C.byref(box.c()) if box is not None else None)

    @property
    def solve(self) -> SolveParams:
        return self._solve

    @solve.setter
    def solve(self, p: SolveParams):
        self._solve = p
        self.lib.solver_set_solve_params(self._h, C.byref(A.SolveParamsC(int(p.max_iters), p.tol)))

    @property
    def needle(self) -> Needle:
        return self._needle

    @needle.setter
    def needle(self, nd: Needle):
        self._needle = nd
        self.lib.solver_set_needle(self._h, C.byref(nd.c()))

    def state(self) -> A.State:
        st = A.State()
        self.lib.solver_get_state(self._h, C.byref(st))
        return st

    def set_state(self, st: A.State):
        self.lib.solver_set_state(self._h, C.byref(st))

    time = property(lambda s: s.state().time)
    steps = property(lambda s: s.state().steps)
    last_pcg_iters = property(lambda s: s.state().last_pcg_iters)
    last_substeps = property(lambda s: s.state().last_substeps)
    drained = property(lambda s: s.state().drained)

    # ---- methods of the reference
    def step(self, dt: float):
        self.lib.solver_step(self._h, dt)

    def substep(self, dt: float):
        self.lib.solver_substep(self._h, dt)

    def run_stage(self, stage: str | int, dt: float = 0.0):
        s = A.STAGE[stage] if isinstance(stage, str) else int(stage)
        self.lib.solver_run_stage(self._h, s, dt)

    def redistribute(self):
        self.lib.solver_redistribute(self._h)

    def particle_count(self) -> int:
        n = C.c_uint64()
        self.lib.solver_particle_count(self._h, C.byref(n))
        return int(n.value)

    def particles(self):
        """(positions, velocities) as (N,3) f64 arrays in the reference particle order."""
        n = self.particle_count()
        pos = np.zeros((n, 3), np.float64)
        vel = np.zeros((n, 3), np.float64)
        self.lib.solver_get_particles(self._h, A.dptr(pos), A.dptr(vel), n)
        return pos, vel

    def set_particles(self, pos: np.ndarray, vel: np.ndarray | None = None):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        self.lib.solver_set_particles(self._h, A.dptr(pos), A.dptr(v), pos.shape[0])

    def face_shapes(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return (nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)

    def grid(self):
        """(u, v, w) face arrays, shaped (z, y, x) of their staggered dims."""
        su, sv, sw = self.face_shapes()
        u, v, w = np.zeros(su), np.zeros(sv), np.zeros(sw)
        self.lib.solver_get_grid(self._h, A.dptr(u), A.dptr(v), A.dptr(w))
        return u, v, w

    def set_grid(self, u, v, w):
        f = lambda a: np.ascontiguousarray(a, np.float64)
        u, v, w = f(u), f(v), f(w)
        self.lib.solver_set_grid(self._h, A.dptr(u), A.dptr(v), A.dptr(w))

    def saved(self):
        su, sv, sw = self.face_shapes()
        u, v, w = np.zeros(su), np.zeros(sv), np.zeros(sw)
        self.lib.solver_get_saved(self._h, A.dptr(u), A.dptr(v), A.dptr(w))
        return u, v, w

    def set_saved(self, u, v, w):
        f = lambda a: np.ascontiguousarray(a, np.float64)
        u, v, w = f(u), f(v), f(w)
        self.lib.solver_set_saved(self._h, A.dptr(u), A.dptr(v), A.dptr(w))

    def cells(self) -> np.ndarray:
        c = np.zeros((self.nz, self.ny, self.nx), np.uint8)
        self.lib.solver_get_cells(self._h, A.u8ptr(c))
        return c

    def set_cells(self, c: np.ndarray):
        c = np.ascontiguousarray(c, np.uint8)
        self.lib.solver_set_cells(self._h, A.u8ptr(c))

    def pressure(self) -> np.ndarray:
        p = np.zeros((self.nz, self.ny, self.nx), np.float64)
        self.lib.solver_get_pressure(self._h, A.dptr(p))
        return p

    def diagnostics(self) -> dict:
        """The reference's pub diagnostics (`particle_volume`, `fluid_volume`,
        `drained_volume`, `cells_fluid_count`, `max_particles_per_cell`, `max_pressure`,
        `mean_pressure`, `max_speed`: src/solver.rs:225-233, 758-891) and
        `ParticleSet::kinetic_energy` (src/particles.rs:52-58), in one call."""
        d = A.Diagnostics()
        self.lib.solver_diagnostics(self._h, C.byref(d))
        return A.struct_dict(d)

    def particle_volume(self) -> float:
        return self.diagnostics()["particle_volume"]

    def fluid_volume(self) -> float:
        return self.diagnostics()["fluid_volume"]

    def drained_volume(self) -> float:
        return self.diagnostics()["drained_volume"]

    def cells_fluid_count(self) -> int:
        return int(self.diagnostics()["cells_fluid_count"])

    def max_particles_per_cell(self) -> int:
        return int(self.diagnostics()["max_particles_per_cell"])

    def max_pressure(self) -> float:
        return self.diagnostics()["max_pressure"]

    def mean_pressure(self) -> float:
        return self.diagnostics()["mean_pressure"]

    def max_speed(self) -> float:
        return self.diagnostics()["max_speed"]

    def cfl_dt(self) -> float:
        """The private `Solver::cfl_dt` (src/solver.rs:262-281) on the current grid."""
        out = C.c_double()
        self.lib.solver_cfl_dt(self._h, C.byref(out))
        return out.value

    # ---- particle surfacing (device; no reference counterpart in-process: the reference
    # hands frames to splashsurf, src/recon.rs:60-88)
    def surface(self, radius: float = 0.0, spacing: float = 0.0) -> dict:
        """Surface the resident particles: the particle distance field
        min(min_j |x - x_j|, 2 r) - r on a node grid of `spacing` (0 = dx) over the grid
        box, polygonised by naive surface nets (src/surface_nets.rs:48-221).  radius 0 is
        dx / cbrt(particles_per_cell).  Returns the counts, area and signed volume; the
        mesh stays resident (`surface_mesh`)."""
        info = A.SurfaceInfo()
        self.lib.solver_surface(self._h, C.byref(A.SurfaceParams(radius, spacing)), C.byref(info))
        d = A.struct_dict(info)
        d["origin"] = (info.origin.x, info.origin.y, info.origin.z)
        self._surface_info = d
        return d

    def surface_mesh(self):
        """The last surface as (vertices[nv, 3] f64, triangles[nt, 3] u32)."""
        d = self._surface_info
        v = np.zeros((int(d["vertices"]), 3))
        t = np.zeros((int(d["triangles"]), 3), np.uint32)
        self.lib.solver_get_surface(self._h, A.dptr(v), v.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    t.shape[0])
        return v, t

    def surface_field(self) -> np.ndarray:
        """The last surface's node field, shape (nz, ny, nx)."""
        d = self._surface_info
        out = np.zeros((int(d["nz"]), int(d["ny"]), int(d["nx"])))
        self.lib.solver_get_surface_field(self._h, A.dptr(out), out.size)
        return out

    def cell_center(self, i, j, k):
        o, d = self.origin, self.dx
        return (o[0] + (i + 0.5) * d, o[1] + (j + 0.5) * d, o[2] + (k + 0.5) * d)

    # ---- device-only extras
    def synchronize(self):
        self.lib.solver_synchronize(self._h)

    def enable_timing(self, on: bool = True, sweeps_only: bool = False):
        """Per-region HIP-event timing (sweeps_only: the MIC(0) sweep regions only, one
        call in 16 of each)."""
        self.lib.solver_enable_timing(self._h, (2 if sweeps_only else 1) if on else 0)

    def timing(self) -> dict:
        cap, nl = 64, 48
        names = C.create_string_buffer(cap * nl)
        ms = np.zeros(cap)
        calls = np.zeros(cap, np.uint64)
        cnt = C.c_uint64()
        self.lib.solver_get_timing(self._h, names, nl, A.dptr(ms),
                                   calls.ctypes.data_as(A.c_u64_p), cap, C.byref(cnt))
        out = {}
        for k in range(cnt.value):
            nm = names.raw[k * nl:(k + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[k]), int(calls[k]))
        return out

    def seed_fill(self, per_cell: int, y_lo: float, y_hi: float, seed: int,
                  vel=(0.0, 0.0, 0.0)) -> int:
        n = C.c_uint64()
        self.lib.solver_seed_fill(self._h, per_cell, y_lo, y_hi, seed, C.byref(A.vec3(vel)),
                                  C.byref(n))
        return int(n.value)


def project_capturing_with_source(nx, ny, nz, dx, origin, u, v, w, cells, rho, dt,
                                  params: SolveParams = SolveParams(), source=None, *,
                                  lib: A.Lib | None = None, dot_mode: int = A.DOT_EXACT,
                                  solve_mode: int = A.SOLVE_REFERENCE, refine_tol: float = 1e-8):
    """`pressure::project_capturing_with_source` (src/pressure.rs:123-239).

    u, v, w are updated in place (f64 arrays of the staggered dims); returns
    (SolveReport, pressure[nz, ny, nx])."""
    lib = lib if lib is not None else _lib_default()
    n = nx * ny * nz
    cells = np.ascontiguousarray(cells, np.uint8)
    if cells.size != n:
        raise ValueError("cell classification must cover the whole grid")
    if source is not None:
        source = np.ascontiguousarray(source, np.float64)
        if source.size != n:
            raise ValueError("volume-control source must cover the grid")
    for a, m in ((u, (nx + 1) * ny * nz), (v, nx * (ny + 1) * nz), (w, nx * ny * (nz + 1))):
        if a.dtype != np.float64 or not a.flags.c_contiguous or a.size != m:
            raise ValueError("face arrays must be contiguous f64 of the staggered dims")
    d = A.Dims(nx, ny, nz, dx, A.vec3(origin))
    p = np.zeros((nz, ny, nx), np.float64)
    rep = A.SolveReportC()
    opts = A.Options(A.PRECISION_F64, dot_mode, 0, refine_tol, solve_mode, 0)
    lib.project_capturing_with_source(C.byref(d), A.dptr(u), A.dptr(v), A.dptr(w), A.u8ptr(cells),
                                      rho, dt, C.byref(A.SolveParamsC(int(params.max_iters), params.tol)),
                                      A.dptr(source), A.dptr(p), C.byref(rep), C.byref(opts))
    return SolveReport(int(rep.iters), float(rep.residual)), p
