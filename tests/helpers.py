"""Shared fixtures-as-functions of the test suites (the reference's own test scenes)."""
from __future__ import annotations

import numpy as np

from sdr_amd import scenes
from sdr_amd.solver import SOLID, FLUID


def fill_divergent(nx, ny, nz):
    """`fill_divergent` (src/pressure.rs:645-670) on a grid of nx*ny*nz cells."""
    u = np.zeros((nz, ny, nx + 1))
    v = np.zeros((nz, ny + 1, nx))
    w = np.zeros((nz + 1, ny, nx))
    i = np.arange(nx + 1, dtype=np.float64)[None, None, :]
    j = np.arange(ny, dtype=np.float64)[None, :, None]
    u[:] = (i - 6.0) * 0.05 + np.sin(j * 0.013)
    j = np.arange(ny + 1, dtype=np.float64)[None, :, None]
    k = np.arange(nz, dtype=np.float64)[:, None, None]
    v[:] = (j - 6.0) * 0.05 + np.cos(k * 0.017)
    k = np.arange(nz + 1, dtype=np.float64)[:, None, None]
    w[:] = (k - 6.0) * 0.03 + np.zeros((1, ny, nx))
    return u, v, w


def radial_field(nx, ny, nz):
    """The radially expanding field of `removes_divergence` (src/pressure.rs:675-704)."""
    u = np.zeros((nz, ny, nx + 1))
    v = np.zeros((nz, ny + 1, nx))
    w = np.zeros((nz + 1, ny, nx))
    u[:] = (np.arange(nx + 1, dtype=np.float64)[None, None, :] - 6.0) * 0.05
    v[:] = (np.arange(ny + 1, dtype=np.float64)[None, :, None] - 6.0) * 0.05
    return u, v, w


def classify_box(nx, ny, nz):
    """`classify_box` (src/pressure.rs:755-770): a 1-cell solid shell around fluid."""
    c = np.full((nz, ny, nx), FLUID, np.uint8)
    c[0, :, :] = c[-1, :, :] = SOLID
    c[:, 0, :] = c[:, -1, :] = SOLID
    c[:, :, 0] = c[:, :, -1] = SOLID
    return c


def interior_divergence(u, v, w):
    div = (u[1:-1, 1:-1, 2:-1] - u[1:-1, 1:-1, 1:-2]) + \
          (v[1:-1, 2:-1, 1:-1] - v[1:-1, 1:-2, 1:-1]) + \
          (w[2:-1, 1:-1, 1:-1] - w[1:-2, 1:-1, 1:-1])
    return np.abs(div).max()


def pack_one_cell(solver, cell, n, jitter):
    """`pack_one_cell` (src/solver.rs:1192-1202): n particles at centre + hash_jitter(idx)*dx."""
    c = np.array(solver.cell_center(*cell))
    pos = np.stack([c + jitter(idx) * solver.dx for idx in range(n)])
    solver.set_particles(pos, np.zeros_like(pos))
    return c


def cell_of(solver, p):
    """`cell_of` (src/solver.rs:897-908)."""
    l = (np.asarray(p) - np.asarray(solver.origin)) / solver.dx
    if (l < 0).any():
        return None
    i, j, k = (int(x) for x in l)
    if i < solver.nx and j < solver.ny and k < solver.nz:
        return (i, j, k)
    return None


def counts(solver, pos):
    out = {}
    for p in pos:
        c = cell_of(solver, p)
        if c is not None:
            out[c] = out.get(c, 0) + 1
    return out


def bits_equal(a, b) -> bool:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def rel_close(a, b, tol):
    """|a - b| <= tol * max(|b|) elementwise: the field-scaled relative bound."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(a - b).max()) <= tol * scale, float(np.abs(a - b).max()) / scale


def paired(scene_name: str, orc_lib, dev_kw=None, stride=1):
    """Oracle and device solvers on the same scene and the same seeded fill."""
    import oracle

    sc = scenes.cavity_scene(scene_name)
    o = scenes.make_solver(sc, lib=orc_lib)
    pos = scenes.fill_particles(o, sc, oracle.cavity_mask(o), stride=stride)
    o.set_particles(pos)
    d = None
    if dev_kw is not None:
        d = scenes.make_solver(sc, **dev_kw)
        d.set_particles(pos)
    return sc, o, d


def progress(msg: str) -> None:
    """A progress line on the real stderr (past pytest's capture), for long GPU tests."""
    import sys

    print(f"[test] {msg}", file=sys.__stderr__, flush=True)


def count_source(solver, scene, cells, f32: bool):
    """volume_source (src/solver.rs:392-423) from the solver's particles, with cell_of in
    the solver's own arithmetic (f32 storage puts a particle within rounding distance of
    a cell face where the device puts it).  None when volume control is off."""
    import ctypes as C

    from sdr_amd import _abi as A

    fp = scene.fluid
    if fp.volume_stiffness <= 0.0:
        return None
    n = solver.particle_count()
    pos = np.zeros((n, 3), np.float64)
    solver.lib.solver_get_particles(solver._h, A.dptr(pos), None, n)
    ft = np.float32 if f32 else np.float64
    g = (pos.astype(ft) - np.asarray(solver.origin).astype(ft)) / ft(solver.dx)
    del pos
    ok = np.all(g >= 0, axis=1)
    ijk = g.astype(np.int64)
    del g
    ok &= (ijk[:, 0] < solver.nx) & (ijk[:, 1] < solver.ny) & (ijk[:, 2] < solver.nz)
    cnt = np.bincount(((ijk[ok, 2] * solver.ny + ijk[ok, 1]) * solver.nx + ijk[ok, 0]),
                      minlength=solver.nx * solver.ny * solver.nz).reshape(cells.shape)
    rest = float(max(fp.particles_per_cell, 1))
    unit = solver.dx / scene.dt
    over = cnt / rest - 1.0
    return np.where((cells == FLUID) & (over > 0.0),
                    np.minimum(fp.volume_stiffness * over * unit, 0.25 * unit), 0.0)


def divergence(u, v, w):
    """Per-cell (u+ - u-) + (v+ - v-) + (w+ - w-) (the projection's div, pressure.rs:152-168)."""
    return (u[:, :, 1:] - u[:, :, :-1]) + (v[:, 1:, :] - v[:, :-1, :]) + (w[1:] - w[:-1])
