"""Synthetic cavities of the benchmark and test workloads.

These are build-defined inputs (the reference builds its cavities from meshes or from
`SinusParams::implicit`, src/sinus.rs:89-119): an analytic signed-distance field on a
node grid (`Sdf::from_fn`, src/sdf.rs:90-116), a needle, an outlet box at the socket
opening (as src/scene.rs:299-305 derives it) and an initial jittered fill.  The same
arrays feed the device library and the CPU oracle, so both start from identical bits.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .solver import Aabb, FluidParams, Needle, Sdf, SolveParams

MM = 1e-3
ML_PER_S = 1e-6


def smooth_min(a, b, k):
    """Polynomial smooth minimum (src/sinus.rs:66-74)."""
    h = np.clip(0.5 + 0.5 * (b - a) / k, 0.0, 1.0)
    return (b * (1.0 - h) + a * h) - k * h * (1.0 - h)


def _lattice(a, b, c, seed, o):
    """Hashed lattice values in [-1, 1] at integer points (a, b, c)."""
    h = (a.astype(np.int64) * 73856093) ^ (b.astype(np.int64) * 19349663) ^ \
        (c.astype(np.int64) * 83492791) ^ (seed * 2654435761 + o * 97)
    h = (h * 0x5DEECE66D + 11) & 0xFFFFFFFF
    return (h / 4294967295.0) * 2.0 - 1.0


def value_noise(x, y, z, scale, seed, octaves=1):
    """Seeded smooth value noise in [-1, 1]: per octave, the hashed lattice values
    blended with smoothstep weights (a tensor product, evaluated axis by axis).  The
    build's stand-in for the wall noise of the larger configs.  x, y, z are the (k, j, i)
    coordinate arrays of a grid of nodes (as `Sdf.from_fn` passes them)."""
    xs, ys, zs = x[0, 0, :], y[0, :, 0], z[:, 0, 0]
    out = np.zeros(x.shape)
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        sc = scale / (2 ** o)
        ax = []
        for c in (xs, ys, zs):
            g = c / sc
            i = np.floor(g)
            f = g - i
            ax.append((i.astype(np.int64), f * f * (3 - 2 * f)))
        (ix, ux), (iy, uy), (iz, uz) = ax
        lo = [a.min() for a in (ix, iy, iz)]
        la = [np.arange(l, a.max() + 2) for l, a in zip(lo, (ix, iy, iz))]
        kk, jj, ii = np.meshgrid(la[2], la[1], la[0], indexing="ij")
        L = _lattice(ii, jj, kk, seed, o)                    # (Lz, Ly, Lx)
        a0 = ix - lo[0]
        Lx = L[:, :, a0] * (1 - ux) + L[:, :, a0 + 1] * ux  # (Lz, Ly, nx)
        b0 = iy - lo[1]
        Ly = Lx[:, b0, :] * (1 - uy)[None, :, None] + Lx[:, b0 + 1, :] * uy[None, :, None]
        c0 = iz - lo[2]
        v = Ly[c0] * (1 - uz)[:, None, None] + Ly[c0 + 1] * uz[:, None, None]
        out = out + amp * v
        tot += amp
        amp *= 0.5
    return out / tot


@dataclass
class Scene:
    name: str
    sdf: Sdf
    sim_dx: float
    fluid: FluidParams
    needle: Needle
    outlet: Aabb | None
    solve: SolveParams
    dims: tuple
    fill_lo: float = -1e9         # seed cavity cells with fill_lo <= centre y <= fill_y
    fill_y: float = 1e9
    fill_ppc: int = 8
    fill_seed: int = 1
    dt: float = 5e-5               # fixed substep of the workload


def _cavity_fn(semi, socket_r, blend, ybot_open=True, noise=None, power=2.0, sockets=None):
    """Signed distance of a (super)ellipsoid body joined by a smooth minimum to socket
    channels running down from y = -ay/2 (one at x = z = 0, or one per (x, z) centre of
    `sockets`), plus optional seeded wall noise."""
    ax, ay, az = semi

    def f(x, y, z):
        if power == 2.0:
            body = np.sqrt((x / ax) ** 2 + (y / ay) ** 2 + (z / az) ** 2) - 1.0
        else:
            body = (np.abs(x / ax) ** power + np.abs(y / ay) ** power +
                    np.abs(z / az) ** power) ** (1.0 / power) - 1.0
        body = body * min(semi)
        axial = y - (-0.5 * ay)
        if sockets is None:
            radial = np.sqrt(x * x + z * z) - socket_r
        else:
            radial = None
            for (sx, sz) in sockets:
                rr = np.sqrt((x - sx) ** 2 + (z - sz) ** 2) - socket_r
                radial = rr if radial is None else np.minimum(radial, rr)
        socket = np.maximum(radial, axial)
        phi = smooth_min(body, socket, blend)
        if noise is not None:
            amp, scale, seed, octaves = noise
            phi = phi + amp * value_noise(x, y, z, scale, seed, octaves)
        return phi

    return f


def _socket_outlet(ybot, r, sim_dx):
    rr = r + 1.5 * sim_dx
    return Aabb((-rr, ybot - 2.0 * sim_dx, -rr), (rr, ybot + 1.5 * sim_dx, rr))


def cavity_scene(level: str) -> Scene:
    """The benchmark cavities.

    c1: 15x18x14 mm ellipsoid + 2 mm socket, 1 mm grid 36x42x34, 2 ml/s 0.5 mm jet.
    c2: the same cavity at 0.5 mm (72x84x68), 0.3 mm seeded wall noise, 4 ml/s.
    c3: 256^3 grid, a box-filling rounded cavity (superellipsoid, power 12) + socket,
        filled to ~130M particles at 8 per cell.
    c3a: 256^3 grid, the literal 60x72x56-cell ellipsoid (~8M particles).
    c5: 512x512x384 grid at 0.1 mm, an ellipsoid filling the box with three octaves of
        seeded wall noise and two socket channels, filled to ~400M particles at 8 per
        cell (the large single scenario).
    """
    fluid = FluidParams()
    solve = SolveParams(200, 1e-6)
    if level in ("c1", "c2"):
        dx = 1.0 * MM if level == "c1" else 0.5 * MM
        semi = (15 * MM, 18 * MM, 14 * MM)
        noise = None if level == "c1" else (0.3 * MM, 3.0 * MM, 7, 2)
        origin = (-18 * MM, -21 * MM, -17 * MM)
        dims = (37, 43, 35) if level == "c1" else (73, 85, 69)
        sdf = Sdf.from_fn(origin, dx, dims, _cavity_fn(semi, 2 * MM, 1.5 * MM, noise=noise))
        flow = (2.0 if level == "c1" else 4.0) * ML_PER_S
        needle = Needle(tip=(0.0, -4 * MM, -10 * MM), axis=(0.0, 0.0, 1.0), radius=0.25 * MM,
                        flow_rate=flow, inject_until=math.inf if level == "c1" else 0.25)
        outlet = _socket_outlet(origin[1], 2 * MM, dx)
        grid = (36, 42, 34) if level == "c1" else (72, 84, 68)
        fill_y = 1.0 * MM if level == "c1" else 1e9
        return Scene(level, sdf, dx, fluid, needle, outlet, solve, grid, fill_lo=-17 * MM,
                     fill_y=fill_y, dt=5e-5 if level == "c1" else 2.4e-5)
    if level in ("c3", "c3a"):
        n = 256
        dx = 0.25 * MM
        half = n * dx / 2.0
        origin = (-half, -half, -half)
        if level == "c3":
            semi = (0.495 * n * dx, 0.495 * n * dx, 0.495 * n * dx)
            fn = _cavity_fn(semi, 8 * dx, 4 * dx, power=12.0)
            fill_y = 0.9 * half
        else:
            semi = (60 * dx, 72 * dx, 56 * dx)
            fn = _cavity_fn(semi, 8 * dx, 4 * dx)
            fill_y = 1e9
        sdf = Sdf.from_fn(origin, dx, (n + 1, n + 1, n + 1), fn)
        needle = Needle(tip=(0.0, -0.25 * half, -0.5 * half), axis=(0.0, 0.0, 1.0),
                        radius=2 * dx, flow_rate=4.0 * ML_PER_S)
        outlet = _socket_outlet(origin[1], 8 * dx, dx)
        return Scene(level, sdf, dx, fluid, needle, outlet, solve, (n, n, n),
                     fill_lo=origin[1] + 4 * dx, fill_y=fill_y, dt=2e-5)
    if level == "c5":
        dims = (512, 512, 384)
        dx = 0.1 * MM
        half = tuple(n * dx / 2.0 for n in dims)
        origin = tuple(-h for h in half)
        semi = tuple(0.49 * n * dx for n in dims)
        socket_r = 12 * dx
        off = 0.3 * semi[0]
        fn = _cavity_fn(semi, socket_r, 6 * dx, noise=(3 * dx, 30 * dx, 11, 3),
                        sockets=((-off, 0.0), (off, 0.0)))
        sdf = Sdf.from_fn(origin, dx, tuple(n + 1 for n in dims), fn)
        # a 0.5 mm bore at 4 ml/s: ~5 m/s, half a cell per 1e-5 s substep
        needle = Needle(tip=(0.0, -0.25 * half[1], -0.5 * half[2]), axis=(0.0, 0.0, 1.0),
                        radius=5 * dx, flow_rate=4.0 * ML_PER_S)
        rr = socket_r + 1.5 * dx
        outlet = Aabb((-off - rr, origin[1] - 2.0 * dx, -rr), (off + rr, origin[1] + 1.5 * dx, rr))
        return Scene(level, sdf, dx, fluid, needle, outlet, solve, dims, fill_lo=origin[1] + 4 * dx,
                     fill_y=1e9, dt=1e-5)
    raise ValueError(level)


def box_scene() -> Scene:
    """`box_solver` of the reference tests (src/solver.rs:1170-1189): a 24 mm box cavity
    in a 16^3-node SDF at 2 mm, needle pointing down, 8 particles per cell."""
    dx = 0.002
    a = 0.012
    sdf = Sdf.from_fn((-0.016, -0.016, -0.016), dx, (16, 16, 16),
                      lambda x, y, z: np.maximum(np.maximum(np.abs(x) - a, np.abs(y) - a),
                                                 np.abs(z) - a))
    needle = Needle(tip=(0.0, 0.010, 0.0), axis=(0.0, -1.0, 0.0), radius=0.0006, flow_rate=5.0e-6)
    return Scene("box", sdf, dx, FluidParams(particles_per_cell=8), needle, None, SolveParams(),
                 (0, 0, 0))


def sinus_scene(sim_dx: float = 0.0025) -> Scene:
    """A coarse parametric-sinus scene in the spirit of `small_solver`
    (src/solver.rs:1016-1043): the default `SinusParams` cavity (src/sinus.rs:47-60)
    sampled analytically (`Sdf::from_fn`) instead of through its mesh, 16 particles per
    cell, needle at the socket, outlet at the socket bottom."""
    semi = np.array([0.017, 0.016, 0.013])
    taper, bump, sx, sz, sr, sdepth = 0.6, 0.12, 0.006, 0.0, 0.0025, 0.006

    def implicit(x, y, z):  # src/sinus.rs:89-119
        qx, qy, qz = x, y, z
        h = ((qy / semi[1]) + 1.0) * 0.5
        scale = taper + (1.0 - taper) * np.clip(h, 0.0, 1.5)
        b = bump * (0.6 * np.sin(4.0 * qx / semi[0]) * np.cos(3.0 * qz / semi[2]) +
                    0.4 * np.sin(5.0 * qy / semi[1] + 1.3))
        body = np.sqrt((qx / (semi[0] * scale)) ** 2 + (qy / semi[1]) ** 2 +
                       (qz / (semi[2] * scale)) ** 2) - (1.0 + b)
        body = body * semi.min()
        radial = np.sqrt((x - sx) ** 2 + (z - sz) ** 2) - sr
        ytop = -0.5 * semi[1]
        ybot = -semi[1] - sdepth
        axial = np.maximum(y - ytop, ybot - y)
        return smooth_min(body, np.maximum(radial, axial), 0.0015)

    pad = 0.006
    lo = np.array([-semi[0] * 1.12 - pad, -semi[1] - sdepth - pad, -semi[2] * 1.12 - pad])
    hi = np.array([semi[0] * 1.12 + pad, semi[1] * 1.12 + pad, semi[2] * 1.12 + pad])
    dims = tuple(int(math.ceil((hi[a] - lo[a]) / sim_dx)) + 1 for a in range(3))
    sdf = Sdf.from_fn(tuple(lo), sim_dx, dims, implicit)
    tip = (sx, -0.5 * semi[1] + semi[1] * 0.15, sz)  # src/sinus.rs:154-160
    needle = Needle(tip=tip, axis=(0.0, 1.0, 0.0), radius=0.0006, flow_rate=5.0e-6)
    yb = -semi[1] - sdepth
    outlet = Aabb((sx - 0.004, yb - 0.002, sz - 0.004), (sx + 0.004, yb + 0.002, sz + 0.004))
    return Scene("sinus", sdf, sim_dx, FluidParams(particles_per_cell=16), needle, outlet,
                 SolveParams(), (0, 0, 0))


def make_solver(scene: Scene, lib=None, fill: bool = True, **kw):
    """Build a `Solver` for a scene on `lib` (device library by default), apply its outlet
    and solve params, and seed the initial fill (device fill kernel when available)."""
    from .solver import Solver

    s = Solver(scene.sdf, scene.sim_dx, scene.fluid, scene.needle, lib=lib, **kw)
    if scene.dims != (0, 0, 0):
        assert (s.nx, s.ny, s.nz) == scene.dims, ((s.nx, s.ny, s.nz), scene.dims)
    s.outlet = scene.outlet
    s.solve = scene.solve
    return s


def fill_particles(solver, scene: Scene, mask: np.ndarray, stride: int = 1) -> np.ndarray:
    """Host-side initial fill (identical to the device `seed_fill` kernel): `fill_ppc`
    particles per cavity cell whose centre y is in [fill_lo, fill_y], at centre +
    splitmix64 jitter of (seed, index).  Returns (N,3) positions; `stride` > 1 returns
    only every stride-th particle of that sequence."""
    nz, ny, nx = mask.shape
    kk, jj, ii = np.nonzero(mask)
    order = np.lexsort((ii, jj, kk))
    ii, jj, kk = ii[order], jj[order], kk[order]
    o, dx = solver.origin, solver.dx
    cy = o[1] + (jj + 0.5) * dx
    sel = (cy >= scene.fill_lo) & (cy <= scene.fill_y)
    ii, jj, kk = ii[sel], jj[sel], kk[sel]
    ppc = scene.fill_ppc
    t = np.arange(0, ii.size * ppc, stride, dtype=np.uint64)
    key = np.uint64(scene.fill_seed) * np.uint64(1 << 32) + t

    def mix(k):
        with np.errstate(over="ignore"):
            z = k + np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return ((z >> np.uint64(11)).astype(np.float64) / float(1 << 53) - 0.5) * 0.6

    with np.errstate(over="ignore"):
        k3 = key * np.uint64(3)
    j = np.stack([mix(k3) / 0.6 * 0.9999, mix(k3 + np.uint64(1)) / 0.6 * 0.9999,
                  mix(k3 + np.uint64(2)) / 0.6 * 0.9999], axis=1)
    q = (t // np.uint64(ppc)).astype(np.int64)
    ci = ii[q].astype(np.float64)
    cj = jj[q].astype(np.float64)
    ck = kk[q].astype(np.float64)
    cx = o[0] + (ci + 0.5) * dx
    cyy = o[1] + (cj + 0.5) * dx
    cz = o[2] + (ck + 0.5) * dx
    return np.stack([cx + j[:, 0] * dx, cyy + j[:, 1] * dx, cz + j[:, 2] * dx], axis=1)


def icosphere(radius: float, subdivisions: int, center=(0.0, 0.0, 0.0)):
    """A closed, outward-wound triangle mesh of a sphere: the icosahedron subdivided
    `subdivisions` times (20 * 4^s triangles), vertices pushed to the sphere.  The mesh
    workload of the GPU mesh -> SDF build.  Returns (vertices (nv, 3), tris (nt, 3))."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    vv = np.array(verts) * radius + np.asarray(center, np.float64)
    return vv, np.array(f, np.uint32)


def box_mesh(half: float = 1.0):
    """The box mesh of the reference's winding test (src/sdf.rs:301-322), scaled."""
    v = np.array([(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1),
                  (-1, 1, 1)], np.float64) * half
    f = np.array([(0, 3, 2), (0, 2, 1), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (3, 7, 6), (3, 6, 2),
                  (0, 4, 7), (0, 7, 3), (1, 2, 6), (1, 6, 5)], np.uint32)
    return v, f
