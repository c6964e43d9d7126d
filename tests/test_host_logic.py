"""Host-side logic: the Python mirror of the reference types, scene construction and the
seeded fill."""
import math

import numpy as np

import oracle
from sdr_amd import scenes
from sdr_amd.solver import Aabb, FluidParams, Needle, SolveParams


def test_reference_defaults():
    f = FluidParams()  # solver.rs:73-85
    assert (f.density, f.gravity, f.flip_ratio, f.particles_per_cell) == (1000.0, (0.0, -9.81, 0.0), 0.95, 8)
    assert (f.volume_stiffness, f.redistribute, f.redist_cap_factor) == (1.0, True, 1.0)
    s = SolveParams()  # pressure.rs:43-50
    assert (s.max_iters, s.tol) == (200, 1e-6)


def test_aabb_contains_is_inclusive():
    b = Aabb((0, 0, 0), (1, 1, 1))
    assert b.contains((1.0, 0.0, 1.0)) and not b.contains((1.0000001, 0, 0))


def test_needle_jet_speed_zero_radius():
    assert Needle((0, 0, 0), (0, 1, 0), 0.0, 1e-6).jet_speed() == 0.0


def test_fill_is_deterministic_and_inside_cells():
    sc = scenes.cavity_scene("c1")
    s = scenes.make_solver(sc, lib=oracle.load())
    m = oracle.cavity_mask(s)
    a = scenes.fill_particles(s, sc, m)
    b = scenes.fill_particles(s, sc, m)
    assert np.array_equal(a, b)
    l = (a - np.array(s.origin)) / s.dx
    frac = l - np.floor(l)
    assert (frac > 0.0).all() and (frac < 1.0).all()
    assert a.shape[0] % sc.fill_ppc == 0
    assert np.array_equal(scenes.fill_particles(s, sc, m, stride=3), a[::3])


def test_workload_scenes_build():
    for lvl, dims in (("c1", (36, 42, 34)), ("c2", (72, 84, 68))):
        sc = scenes.cavity_scene(lvl)
        s = scenes.make_solver(sc, lib=oracle.load())
        assert (s.nx, s.ny, s.nz) == dims
        assert sc.outlet.contains((0.0, s.origin[1], 0.0))


def test_sdf_slabs_and_socket_channels():
    """Sdf.from_fn evaluates slab by slab with the same bits as one evaluation over the
    whole node grid; the two-socket cavity of the large scenario with one socket at the
    origin is the single-socket cavity, and each of two channels reaches the floor."""
    from sdr_amd.solver import Sdf
    dx = 0.1 * scenes.MM
    dims = (70, 60, 31)
    origin = (-3.5 * dx * 10, -3.0 * dx * 10, -15.0 * dx * 10)
    semi = (3.2e-3, 2.7e-3, 13e-3)
    noise = (3 * dx, 30 * dx, 11, 3)
    one = scenes._cavity_fn(semi, 0.6e-3, 6 * dx, noise=noise)
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nx, dtype=np.float64), indexing="ij")
    whole = one(origin[0] + i * dx, origin[1] + j * dx, origin[2] + k * dx).reshape(-1)
    for slab_nodes in (1 << 22, 4200 * 4):  # one slab; slabs of 4 planes, the last of 3
        sd = Sdf.from_fn(origin, dx, dims, one, slab_nodes=slab_nodes)
        assert np.array_equal(sd.data, whole)
    same = Sdf.from_fn(origin, dx, dims, scenes._cavity_fn(semi, 0.6e-3, 6 * dx, noise=noise,
                                                           sockets=((0.0, 0.0),)))
    assert np.array_equal(same.data, whole)
    two = scenes._cavity_fn(semi, 0.6e-3, 6 * dx, sockets=((-1.5e-3, 0.0), (1.5e-3, 0.0)))
    ybot = origin[1]
    for sx in (-1.5e-3, 1.5e-3):
        assert two(np.array([sx]), np.array([ybot]), np.array([0.0]))[0] < 0.0
    assert two(np.array([0.0]), np.array([ybot]), np.array([0.0]))[0] > 0.0
