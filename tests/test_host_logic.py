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
