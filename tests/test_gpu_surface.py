"""Device particle surfacing and frame output against the oracle.

Bar: surface nets on the device give the oracle's vertices bit for bit and its
triangles index for index (the same f64 operations per element, the reference's
numbering); area and volume agree to 1e-12 relative (a fixed tree against the
reference's sequential sum).  The particle field is bit-identical to the oracle's for
the same particle positions, in both precisions (distances are formed in f64 from the
stored coordinates).  Surfacing leaves the run unchanged, and the asynchronous frame
writer's files are byte-identical to the synchronous writer's on the same frames."""
import numpy as np
import pytest

from sdr_amd import _abi as A
from sdr_amd import output, scenes
from sdr_amd.surface import surface_nets

from helpers import bits_equal, paired, progress

pytestmark = pytest.mark.gpu
EXACT = {"dot_mode": A.DOT_EXACT}


def _nets_pair(dev, orc, origin, dx, dims, vals):
    a, ia = surface_nets(origin, dx, dims, vals, lib=orc)
    b, ib = surface_nets(origin, dx, dims, vals, lib=dev)
    assert bits_equal(a.vertices, b.vertices)
    assert np.array_equal(a.indices, b.indices)
    assert (ia["vertices"], ia["triangles"]) == (ib["vertices"], ib["triangles"])
    assert ib["area"] == pytest.approx(ia["area"], rel=1e-12, abs=1e-300)
    assert ib["volume"] == pytest.approx(ia["volume"], rel=1e-12, abs=1e-300)
    return b, ib


def test_surface_nets_sphere_matches_oracle(dev, orc):  # surface_nets.rs:228-248
    n, dx = 41, 0.1
    k, j, i = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    p = np.stack([-2.0 + i * dx, -2.0 + j * dx, -2.0 + k * dx], -1)
    vals = np.sqrt(p[..., 0] * p[..., 0] + p[..., 1] * p[..., 1] + p[..., 2] * p[..., 2]) - 1.0
    mesh, info = _nets_pair(dev, orc, (-2.0, -2.0, -2.0), dx, (n, n, n), vals)
    assert info["volume"] > 0 and abs(info["volume"] - 4.0 / 3.0 * np.pi) <= 0.15
    assert abs(info["area"] - 4.0 * np.pi) <= 1.2


@pytest.mark.parametrize("case", ["node", "plane", "noise", "max_faces"])
def test_surface_nets_fields_match_oracle(dev, orc, case):
    if case == "node":
        vals = np.ones((5, 5, 5))
        vals[2, 2, 2] = -1.0
    elif case == "plane":
        vals = np.fromfunction(lambda k, j, i: k - 1.37 + 0.0 * i, (5, 6, 7))
    elif case == "noise":
        k, j, i = np.meshgrid(*(np.arange(48, dtype=np.float64),) * 3, indexing="ij")
        vals = scenes.value_noise(i * 0.05, j * 0.05, k * 0.05, 0.3, 17, octaves=3) - 0.5
    else:
        vals = np.fromfunction(lambda k, j, i: (i + j + k) - 5.5, (4, 4, 4))
    nz, ny, nx = vals.shape
    _nets_pair(dev, orc, (0.25, -0.5, 1.0), 0.125, (nx, ny, nz), np.ascontiguousarray(vals, np.float64))


def _surface_pair(o, d, **kw):
    io, idv = o.surface(**kw), d.surface(**kw)
    for key in ("nx", "ny", "nz", "origin", "spacing", "radius", "vertices", "triangles"):
        assert io[key] == idv[key], key
    assert bits_equal(o.surface_field(), d.surface_field())
    vo, to = o.surface_mesh()
    vd, td = d.surface_mesh()
    assert bits_equal(vo, vd) and np.array_equal(to, td)
    assert idv["area"] == pytest.approx(io["area"], rel=1e-12)
    assert idv["volume"] == pytest.approx(io["volume"], rel=1e-12)
    return idv


@pytest.mark.parametrize("level", ["c1", "c2"])
def test_particle_surface_matches_oracle_f64(dev, orc, level):
    sc, o, d = paired(level, orc, EXACT, stride=1 if level == "c1" else 2)
    for _ in range(3):
        o.substep(sc.dt)
        d.substep(sc.dt)
    progress(f"{level}: surfacing after 3 substeps")
    info = _surface_pair(o, d)
    assert info["triangles"] > 1000 and info["volume"] > 0.0
    _surface_pair(o, d, radius=0.7 * d.dx, spacing=0.5 * d.dx)


def test_particle_surface_matches_oracle_f32(dev, orc):
    """The precision form: the same field and mesh from its f32 positions."""
    sc = scenes.cavity_scene("c1")
    d = scenes.make_solver(sc, precision=A.PRECISION_F32, solve_mode=A.SOLVE_RELATIVE, refine_tol=1e-8)
    import oracle

    o = scenes.make_solver(sc, lib=orc)
    pos = scenes.fill_particles(o, sc, oracle.cavity_mask(o))
    d.set_particles(pos)
    for _ in range(3):
        d.substep(sc.dt)
    pd, vd = d.particles()  # f32 values, exactly representable in f64
    o.set_particles(pd, vd)
    _surface_pair(o, d)


def test_surfacing_leaves_the_run_unchanged(dev, orc):
    sc = scenes.cavity_scene("c1")
    import oracle

    o = scenes.make_solver(sc, lib=orc)
    pos = scenes.fill_particles(o, sc, oracle.cavity_mask(o))
    a = scenes.make_solver(sc, **EXACT)
    b = scenes.make_solver(sc, **EXACT)
    a.set_particles(pos)
    b.set_particles(pos)
    for step in range(12):
        a.substep(sc.dt)
        b.substep(sc.dt)
        if step % 5 == 4:
            info = b.surface()
            assert info["triangles"] > 0
    pa, va = a.particles()
    pb, vb = b.particles()
    assert bits_equal(pa, pb) and bits_equal(va, vb)
    assert bits_equal(a.pressure(), b.pressure())


@pytest.mark.parametrize("fmt", ["vtk", "xyz"])
def test_frame_writer_matches_synchronous_writes(dev, tmp_path, fmt):
    sc = scenes.cavity_scene("c1")
    import oracle
    from sdr_amd.solver import Solver  # noqa: F401

    orc_lib = oracle.load()
    o = scenes.make_solver(sc, lib=orc_lib)
    pos = scenes.fill_particles(o, sc, oracle.cavity_mask(o))
    del o
    d = scenes.make_solver(sc, precision=A.PRECISION_F32, solve_mode=A.SOLVE_RELATIVE, refine_tol=1e-8)
    d.set_particles(pos)
    (tmp_path / "async").mkdir()
    with output.FrameWriter(tmp_path / "async", "particles", fmt, slots=2, lib=dev) as w:
        for f in range(4):
            w.submit(d, f)
            p, _ = d.particles()
            ref = output.frame_path(tmp_path, "ref", f, fmt, lib=dev)
            (output.write_points_vtk if fmt == "vtk" else output.write_points_xyz)(ref, p, lib=dev)
            d.substep(sc.dt)
            d.substep(sc.dt)
        assert w.flush() == 4
    for f in range(4):
        got = open(output.frame_path(tmp_path / "async", "particles", f, fmt, lib=dev), "rb").read()
        want = open(output.frame_path(tmp_path, "ref", f, fmt, lib=dev), "rb").read()
        assert got == want, f
    if fmt == "vtk":
        back = output.read_points_vtk(output.frame_path(tmp_path / "async", "particles", 3, fmt, lib=dev), lib=dev)
        assert back.shape[0] > 1000
