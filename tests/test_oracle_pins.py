"""Independent pins of the oracle stages that no reference test covers.

Each stage of the CPU oracle is checked by a route other than its own formula: numpy
brute force written from the reference's rules (classify, extrapolation masks,
diagnostics, metrics wall cells), closed forms (volume source through an explicit
source, single-particle P2G face sets and the normalize threshold, constant-field
P2G), invariants of the CFL rule, and the published xorshift stream for emission.
A plausible slip in the oracle (a dropped term, a wrong half-offset, index or
threshold) fails one of them.
"""
import math

import numpy as np
import pytest

import oracle
from sdr_amd import scenes
from sdr_amd.metrics import MetricsCollector
from sdr_amd.solver import AIR, FLUID, SOLID, Aabb, Needle, SolveParams, project_capturing_with_source

from helpers import bits_equal
from test_oracle_math import numpy_extrapolate


def _c1(orc, stride=1):
    sc = scenes.cavity_scene("c1")
    o = scenes.make_solver(sc, lib=orc)
    pos = scenes.fill_particles(o, sc, oracle.cavity_mask(o), stride=stride)
    return sc, o, pos


def _cell_of(o, pos):
    """cell_of (src/solver.rs:897-908): truncating index after the >= 0 checks."""
    l = (np.asarray(pos) - np.asarray(o.origin)) / o.dx
    ok = np.all(l >= 0.0, axis=1)
    ijk = np.where(ok[:, None], l, 0.0).astype(np.int64)
    ok &= (ijk[:, 0] < o.nx) & (ijk[:, 1] < o.ny) & (ijk[:, 2] < o.nz)
    c = (ijk[:, 2] * o.ny + ijk[:, 1]) * o.nx + ijk[:, 0]
    return np.where(ok, c, -1)


def _numpy_mask(o, sc):
    """Cavity mask (src/solver.rs:194-199): the SDF sampled trilinearly at each cell
    centre is negative.  Returns (mask, |phi|) so borderline cells can be excluded."""
    sd = sc.sdf
    data = np.asarray(sd.data).reshape(sd.nz, sd.ny, sd.nx)
    k, j, i = np.meshgrid(np.arange(o.nz), np.arange(o.ny), np.arange(o.nx), indexing="ij")
    p = [np.asarray(o.origin)[a] + (idx + 0.5) * o.dx for a, idx in enumerate((i, j, k))]
    g = [(p[a] - sd.origin[a]) / sd.dx for a in range(3)]
    n = (sd.nx, sd.ny, sd.nz)
    b, f = [], []
    for a in range(3):
        fl = np.clip(np.floor(g[a]), 0, n[a] - 2).astype(np.int64)
        b.append(fl)
        f.append(np.clip(g[a] - fl, 0.0, 1.0))
    phi = 0.0
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                w = (f[0] if di else 1 - f[0]) * (f[1] if dj else 1 - f[1]) * (f[2] if dk else 1 - f[2])
                phi = phi + w * data[b[2] + dk, b[1] + dj, b[0] + di]
    return phi < 0.0, np.abs(phi)


# ---------------------------------------------------------------- classify (solver.rs:342-371)
def test_classify_matches_numpy_brute_force(orc):
    sc, o, pos = _c1(orc, stride=3)
    ob = sc.outlet
    # a few particles placed in the outlet box and in solid cells as well
    extra = np.array([[0.5 * (ob.min[0] + ob.max[0]), o.origin[1] + 0.7 * o.dx, 0.5 * (ob.min[2] + ob.max[2])],
                      [o.origin[0] + 0.5 * o.dx, o.origin[1] + 0.5 * o.dx, o.origin[2] + 0.5 * o.dx],
                      [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    pos = np.concatenate([pos, extra])
    o.set_particles(pos)
    o.run_stage("classify")
    got = o.cells()
    mask, aphi = _numpy_mask(o, sc)
    assert np.array_equal(mask[aphi > 1e-9], oracle.cavity_mask(o).astype(bool)[aphi > 1e-9])
    mask = oracle.cavity_mask(o).astype(bool)
    k, j, i = np.meshgrid(np.arange(o.nz), np.arange(o.ny), np.arange(o.nx), indexing="ij")
    cc = [np.asarray(o.origin)[a] + (idx + 0.5) * o.dx for a, idx in enumerate((i, j, k))]
    in_box = lambda x, y, z: ((x >= ob.min[0]) & (x <= ob.max[0]) & (y >= ob.min[1]) & (y <= ob.max[1]) &
                              (z >= ob.min[2]) & (z <= ob.max[2]))
    want = np.where(in_box(*cc), AIR, np.where(~mask, SOLID, AIR)).astype(np.uint8).reshape(-1)
    c = _cell_of(o, pos)
    for t in range(pos.shape[0]):
        if c[t] >= 0 and want[c[t]] == AIR and not in_box(*pos[t]):
            want[c[t]] = FLUID
    assert np.array_equal(got.reshape(-1), want)
    assert (want == FLUID).sum() > 1000 and (want == SOLID).sum() > 1000
    # an outlet cell stays Air although a particle sits in it
    assert c[-4] >= 0 and got.reshape(-1)[c[-4]] == AIR


# ---------------------------------------------------------------- volume_source (solver.rs:392-423)
def test_volume_source_closed_form(orc):
    """The project stage's internal source equals min(k (cnt/ppc - 1) dx/dt, 0.25 dx/dt)
    on over-full Fluid cells: passing that closed form explicitly to the projection gives
    the same pressure bits.  Clumps make some cells far over the cap (the 0.25 clamp)."""
    sc, o, pos = _c1(orc, stride=2)
    g = np.random.default_rng(3)
    centres = pos[g.choice(pos.shape[0], 40, replace=False)]
    clumps = np.repeat(centres, 30, axis=0) + g.uniform(-0.3, 0.3, (1200, 3)) * o.dx
    pos = np.concatenate([pos, clumps])
    o.set_particles(pos)
    dt = sc.dt
    for st in ("classify", "p2g", "save", "forces"):
        o.run_stage(st, dt)
    u, v, w = o.grid()
    cells = o.cells()
    c = _cell_of(o, pos)
    fl = cells.reshape(-1) == FLUID
    cnt = np.bincount(c[(c >= 0)], minlength=cells.size)
    cnt = np.where(fl, cnt, 0)
    fp = sc.fluid
    rest = float(max(fp.particles_per_cell, 1))
    unit = o.dx / dt
    over = cnt / rest - 1.0
    src = np.where(fl & (over > 0.0), np.minimum(fp.volume_stiffness * over * unit, 0.25 * unit), 0.0)
    assert (src == 0.25 * unit).sum() > 5 and ((src > 0) & (src < 0.25 * unit)).sum() > 5
    uo, vo, wo = u.copy(), v.copy(), w.copy()
    rep, p_ref = project_capturing_with_source(o.nx, o.ny, o.nz, o.dx, o.origin, uo, vo, wo, cells,
                                               fp.density, dt, sc.solve, src.reshape(cells.shape), lib=orc)
    o.run_stage("project", dt)
    assert o.last_pcg_iters == rep.iters
    assert bits_equal(o.pressure(), p_ref)
    for a, b in zip(o.grid(), (uo, vo, wo)):
        assert bits_equal(a, b)


# ---------------------------------------------------------------- p2g (solver.rs:560-585, 944-952)
def _faces_of(o, p, comp):
    """The 8 faces (z, y, x index of the comp's face array) a particle at p scatters
    to, with their weights: half-offset 0 along the component axis, 0.5 across, then
    the clamped base of grid.rs:26-35."""
    shape = o.face_shapes()[comp]
    dims = (shape[2], shape[1], shape[0])
    l = (np.asarray(p) - np.asarray(o.origin)) / o.dx
    g = [l[a] - (0.0 if a == comp else 0.5) for a in range(3)]
    base, frac = [], []
    for a in range(3):
        if g[a] <= 0.0:
            b, f = 0, 0.0
        elif g[a] >= dims[a] - 1:
            b, f = dims[a] - 2, 1.0
        else:
            b = int(math.floor(g[a]))
            f = g[a] - b
        base.append(b)
        frac.append(f)
    out = {}
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                w = ((frac[0] if di else 1 - frac[0]) * (frac[1] if dj else 1 - frac[1])) * \
                    (frac[2] if dk else 1 - frac[2])
                out[(base[2] + dk, base[1] + dj, base[0] + di)] = w
    return out


def test_p2g_single_particle_face_sets_per_component(orc):
    """One particle: exactly the faces with weight > 1e-12 get its velocity component
    (the normalisation w v / w), every other face is 0.  Particles in the cavity interior,
    against the low boundary (clamped base, fraction 0) and at a face-aligned position
    (fractions 0 / 0.5 hit the threshold rule)."""
    sc, o, _ = _c1(orc)
    o.set_particles(np.zeros((1, 3)))
    pts = [(0.0011, -0.0023, 0.0037),
           (o.origin[0] + 0.2 * o.dx, o.origin[1] + 0.3 * o.dx, o.origin[2] + 0.6 * o.dx),
           (o.origin[0] + 7.0 * o.dx, o.origin[1] + 9.5 * o.dx, o.origin[2] + 11.25 * o.dx)]
    for p in pts:
        vel = np.array([[1.5, -2.25, 0.75]])
        o.set_particles(np.array([p]), vel)
        o.run_stage("classify")
        o.run_stage("p2g")
        for comp, arr in enumerate(o.grid()):
            want = np.zeros(arr.shape)
            for idx, w in _faces_of(o, p, comp).items():
                if w > 1e-12:
                    want[idx] = vel[0, comp]
            assert np.array_equal(np.abs(arr - want) <= 1e-15 * abs(vel[0, comp]), np.ones(arr.shape, bool)), \
                (p, comp)
            assert (arr != 0).sum() == sum(1 for w in _faces_of(o, p, comp).values() if w > 1e-12)


def test_p2g_normalize_threshold_is_1e_12(orc):
    """normalize (solver.rs:944-952) keeps a face whose weight sum exceeds 1e-12 and zeroes
    it otherwise: a particle whose fraction along x puts 1e-13 or 1e-11 of its weight on
    the far u-face row."""
    sc, o, _ = _c1(orc)
    x0 = o.origin[0] + 10.0 * o.dx
    base = (o.origin[1] + 10.5 * o.dx, o.origin[2] + 10.5 * o.dx)  # across fractions 0 -> weight 1 on one row
    for eps, kept in ((1e-13, False), (1e-11, True)):
        p = np.array([[x0 + eps * o.dx, base[0], base[1]]])
        o.set_particles(p, np.array([[3.0, 0.0, 0.0]]))
        o.run_stage("classify")
        o.run_stage("p2g")
        u = o.grid()[0]
        i = int(round((x0 - o.origin[0]) / o.dx))
        assert u[10, 10, i] == 3.0
        far = u[10, 10, i + 1]
        l = (p[0, 0] - o.origin[0]) / o.dx
        w_far = l - math.floor(l)
        assert (w_far > 1e-12) == kept
        assert (far != 0.0) == kept, (eps, far, w_far)


def test_p2g_constant_field_is_reproduced(orc):
    """grid.rs:49-51: splatting a constant field and normalising gives the constant back
    on every face with weight, 0 elsewhere."""
    sc, o, pos = _c1(orc, stride=5)
    vel = np.tile([[0.7, -1.3, 2.1]], (pos.shape[0], 1))
    o.set_particles(pos, vel)
    o.run_stage("classify")
    o.run_stage("p2g")
    for comp, arr in enumerate(o.grid()):
        nz = arr != 0
        assert nz.sum() > 1000
        assert np.allclose(arr[nz], vel[0, comp], rtol=1e-14, atol=0)


# ---------------------------------------------------------------- extrapolation masks (solver.rs:608-647)
def test_extrapolation_valid_masks_from_numpy(orc):
    """valid = the face borders a Fluid cell; the stage equals the numpy flood fill run
    on masks built here from the cell classification."""
    sc, o, pos = _c1(orc, stride=2)
    o.set_particles(pos)
    o.run_stage("classify")
    cells = o.cells()
    g = np.random.default_rng(4)
    faces = [g.standard_normal(sh) for sh in o.face_shapes()]
    o.set_grid(*faces)
    o.run_stage("extrapolate")
    fl = cells == FLUID
    pad = lambda a, ax: np.concatenate([np.zeros_like(a.take([0], axis=ax)), a,
                                        np.zeros_like(a.take([0], axis=ax))], axis=ax)
    masks = []
    for ax in (2, 1, 0):
        f = pad(fl, ax)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        masks.append((f[tuple(lo)] | f[tuple(hi)]).astype(np.uint8))
    for got, f0, m in zip(o.grid(), faces, masks):
        assert m.sum() > 1000
        want = numpy_extrapolate(f0, m, 4)
        assert bits_equal(got, want)


# ---------------------------------------------------------------- add_forces (solver.rs:593-604)
def test_add_forces_adds_gravity_dt_per_nonzero_component(orc):
    sc, o, _ = _c1(orc)
    g = np.random.default_rng(6)
    faces = [g.standard_normal(sh) for sh in o.face_shapes()]
    faces[0][0, 0, 0] = -0.0  # a zero component leaves the array untouched, sign of zero included
    o.set_grid(*faces)
    dt = 3.7e-5
    o.run_stage("forces", dt)
    gx, gy, gz = sc.fluid.gravity
    assert gx == 0.0 and gz == 0.0 and gy != 0.0
    u, v, w = o.grid()
    assert bits_equal(u, faces[0]) and bits_equal(w, faces[2])
    assert bits_equal(v, faces[1] + gy * dt)


# ---------------------------------------------------------------- cfl_dt (solver.rs:262-281, grid.rs:202-206)
def test_cfl_dt_properties(orc):
    sc = scenes.cavity_scene("c1")
    sc.fluid.gravity = (0.0, 0.0, 0.0)
    sc.needle = Needle(sc.needle.tip, sc.needle.axis, sc.needle.radius, 2e-6, inject_until=0.0)
    o = scenes.make_solver(sc, lib=orc)
    shapes = o.face_shapes()
    zero = [np.zeros(sh) for sh in shapes]
    # empty, still, no gravity, no jet: the 1e-6 speed floor, then the 0.01 cap
    o.set_grid(*zero)
    assert o.cfl_dt() == 0.01
    # the bound is per component: u's largest |u| and v's largest |v| on different faces
    f = [z.copy() for z in zero]
    f[0][3, 4, 5] = -30.0
    f[1][7, 8, 9] = 40.0
    o.set_grid(*f)
    dt = o.cfl_dt()
    assert abs(dt * 50.0 - o.dx) <= 1e-16 * o.dx * 4
    # homogeneous of degree -1 in the field
    o.set_grid(*[a * 2.0 for a in f])
    assert abs(o.cfl_dt() * 2.0 - dt) <= 4e-16 * dt
    # gravity's sqrt(|g| dx) dominates a slow field (a strong g: at 9.81 the 0.01 cap binds)
    sc.fluid.gravity = (0.0, -100.0, 0.0)
    og = scenes.make_solver(sc, lib=orc)
    og.set_grid(*zero)
    assert abs(og.cfl_dt() - og.dx / math.sqrt(100.0 * og.dx)) <= 4e-16
    sc.fluid.gravity = (0.0, -9.81, 0.0)
    og = scenes.make_solver(sc, lib=orc)
    og.set_grid(*zero)
    assert og.cfl_dt() == 0.01
    # the jet counts only while injecting (time < inject_until)
    sc.fluid.gravity = (0.0, 0.0, 0.0)
    sc.needle = Needle(sc.needle.tip, sc.needle.axis, 0.25e-3, 2e-6, inject_until=1.0)
    oj = scenes.make_solver(sc, lib=orc)
    oj.set_grid(*zero)
    v_jet = 2e-6 / (math.pi * 0.25e-3 ** 2)
    assert abs(oj.cfl_dt() - oj.dx / v_jet) <= 4e-16 * oj.dx / v_jet


# ---------------------------------------------------------------- emit (solver.rs:313-339)
def _xorshift_units(state, n):
    out = []
    m = (1 << 64) - 1
    for _ in range(n):
        x = state
        x ^= (x << 13) & m
        x ^= x >> 7
        x ^= (x << 17) & m
        state = x
        out.append((x >> 11) / float(1 << 53))
    return out, state


def test_emit_positions_from_the_published_stream(orc):
    """The emitted particles follow from the xorshift64 stream (3 draws each: radius,
    angle, axial jitter) and an orthonormal basis of the axis; numpy's cos/sin agree with
    glibc's to an ulp, so positions agree to a few ulps of the tip coordinates."""
    sc = scenes.cavity_scene("c1")
    axis = np.array([0.3, -0.2, 0.9])
    sc.needle = Needle(sc.needle.tip, tuple(axis), 0.25e-3, 2e-6)
    o = scenes.make_solver(sc, lib=orc)
    o.set_particles(np.zeros((0, 3)))
    dt = 5e-4  # 8 particles of 1.25e-10 m^3 at 2 ml/s
    st0 = o.state().rng
    o.run_stage("emit", dt)
    pos, vel = o.particles()
    pv = o.dx ** 3 / 8
    n = math.floor(2e-6 * dt / pv)
    assert pos.shape[0] == n == 8
    us, st1 = _xorshift_units(st0, 3 * n)
    assert o.state().rng == st1
    a = axis / np.linalg.norm(axis)
    ref = np.array([1.0, 0.0, 0.0]) if abs(a[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(a, ref)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(a, e1)
    e2 /= np.linalg.norm(e2)
    tip = np.array(sc.needle.tip)
    for t in range(n):
        r = 0.25e-3 * math.sqrt(us[3 * t])
        th = 2.0 * math.pi * us[3 * t + 1]
        along = (us[3 * t + 2] - 0.5) * o.dx
        want = tip + e1 * (r * math.cos(th)) + e2 * (r * math.sin(th)) + a * along
        assert np.all(np.abs(pos[t] - want) <= 8 * np.spacing(np.abs(tip) + 1e-3)), (t, pos[t] - want)
    speed = 2e-6 / (math.pi * 0.25e-3 ** 2)
    assert np.allclose(vel, a * speed, rtol=1e-15, atol=0)


# ---------------------------------------------------------------- diagnostics (solver.rs:225-233, 758-891)
def test_diagnostics_match_numpy(orc):
    sc, o, pos = _c1(orc, stride=1)
    o.set_particles(pos)
    for _ in range(2):
        o.substep(sc.dt)
    d = o.diagnostics()
    pos, vel = o.particles()
    cells = o.cells().reshape(-1)
    p = o.pressure().reshape(-1)
    c = _cell_of(o, pos)
    cnt = np.bincount(c[c >= 0], minlength=cells.size)
    pv = o.dx * o.dx * o.dx / 8.0
    assert d["particle_volume"] == pv
    assert d["fluid_volume"] == pos.shape[0] * pv
    assert d["drained_volume"] == o.state().drained * pv
    assert d["cells_fluid_count"] == int((cells == FLUID).sum())
    assert d["max_particles_per_cell"] == int(cnt.max())
    assert d["max_pressure"] == max(0.0, float(p[cells == FLUID].max()))
    assert math.isclose(d["mean_pressure"], float(p[cells == FLUID].mean()), rel_tol=1e-12)
    assert d["max_speed"] == float(np.sqrt((vel * vel).sum(1)).max())
    assert math.isclose(d["kinetic_energy"], 0.5 * float((vel * vel).sum()), rel_tol=1e-12)


# ---------------------------------------------------------------- metrics (metrics.rs:119-479)
def test_membrane_stats_known_answers():
    """metrics.rs:537-551."""
    import ctypes as C

    dll = oracle.load().dll

    def ms(vals, cap):
        a = np.ascontiguousarray(vals, np.float64)
        mean, peak, art = C.c_double(), C.c_double(), C.c_uint64()
        dll.orc_membrane_stats(a.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(a.size), C.c_double(cap),
                               C.byref(mean), C.byref(peak), C.byref(art))
        return mean.value, peak.value, art.value

    mean, peak, art = ms([100.0, 200.0, 500.0, 900.0, 2_000_000.0], 1000.0)
    assert art == 1 and peak == 900.0 and abs(mean - (100.0 + 200.0 + 500.0 + 900.0) / 4.0) < 1e-9
    assert ms([10.0, 950.0], 1000.0) == (480.0, 950.0, 0)
    assert ms([5000.0], 1000.0) == (0.0, 0.0, 1)
    assert ms(np.zeros(0), 1000.0) == (0.0, 0.0, 0)


def test_metrics_wall_cells_and_frames_match_numpy(orc):
    """Wall cells (metrics.rs:140-182) by numpy; coverage, membrane load and the summary
    relations of metrics.rs:487-534 over several frames; the integer mass balance."""
    sc, o, pos = _c1(orc, stride=1)
    o.set_particles(pos)
    cavity_ml = float(oracle.cavity_mask(o).sum()) * o.dx ** 3
    m = MetricsCollector(o, cavity_ml)
    inside = oracle.cavity_mask(o).astype(bool)
    pad = np.pad(inside, 1, constant_values=False)
    nb_out = np.zeros_like(inside)
    for ax in range(3):
        for sh in (-1, 1):
            nb_out |= ~np.roll(pad, sh, axis=ax)[1:-1, 1:-1, 1:-1]
    wall = inside & nb_out
    assert m.wall_count == int(wall.sum())
    washed_want = np.zeros(wall.size, bool)
    v_jet = sc.needle.jet_speed()
    cap = 3.0 * (0.5 * sc.fluid.density * v_jet * v_jet) + \
        sc.fluid.density * float(np.linalg.norm(sc.fluid.gravity)) * (o.ny * o.dx) + 500.0
    for f in range(4):
        o.substep(sc.dt)
        r = m.record(o, f)
        pos, _ = o.particles()
        c = _cell_of(o, pos)
        wet = np.zeros(wall.size, bool)
        wet[c[c >= 0]] = True
        wet &= wall.reshape(-1)
        washed_want |= wet
        p = o.pressure().reshape(-1)[wet]
        acc = p[p <= cap]
        assert r["wet_wall_cells"] == int(wet.sum())
        assert r["membrane_artifacts"] == int((p > cap).sum())
        assert r["wall_coverage"] == washed_want.sum() / wall.sum()
        assert r["peak_wall_pressure_pa"] == (max(0.0, float(acc.max())) if acc.size else 0.0)
        assert math.isclose(r["wall_pressure_pa"], float(acc.mean()) if acc.size else 0.0, rel_tol=1e-12)
        assert 0.0 <= r["wall_coverage"] <= 1.0
        assert r["peak_wall_pressure_pa"] <= r["max_pressure_pa"] + 1e-6
        assert r["wall_pressure_pa"] <= r["peak_wall_pressure_pa"] + 1e-6
        assert r["fill_fraction"] == pos.shape[0] * o.dx ** 3 / 8.0 / cavity_ml
    washed, iw = m.cells()
    assert np.array_equal(iw.reshape(-1).astype(bool), wall.reshape(-1))
    assert np.array_equal(washed.reshape(-1).astype(bool), washed_want)
    s = m.summary()
    assert s["frames"] == 4
    assert s["peak_membrane_pressure_pa"] <= s["peak_pressure_pa"] + 1e-6
    assert s["mean_membrane_pressure_pa"] <= s["peak_membrane_pressure_pa"] + 1e-6
    assert s["emitted"] == s["resident"] + s["drained"] + s["culled"]
    assert s["mass_balance"] == 1.0
    assert s["peak_wall_coverage"] == max(h["wall_coverage"] for h in m.history)
