"""Device parity: the gfx950 path against the CPU oracle on identical seeded inputs.

Bar (the parity contract of this build):
  * bit-exact: cell classification, particle order and counts (sort / compaction /
    drained / culled), and — in exact mode (reference-order P2G merge, 4096-chunk dot
    tree) — every float of every stage of the f64 substep;
  * fast mode (cell-ordered P2G sums, tree dot): fields within 1e-5 relative to the
    field's max magnitude after one substep from the same state, integers exact.
"""
import numpy as np
import pytest

import oracle
import sdr_amd
from sdr_amd import _abi as A
from sdr_amd import scenes
from sdr_amd.solver import FLUID, Needle, SolveParams, project_capturing_with_source

from helpers import (bits_equal, cell_of, classify_box, counts, fill_divergent,
                     interior_divergence, pack_one_cell, paired, radial_field, rel_close)

pytestmark = pytest.mark.gpu
EXACT = {"dot_mode": A.DOT_EXACT}
FAST = {"dot_mode": A.DOT_FAST}
FIELD_TOL = 1e-5


def test_device_is_gfx950(dev):
    n, arch = sdr_amd.device_info()
    assert n >= 1 and arch.startswith("gfx950")


def _compare_all(o, d, exact=True):
    po, vo = o.particles()
    pd, vd = d.particles()
    assert po.shape == pd.shape
    assert np.array_equal(o.cells(), d.cells())
    so, sd = o.state(), d.state()
    assert (so.drained, so.culled, so.emitted) == (sd.drained, sd.culled, sd.emitted)
    if exact:
        assert bits_equal(po, pd) and bits_equal(vo, vd)
        for a, b in zip(o.grid(), d.grid()):
            assert bits_equal(a, b)
        assert bits_equal(o.pressure(), d.pressure())
        assert so.last_pcg_iters == sd.last_pcg_iters


@pytest.mark.parametrize("level", ["c1", "c2"])
def test_substep_stages_bit_exact(dev, orc, level):
    """Every stage of two substeps, stage by stage, bit for bit (exact mode)."""
    sc, o, d = paired(level, orc, EXACT, stride=1 if level == "c1" else 2)
    for sub in range(2):
        for st in A.STAGES:
            o.run_stage(st, sc.dt)
            d.run_stage(st, sc.dt)
            po, vo = o.particles()
            pd, vd = d.particles()
            assert bits_equal(po, pd), (sub, st, "positions")
            assert bits_equal(vo, vd), (sub, st, "velocities")
            if st in ("classify", "p2g", "forces", "project", "extrapolate"):
                assert np.array_equal(o.cells(), d.cells()), (sub, st)
                for a, b in zip(o.grid(), d.grid()):
                    assert bits_equal(a, b), (sub, st, "grid")
            if st == "project":
                assert bits_equal(o.pressure(), d.pressure()), (sub, st)
                assert o.last_pcg_iters == d.last_pcg_iters
        so, sd = o.state(), d.state()
        assert (so.drained, so.culled, so.emitted) == (sd.drained, sd.culled, sd.emitted)


def test_multi_substep_run_bit_exact(dev, orc):
    sc, o, d = paired("c1", orc, EXACT)
    for _ in range(10):
        o.substep(sc.dt)
        d.substep(sc.dt)
    _compare_all(o, d, exact=True)


def test_timing_modes(dev, orc):
    """Timing mode 1 brackets every region with events; mode 2 only the MIC(0) sweep
    regions, one call in 16 of each (the 16th, the 32nd, ...), on the same substep."""
    sc, o, d1 = paired("c1", orc, FAST)
    d2 = scenes.make_solver(sc, **FAST)
    d2.set_particles(o.particles()[0])
    d1.enable_timing(True)
    d1.substep(sc.dt)
    t1 = d1.timing()
    d2.enable_timing(True, sweeps_only=True)
    d2.substep(sc.dt)
    t2 = d2.timing()
    assert d1.last_pcg_iters == d2.last_pcg_iters > 16
    assert any(not k.startswith("mic_") for k in t1)
    sampled = {k for k in t1 if k.startswith("mic_") and t1[k][1] >= 16}
    assert set(t2) == sampled and "mic_bwd_dot" in t2
    for k, (ms, calls) in t2.items():
        assert calls == t1[k][1] // 16 and ms > 0.0, (k, calls, t1[k])


def test_cfl_step_bit_exact(dev, orc):
    """Solver::step's CFL sub-stepping (solver.rs:237-281) picks the same substeps."""
    sc, o, d = paired("c1", orc, EXACT, stride=2)
    o.step(0.0004)
    d.step(0.0004)
    assert o.last_substeps == d.last_substeps
    assert o.time == d.time
    _compare_all(o, d, exact=True)


def test_fast_mode_within_tolerance(dev, orc):
    """Fast mode from the same state: integers exact, floats within 1e-5 relative."""
    sc, o, d = paired("c1", orc, FAST)
    o.substep(sc.dt)
    d.substep(sc.dt)
    _compare_all(o, d, exact=False)
    po, vo = o.particles()
    pd, vd = d.particles()
    assert rel_close(pd, po, FIELD_TOL)[0]
    assert rel_close(vd, vo, FIELD_TOL)[0], rel_close(vd, vo, FIELD_TOL)
    assert rel_close(d.pressure(), o.pressure(), FIELD_TOL)[0]
    for a, b in zip(d.grid(), o.grid()):
        ok, err = rel_close(a, b, FIELD_TOL)
        assert ok, err


@pytest.mark.parametrize("precision", [A.PRECISION_F32, A.PRECISION_F64])
def test_p2g_fused_matches_face_gather(dev, precision, monkeypatch):
    """The one-pass LDS-staged P2G (slabs staged once for all three components, several
    z-levels accumulated at a time) sums every face in the same order as the
    face-by-face gather, so the two agree bit for bit; a 5000-particle clump makes some
    slabs overflow the staging area (several passes).  Heavy cells off: every cell is
    walked."""
    sc = scenes.cavity_scene("c2")
    g = np.random.default_rng(5)
    out = []
    monkeypatch.setenv("SDR_P2G_HEAVY", "0")
    for perface in ("0", "1"):
        monkeypatch.setenv("SDR_P2G_PERFACE", perface)  # read when the solver is made
        d = scenes.make_solver(sc, precision=precision, **FAST)
        if not out:
            pos = scenes.fill_particles(d, sc, d.cells().reshape(d.nz, d.ny, d.nx) != 0, stride=1)
            c = np.array(d.cell_center(d.nx // 2, d.ny // 3, d.nz // 2))
            clump = c + g.uniform(-0.45, 0.45, (5000, 3)) * d.dx
            pos = np.concatenate([pos, clump])
            vel = g.standard_normal(pos.shape)
        d.set_particles(pos, vel)
        d.run_stage("classify", sc.dt)
        d.run_stage("p2g", sc.dt)
        out.append(d.grid())
        d.close()
    for a, b in zip(*out):
        assert np.abs(a).max() > 0
        assert bits_equal(a, b)


def _dilate(a, axis):
    """a | a shifted by one cell either way along `axis`."""
    out = a.copy()
    sl = [slice(None)] * a.ndim
    lo, hi = list(sl), list(sl)
    lo[axis], hi[axis] = slice(1, None), slice(None, -1)
    out[tuple(lo)] |= a[tuple(hi)]
    out[tuple(hi)] |= a[tuple(lo)]
    return out


def _faces_reached(heavy):
    """Faces whose P2G stencil covers a cell of `heavy` (z, y, x): along a component's axis
    the cells f - 1 and f, across it f - 1 .. f + 1."""
    nz, ny, nx = heavy.shape
    out = []
    for ax in (2, 1, 0):  # u (x), v (y), w (z)
        h = heavy
        for other in (0, 1, 2):
            if other != ax:
                h = _dilate(h, other)
        shape = list(heavy.shape)
        shape[ax] += 1
        f = np.zeros(shape, bool)
        lo, hi = [slice(None)] * 3, [slice(None)] * 3
        lo[ax], hi[ax] = slice(0, -1), slice(1, None)
        f[tuple(lo)] |= h  # face f = cell index (the cell after the face)
        f[tuple(hi)] |= h  # face f = cell index + 1 (the cell before)
        out.append(f)
    return out


@pytest.mark.parametrize("precision", [A.PRECISION_F32, A.PRECISION_F64])
def test_p2g_heavy_cells_match_face_gather(dev, precision, monkeypatch):
    """Cells holding at least SDR_P2G_HEAVY particles (32) are pre-reduced by a workgroup
    of their own and enter each face's sums as one term: faces no such cell reaches keep
    the face-by-face gather's bits, the others agree to rounding (the same terms summed in
    another order).  Clumps of 31, 32, 40-600 and 5000 particles on top of the fill; the
    faces counted as reached are those of every cell holding 30 or more (a superset, so
    that a particle rounding into the next cell in f32 cannot move the boundary)."""
    sc = scenes.cavity_scene("c2")
    g = np.random.default_rng(7)
    out = []
    for perface, heavy in (("1", "0"), ("0", "32")):
        monkeypatch.setenv("SDR_P2G_PERFACE", perface)
        monkeypatch.setenv("SDR_P2G_HEAVY", heavy)
        d = scenes.make_solver(sc, precision=precision, **FAST)
        if not out:
            fluid = d.cells().reshape(d.nz, d.ny, d.nx) != 0
            pos = scenes.fill_particles(d, sc, fluid, stride=1)
            kk, jj, ii = np.nonzero(fluid)
            pick = g.choice(kk.size, size=40, replace=False)
            sizes = [31, 32, 5000] + list(g.integers(40, 600, size=37))
            clumps = [np.array(d.cell_center(ii[p], jj[p], kk[p])) + g.uniform(-0.45, 0.45, (n, 3)) * d.dx
                      for p, n in zip(pick, sizes)]
            pos = np.concatenate([pos] + clumps)
            vel = g.standard_normal(pos.shape)
            loc = np.clip(((pos - np.asarray(d.origin)) / d.dx).astype(np.int64), 0,
                          [d.nx - 1, d.ny - 1, d.nz - 1])
            cnt = np.zeros((d.nz, d.ny, d.nx), np.int64)
            np.add.at(cnt, (loc[:, 2], loc[:, 1], loc[:, 0]), 1)
            heavy_cells = cnt >= 30
            assert (cnt >= 34).sum() >= 38
        d.set_particles(pos, vel)
        d.run_stage("classify", sc.dt)
        d.run_stage("p2g", sc.dt)
        out.append(d.grid())
        d.close()
    tol = FIELD_TOL if precision == A.PRECISION_F32 else 1e-12
    for ref, got, reach in zip(out[0], out[1], _faces_reached(heavy_cells)):
        assert reach.any() and not reach.all()
        assert bits_equal(ref[~reach], got[~reach])
        assert np.abs(got[reach] - ref[reach]).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_bounds_removals_match_oracle(dev, orc, mode):
    """enforce_particle_bounds + retain (solver.rs:736-755, particles.rs:38-49) with many
    removals: a c1 fill plus particles deep in the wall, outside the SDF box and in the
    outlet, shuffled.  The device compacts in place (survivors from the end fill the
    removed slots) and fixes the ranks from a bitmap of the removed ones; the survivors
    in reference order, the counts and a following classify must equal the oracle's."""
    sc, o, _ = paired("c1", orc, None, stride=2)
    d = scenes.make_solver(sc, **(EXACT if mode == "exact" else FAST))
    g = np.random.default_rng(21)
    pos, _ = o.particles()
    lo = np.asarray(sc.sdf.origin)
    hi = lo + sc.sdf.dx * (np.array([sc.sdf.nx, sc.sdf.ny, sc.sdf.nz]) - 1)
    extra = [lo + (hi - lo) * g.uniform(-0.2, 1.2, (3000, 3))]  # walls, and outside the box
    ob = sc.outlet
    extra.append(np.asarray(ob.min) + (np.asarray(ob.max) - np.asarray(ob.min)) * g.uniform(0, 1, (500, 3)))
    allp = np.concatenate([pos] + extra)
    allp = allp[g.permutation(allp.shape[0])]
    vel = g.standard_normal(allp.shape)
    for s_ in (o, d):
        s_.set_particles(allp, vel)
        s_.run_stage("bounds", sc.dt)
    so, sd = o.state(), d.state()
    assert so.culled > 1000 and (so.drained, so.culled) == (sd.drained, sd.culled)
    po, vo = o.particles()
    pd, vd = d.particles()
    assert bits_equal(po, pd) and bits_equal(vo, vd)
    for s_ in (o, d):
        s_.run_stage("classify", sc.dt)
    assert np.array_equal(o.cells(), d.cells())


def test_deterministic_run_to_run(dev):
    """No float atomics anywhere: two device runs agree bit for bit, fast mode included."""
    out = []
    for _ in range(2):
        sc = scenes.cavity_scene("c1")
        d = scenes.make_solver(sc, **FAST)
        d.seed_fill(sc.fill_ppc, sc.fill_lo, sc.fill_y, sc.fill_seed)
        for _ in range(3):
            d.substep(sc.dt)
        out.append((d.particles(), d.pressure()))
    (pa, va), qa = out[0]
    (pb, vb), qb = out[1]
    assert bits_equal(pa, pb) and bits_equal(va, vb) and bits_equal(qa, qb)


def test_device_seed_fill_matches_host_fill(dev, orc):
    sc = scenes.cavity_scene("c1")
    o = scenes.make_solver(sc, lib=orc)
    want = scenes.fill_particles(o, sc, oracle.cavity_mask(o))
    d = scenes.make_solver(sc, **EXACT)
    n = d.seed_fill(sc.fill_ppc, sc.fill_lo, sc.fill_y, sc.fill_seed)
    assert n == want.shape[0]
    assert bits_equal(d.particles()[0], want)


# ---------------------------------------------------------------- projection surface
def _proj(lib, n, fill, tol, iters=400, dot=A.DOT_EXACT):
    u, v, w = fill(n, n, n)
    cells = classify_box(n, n, n)
    rep, p = project_capturing_with_source(n, n, n, 0.01, (0.0, 0.0, 0.0), u, v, w, cells, 1000.0,
                                           1e-3, SolveParams(iters, tol), lib=lib, dot_mode=dot)
    return rep, p, (u, v, w)


def test_projection_bit_exact_vs_oracle(dev, orc):
    ro, po, fo = _proj(orc, 24, fill_divergent, 1e-12)
    rd, pd, fd = _proj(dev, 24, fill_divergent, 1e-12)
    assert ro.iters == rd.iters and ro.residual == rd.residual
    assert bits_equal(po, pd)
    for a, b in zip(fo, fd):
        assert bits_equal(a, b)


def test_projection_reference_known_answers_on_device(dev):
    # pressure.rs:619-642 idempotence
    _, _, (u, v, w) = _proj(dev, 12, fill_divergent, 1e-12)
    u1, v1, w1 = u.copy(), v.copy(), w.copy()
    project_capturing_with_source(12, 12, 12, 0.01, (0, 0, 0), u, v, w, classify_box(12, 12, 12),
                                  1000.0, 1e-3, SolveParams(400, 1e-12), lib=dev)
    assert max(np.abs(u - u1).max(), np.abs(v - v1).max(), np.abs(w - w1).max()) < 1e-7
    # pressure.rs:675-723 divergence removal, both dot modes
    for dot in (A.DOT_EXACT, A.DOT_FAST):
        _, _, (u, v, w) = _proj(dev, 12, radial_field, 1e-9, dot=dot)
        assert interior_divergence(u, v, w) < 1e-5
    # pressure.rs:731-752 bit reproducibility
    _, pa, _ = _proj(dev, 24, fill_divergent, 1e-12, dot=A.DOT_FAST)
    _, pb, _ = _proj(dev, 24, fill_divergent, 1e-12, dot=A.DOT_FAST)
    assert bits_equal(pa, pb)


def test_projection_edge_cases(dev, orc):
    n = 8
    cells = classify_box(n, n, n)
    zero = lambda: (np.zeros((n, n, n + 1)), np.zeros((n, n + 1, n)), np.zeros((n + 1, n, n)))
    u, v, w = zero()
    rep, p = project_capturing_with_source(n, n, n, 0.01, (0, 0, 0), u, v, w, cells, 1000.0, 1e-3,
                                           lib=dev)
    assert rep.iters == 0 and not p.any()  # zero RHS: pressure.rs:345-352
    # max_iters exhaustion (tol 0) and the all-air breakdown guard match the oracle
    for cl in (cells, np.full((n, n, n), 2, np.uint8)):
        res = []
        for lib in (orc, dev):
            u, v, w = fill_divergent(n, n, n)
            r, p = project_capturing_with_source(n, n, n, 0.01, (0, 0, 0), u, v, w, cl, 1000.0,
                                                 1e-3, SolveParams(5, 0.0), lib=lib)
            res.append((r.iters, r.residual, p, u))
        assert res[0][0] == res[1][0] and res[0][1] == res[1][1]
        assert bits_equal(res[0][2], res[1][2]) and bits_equal(res[0][3], res[1][3])


# ---------------------------------------------------------------- reference solver tests on device
def test_redistribution_known_answers_on_device(dev, orc):
    sc = scenes.box_scene()
    for n, here, below in ((12, 8, 4), (100, None, None)):
        d = scenes.make_solver(sc)
        o = scenes.make_solver(sc, lib=orc)
        pack_one_cell(d, (8, 8, 8), n, oracle.hash_jitter)
        pack_one_cell(o, (8, 8, 8), n, oracle.hash_jitter)
        d.redistribute()
        o.redistribute()
        pd, vd = d.particles()
        assert pd.shape[0] == n
        cnt = counts(d, pd)
        if here is not None:  # solver.rs:1212-1243
            assert cnt.get((8, 8, 8), 0) == here and cnt.get((8, 7, 8), 0) == below
            assert cnt.get((8, 9, 8), 0) == 0
        else:  # solver.rs:1251-1335
            assert max(cnt.values()) < n
            assert sum(1 for c in cnt.values() if c > 8) <= 1
        po, vo = o.particles()
        assert bits_equal(pd, po) and bits_equal(vd, vo)


def test_injection_and_walls_on_device(dev, orc):
    # solver.rs:1075-1099
    sc = scenes.box_scene()
    d = scenes.make_solver(sc)
    nd = sc.needle
    d.needle = Needle(nd.tip, nd.axis, nd.radius, nd.flow_rate, inject_until=0.02)
    for _ in range(10):
        d.step(0.005)
    after = d.particle_count()
    assert after > 0 and d.time > 0.02
    for _ in range(10):
        d.step(0.005)
    assert d.particle_count() == after
    # solver.rs:1102-1118 on the parametric sinus, and bit parity with the oracle
    sc = scenes.sinus_scene()
    d = scenes.make_solver(sc, **EXACT)
    o = scenes.make_solver(sc, lib=orc)
    for _ in range(12):
        d.step(0.005)
        o.step(0.005)
    pd, _ = d.particles()
    assert pd.shape[0] > 0
    assert max(oracle.sdf_sample(sc.sdf, p)[0] for p in pd) < sc.sdf.dx
    _compare_all(o, d, exact=True)


def test_mass_invariant_on_device(dev):
    sc = scenes.cavity_scene("c1")
    d = scenes.make_solver(sc, **FAST)
    n0 = d.seed_fill(sc.fill_ppc, sc.fill_lo, sc.fill_y, sc.fill_seed)
    for _ in range(5):
        d.substep(sc.dt)
    st = d.state()
    assert st.emitted == d.particle_count() + st.drained + st.culled
    assert st.emitted >= n0


@pytest.mark.parametrize("shape", [(150, 37, 41), (130, 19, 256)])
@pytest.mark.parametrize("variant", ["-1", "0", "1", "2"])
def test_mic_sweep_variants_bit_exact(dev, orc, variant, shape, monkeypatch):
    """Every built MIC(0) sweep shape reproduces the oracle's projection bit for bit on
    grids spanning several i-tiles and k-tiles, with solid, air and fluid cells mixed:
    the plane-by-plane kernels (-1), the one-plane-per-wave tiles of 16 and 8 planes
    (0, 1: the factor's shapes) and the default 7-plane tiles with an I/O wave (2)."""
    monkeypatch.setenv("SDR_MIC_VARIANT", variant)
    nx, ny, nz = shape
    cells, faces, (ro, po, *fo) = _multi_case(orc, shape)
    u, v, w = (a.copy() for a in faces)
    rd, pd = project_capturing_with_source(nx, ny, nz, 0.01, (0, 0, 0), u, v, w, cells, 1000.0, 1e-3,
                                           SolveParams(25, 1e-9), lib=dev)
    assert ro.iters == rd.iters and ro.residual == rd.residual
    assert bits_equal(po, pd)
    for a, b in zip(fo, (u, v, w)):
        assert bits_equal(a, b)


_MULTI_GRIDS = {}


def _multi_case(orc, shape):
    """Seeded projection case on `shape` and the oracle's result (computed once)."""
    if shape not in _MULTI_GRIDS:
        nx, ny, nz = shape
        g = np.random.default_rng(11)
        r = g.random((nz, ny, nx))
        cells = np.where(r < 0.8, 1, np.where(r < 0.9, 0, 2)).astype(np.uint8)
        faces = (g.standard_normal((nz, ny, nx + 1)), g.standard_normal((nz, ny + 1, nx)),
                 g.standard_normal((nz + 1, ny, nx)))
        u, v, w = (a.copy() for a in faces)
        rep, p = project_capturing_with_source(nx, ny, nz, 0.01, (0, 0, 0), u, v, w, cells, 1000.0,
                                               1e-3, SolveParams(25, 1e-9), lib=orc)
        _MULTI_GRIDS[shape] = (cells, faces, (rep, p, u, v, w))
    return _MULTI_GRIDS[shape]


@pytest.mark.parametrize("shape", [(150, 37, 41), (64, 64, 64)])
def test_fused_fast_solve_matches_unfused(dev, shape, monkeypatch):
    """Fast reduction mode with the default sweeps (dot(z, r) fused into the backward
    sweep) against the unfused kernels (variant 0): the same relative residual, and
    pressures equal to the solve's tolerance."""
    nx, ny, nz = shape
    g = np.random.default_rng(3)
    r = g.random((nz, ny, nx))
    cells = np.where(r < 0.85, 1, np.where(r < 0.93, 0, 2)).astype(np.uint8)
    faces = (g.standard_normal((nz, ny, nx + 1)), g.standard_normal((nz, ny + 1, nx)),
             g.standard_normal((nz + 1, ny, nx)))
    out = {}
    for variant in ("2", "0"):
        monkeypatch.setenv("SDR_MIC_VARIANT", variant)
        u, v, w = (a.copy() for a in faces)
        rep, p = project_capturing_with_source(nx, ny, nz, 0.01, (0, 0, 0), u, v, w, cells, 1000.0, 1e-3,
                                               SolveParams(5000, 0.0), dot_mode=A.DOT_FAST,
                                               solve_mode=A.SOLVE_RELATIVE, refine_tol=1e-10)
        out[variant] = (rep, p)
    (r2, p2), (r0, p0) = out["2"], out["0"]
    assert abs(r2.iters - r0.iters) <= max(3, r0.iters // 20), (r2.iters, r0.iters)
    scale = np.abs(p0).max()
    assert np.abs(p2 - p0).max() <= 1e-8 * scale


def _random_clumps(solver, seed, lo=0, hi=16):
    """Particles with a random count in [lo, hi] per cavity cell, at random offsets inside
    the cell and in shuffled reference order, so sources and receivers interleave and
    many sources share receivers (the order-dependent case of redistribute)."""
    g = np.random.default_rng(seed)
    mask = (solver.cells().reshape(solver.nz, solver.ny, solver.nx) != 0)
    kk, jj, ii = np.nonzero(mask)
    cnt = g.integers(lo, hi + 1, size=kk.size)
    ci = np.repeat(np.stack([ii, jj, kk], 1), cnt, axis=0).astype(np.float64)
    pos = np.asarray(solver.origin) + (ci + 0.5 + g.uniform(-0.45, 0.45, ci.shape)) * solver.dx
    pos = pos[g.permutation(pos.shape[0])]
    return pos


@pytest.mark.parametrize("sequential", ["0", "1"])
def test_redistribute_decision_paths_bit_exact(dev, orc, sequential, monkeypatch):
    """The order-dependent redistribution pass gives the oracle's bits both by the
    parallel reservation rounds and by the single-thread reference loop: one packed cell
    (12 particles; 100 and 3000, whose buckets are sorted in one workgroup; 12000 and
    20000, which also take one and two merge passes), random clumps on the box and on
    the 36x42x34 cavity, and four full substeps."""
    monkeypatch.setenv("SDR_REDIST_SEQUENTIAL", sequential)
    for n in (12, 100, 3000, 12000, 20000):
        sc = scenes.box_scene()
        d = scenes.make_solver(sc)
        o = scenes.make_solver(sc, lib=orc)
        pack_one_cell(d, (8, 8, 8), n, oracle.hash_jitter)
        pack_one_cell(o, (8, 8, 8), n, oracle.hash_jitter)
        d.redistribute()
        o.redistribute()
        assert bits_equal(d.particles()[0], o.particles()[0])
        assert d.state().redist_moved == o.state().redist_moved
        assert d.state().redist_active == o.state().redist_active
    for name, seed in (("box", 1), ("c1", 2), ("c1", 3)):
        sc = scenes.box_scene() if name == "box" else scenes.cavity_scene(name)
        d = scenes.make_solver(sc)
        o = scenes.make_solver(sc, lib=orc)
        d.run_stage("classify")
        pos = _random_clumps(d, seed)
        vel = np.random.default_rng(seed).normal(size=pos.shape)
        d.set_particles(pos, vel)
        o.set_particles(pos, vel)
        d.redistribute()
        o.redistribute()
        assert d.state().redist_moved == o.state().redist_moved > 0
        assert d.state().redist_active == o.state().redist_active
        pd, vd = d.particles()
        po, vo = o.particles()
        assert bits_equal(pd, po) and bits_equal(vd, vo)
    sc, o, d = paired("c1", orc, EXACT, stride=1)
    for _ in range(4):
        o.substep(sc.dt)
        d.substep(sc.dt)
    _compare_all(o, d, exact=True)
    assert d.state().redist_moved == o.state().redist_moved


def _precision_problem(seed=5):
    """A cavity-shaped projection problem on the 36x42x34 grid: fluid below a plane in
    the cavity mask, air above, solid outside; seeded random faces rounded to f32 (the
    storage of the precision path), so the f64 oracle sees the same inputs."""
    sc = scenes.cavity_scene("c1")
    s = scenes.make_solver(sc)
    s.run_stage("classify")
    mask = s.cells() != 0
    s.close()
    nz, ny, nx = mask.shape
    cells = np.where(mask, np.uint8(2), np.uint8(0))
    jj = np.arange(ny)[None, :, None] * np.ones((nz, 1, nx), int)
    cells[mask & (jj < ny * 2 // 3)] = FLUID
    g = np.random.default_rng(seed)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    u = f32(g.normal(size=(nz, ny, nx + 1)))
    v = f32(g.normal(size=(nz, ny + 1, nx)))
    w = f32(g.normal(size=(nz + 1, ny, nx)))
    return sc, cells, u, v, w


def _oracle_converged(orc, sc, cells, u, v, w, dt, rel):
    nz, ny, nx = cells.shape
    s = scenes.make_solver(sc, lib=orc)
    args = (nx, ny, nz, s.dx, s.origin)
    s.close()
    rho = sc.fluid.density
    rep0, _ = project_capturing_with_source(*args, u.copy(), v.copy(), w.copy(), cells, rho, dt,
                                            SolveParams(0, 0.0), lib=orc)
    bmax = rep0.residual  # zero iterations report max|b|
    uo, vo, wo = u.copy(), v.copy(), w.copy()
    rep, p = project_capturing_with_source(*args, uo, vo, wo, cells, rho, dt,
                                           SolveParams(20000, rel * bmax), lib=orc)
    assert rep.residual <= rel * bmax
    return bmax, p, (uo, vo, wo)


@pytest.mark.parametrize("precision", [A.PRECISION_F32, A.PRECISION_F64])
def test_relative_solve_matches_converged_oracle(dev, orc, precision):
    """SDR_SOLVE_RELATIVE: the f32 path (f64 residual, iterative refinement) and the f64
    path stop at max|r| <= 1e-8 max|b|; their pressure is within 1e-6 (relative, inf-norm)
    of the f64 oracle's solve of the same inputs run to 1e-12, and the projected faces
    within 1e-5."""
    sc, cells, u, v, w = _precision_problem()
    dt = sc.dt
    bmax, p_ref, faces_ref = _oracle_converged(orc, sc, cells, u, v, w, dt, 1e-12)
    fluid = sc.fluid
    fluid.volume_stiffness = 0.0
    sc.solve = SolveParams(5000, 1e-6)
    d = scenes.make_solver(sc, precision=precision, dot_mode=A.DOT_FAST,
                           solve_mode=A.SOLVE_RELATIVE, refine_tol=1e-8)
    d.set_cells(cells)
    d.set_grid(u, v, w)
    d.run_stage("project", dt)
    st = d.state()
    assert 0 < st.last_pcg_iters < 5000
    assert st.last_residual <= 1e-8 * bmax * (1 + 1e-12)
    p = d.pressure()
    err = np.abs(p - p_ref).max() / np.abs(p_ref).max()
    assert err <= 1e-6, err
    for a, b in zip(d.grid(), faces_ref):
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()
    d.close()
