"""Device parity on the workload configurations, the precision form against the oracle,
diagnostics, clinical metrics and the NaN guard.

Bars (stated per test):
  * f64 exact mode: bit-exact against the oracle (C1 over 200 substeps, C2 at its full
    8-per-cell fill), diagnostics and metrics integer-exact, maxima bit-exact, sums
    within 1e-12 relative (the device sums in a fixed tree, the reference sequentially);
  * precision form (f32 storage, f64 dots/norms/refinement to a 1e-8 relative residual)
    after one substep from the oracle's state, the oracle solved to 1e-11: particle
    count, cells, drained/culled exact; positions within 1e-5 dx; velocities, faces and
    pressure within 1e-5 of the field's max magnitude, except particles whose
    redistribution or cull decision flipped on f32 rounding, which are counted and must
    stay below 0.1 %; pressure-based metrics within 1e-5;
  * full sizes (256^3, 512x512x384): the integer mass invariant, run-to-run bits, the
    projection within 1e-6 of an f64 solve to 1e-12, and the
    projected divergence equal to the volume-control target to 1e-6 of the input's.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from sdr_amd import _abi as A
from sdr_amd import scenes, serving
from sdr_amd.metrics import MetricsCollector
from sdr_amd.solver import FLUID, SolveParams, project_capturing_with_source

from helpers import bits_equal, classify_box, count_source, divergence, fill_divergent, paired, progress

pytestmark = pytest.mark.gpu
EXACT = {"dot_mode": A.DOT_EXACT}
F32 = {"precision": A.PRECISION_F32, "dot_mode": A.DOT_FAST, "solve_mode": A.SOLVE_RELATIVE, "refine_tol": 1e-8}


def _compare_exact(o, d):
    po, vo = o.particles()
    pd, vd = d.particles()
    assert po.shape == pd.shape
    assert bits_equal(po, pd) and bits_equal(vo, vd)
    assert np.array_equal(o.cells(), d.cells())
    for a, b in zip(o.grid(), d.grid()):
        assert bits_equal(a, b)
    assert bits_equal(o.pressure(), d.pressure())
    so, sd = o.state(), d.state()
    assert (so.drained, so.culled, so.emitted, so.last_pcg_iters) == \
        (sd.drained, sd.culled, sd.emitted, sd.last_pcg_iters)


def test_c1_200_substeps_bit_exact(dev, orc):
    """C1 as BASELINE states it: 200 substeps of the 36x42x34 cavity, bit for bit."""
    sc, o, d = paired("c1", orc, EXACT)
    for k in range(200):
        o.substep(sc.dt)
        d.substep(sc.dt)
        if k % 50 == 49:
            progress(f"c1 substep {k + 1}")
            _compare_exact(o, d)
    st = d.state()
    assert st.emitted == d.particle_count() + st.drained + st.culled
    assert st.emitted > 60000


def test_c2_full_fill_substeps_bit_exact(dev, orc):
    """C2 at its full 8-per-cell fill (~1M particles, 72x84x68): four substeps."""
    sc, o, d = paired("c2", orc, EXACT, stride=1)
    assert o.particle_count() > 900_000
    for k in range(4):
        o.substep(sc.dt)
        d.substep(sc.dt)
        progress(f"c2 substep {k + 1}")
    _compare_exact(o, d)


# ---------------------------------------------------------------- precision form vs oracle
def _tight_oracle_substep(o, sc, rel):
    """One oracle substep whose projection is solved to `rel` of max|b| (the reference's
    absolute rule at tol = rel * max|b|)."""
    dt = sc.dt
    for st in ("emit", "classify", "p2g", "save", "forces"):
        o.run_stage(st, dt)
    u, v, w = o.grid()
    cells = o.cells()
    src = count_source(o, sc, cells, f32=False)
    rep0, _ = project_capturing_with_source(o.nx, o.ny, o.nz, o.dx, o.origin, u.copy(), v.copy(), w.copy(),
                                            cells, sc.fluid.density, dt, SolveParams(0, 0.0), src, lib=o.lib)
    keep = o.solve
    o.solve = SolveParams(50000, rel * rep0.residual)
    o.run_stage("project", dt)
    o.solve = keep
    for st in ("extrapolate", "g2p", "advect", "bounds", "redistribute", "finish"):
        o.run_stage(st, dt)
    return rep0.residual


@pytest.mark.parametrize("level", ["c1", "c2"])
def test_f32_substep_matches_oracle(dev, orc, level):
    sc, o, d = paired(level, orc, F32, stride=1)
    sc.solve = SolveParams(5000, 1e-6)
    d.solve = sc.solve
    mo = MetricsCollector(o, 1.0e-5)
    md = MetricsCollector(d, 1.0e-5)
    _tight_oracle_substep(o, sc, 1e-11)
    d.substep(sc.dt)
    po, vo = o.particles()
    pd, vd = d.particles()
    assert po.shape == pd.shape
    assert np.array_equal(o.cells(), d.cells())
    so, sd = o.state(), d.state()
    assert (so.drained, so.culled, so.emitted) == (sd.drained, sd.culled, sd.emitted)
    dpos = np.abs(pd - po).max(axis=1)
    vscale = np.abs(vo).max()
    dvel = np.abs(vd - vo).max(axis=1)
    bad = (dpos > 1e-5 * o.dx) | (dvel > 1e-5 * vscale)
    progress(f"{level} f32: {bad.sum()} of {bad.size} particles past the bound; "
             f"max dpos {dpos.max() / o.dx:.2e} dx, max dvel {dvel.max() / vscale:.2e}")
    assert bad.sum() <= 1e-3 * bad.size
    for a, b in zip(d.grid(), o.grid()):
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()
    p_o, p_d = o.pressure(), d.pressure()
    assert np.abs(p_d - p_o).max() <= 1e-5 * np.abs(p_o).max()
    fo, fd = mo.record(o, 0), md.record(d, 0)
    for k in ("particles", "emitted", "drained", "culled"):
        assert fo[k] == fd[k], k
    assert fo["fill_fraction"] == fd["fill_fraction"]
    # a particle within f32 rounding of a cell face may mark a different wall cell
    for k in ("wet_wall_cells", "membrane_artifacts"):
        assert abs(int(fo[k]) - int(fd[k])) <= max(2, 1e-4 * fo[k]), (k, fo[k], fd[k])
    assert abs(fo["wall_coverage"] - fd["wall_coverage"]) <= max(2.0 / md.wall_count, 1e-4)
    for k in ("mean_pressure_pa", "max_pressure_pa", "wall_pressure_pa", "peak_wall_pressure_pa",
              "max_speed_m_s", "kinetic_energy"):
        assert math.isclose(fd[k], fo[k], rel_tol=1e-5), (k, fd[k], fo[k])


# ---------------------------------------------------------------- diagnostics and metrics
def test_diagnostics_and_metrics_match_oracle(dev, orc):
    sc, o, d = paired("c1", orc, EXACT)
    cv = float(oracle.cavity_mask(o).sum()) * o.dx ** 3
    mo, md = MetricsCollector(o, cv), MetricsCollector(d, cv)
    assert mo.wall_count == md.wall_count > 0
    for f in range(5):
        for _ in range(3):
            o.substep(sc.dt)
            d.substep(sc.dt)
        a, b = o.diagnostics(), d.diagnostics()
        for k in ("particle_volume", "fluid_volume", "drained_volume", "cells_fluid_count",
                  "max_particles_per_cell", "max_pressure", "max_speed", "particles"):
            assert a[k] == b[k], (k, a[k], b[k])
        for k in ("mean_pressure", "kinetic_energy"):
            assert math.isclose(a[k], b[k], rel_tol=1e-12), (k, a[k], b[k])
        ro, rd = mo.record(o, f), md.record(d, f)
        for k, va in ro.items():
            if k in ("mean_pressure_pa", "wall_pressure_pa", "kinetic_energy"):
                assert math.isclose(va, rd[k], rel_tol=1e-12), (k, va, rd[k])
            else:
                assert va == rd[k], (k, va, rd[k])
    wo, io = mo.cells()
    wd, id_ = md.cells()
    assert np.array_equal(wo, wd) and np.array_equal(io, id_)
    so, sd = mo.summary(), md.summary()
    for k, va in so.items():
        if k.startswith("mean_membrane"):
            assert math.isclose(va, sd[k], rel_tol=1e-12)
        else:
            assert va == sd[k], (k, va, sd[k])
    assert sd["mass_balance"] == 1.0


# ---------------------------------------------------------------- NaN guard
def test_nan_residual_is_an_error(dev, orc):
    """A NaN face reaches the residual; the reference's NaN-blind inf_norm would report
    convergence, the device returns SDR_E_NUMERIC (standalone projection and the solver's
    project stage), and the solver stays usable."""
    n = 12
    cells = classify_box(n, n, n)
    u, v, w = fill_divergent(n, n, n)
    u[5, 5, 5] = np.nan
    with pytest.raises(A.AbiError) as e:
        project_capturing_with_source(n, n, n, 0.01, (0, 0, 0), u, v, w, cells, 1000.0, 1e-3,
                                      SolveParams(200, 1e-9), lib=dev)
    assert e.value.code == A.SDR_E_NUMERIC
    sc, o, d = paired("c1", orc, EXACT)
    for st in ("emit", "classify", "p2g", "save", "forces"):
        d.run_stage(st, sc.dt)
    gu, gv, gw = d.grid()
    cl = d.cells()
    # a u-face between two Fluid cells (faces touching Solid are zeroed before the solve)
    pairs = np.argwhere((cl[:, :, 1:] == FLUID) & (cl[:, :, :-1] == FLUID))
    k, j, i = pairs[len(pairs) // 2]
    gu[k, j, i + 1] = np.nan
    d.set_grid(gu, gv, gw)
    with pytest.raises(A.AbiError) as e:
        d.run_stage("project", sc.dt)
    assert e.value.code == A.SDR_E_NUMERIC
    d.set_grid(*o.grid())
    d.run_stage("project", sc.dt)  # a finite grid solves again


# ---------------------------------------------------------------- full sizes
def _full_size_check(level, rel_ref, n_run=2):
    sc = scenes.cavity_scene(level)
    progress(f"{level}: scene built")
    sc.solve = SolveParams(4000, sc.solve.tol)
    runs = []
    for q in range(n_run):
        d = scenes.make_solver(sc, **F32)
        d.solve = sc.solve
        n0 = d.seed_fill(sc.fill_ppc, sc.fill_lo, sc.fill_y, sc.fill_seed)
        dt = sc.dt
        for st in ("emit", "classify", "p2g", "save", "forces"):
            d.run_stage(st, dt)
        if q == 0:
            u, v, w = d.grid()
            cells = d.cells()
            src = count_source(d, sc, cells, f32=True)
            uo, vo, wo = u.copy(), v.copy(), w.copy()
            rep, p_ref = project_capturing_with_source(
                d.nx, d.ny, d.nz, d.dx, d.origin, uo, vo, wo, cells, sc.fluid.density, dt,
                SolveParams(50000, 0.0), src, dot_mode=A.DOT_FAST, solve_mode=A.SOLVE_RELATIVE,
                refine_tol=rel_ref)
            progress(f"{level}: f64 reference solve, {rep.iters} iterations")
        d.run_stage("project", dt)
        p = d.pressure()
        ud, vd, wd = d.grid()
        for st in ("extrapolate", "g2p", "advect", "bounds", "redistribute", "finish"):
            d.run_stage(st, dt)
        st_ = d.state()
        assert st_.emitted == d.particle_count() + st_.drained + st_.culled
        assert st_.emitted >= n0
        runs.append((p, (ud, vd, wd), d.diagnostics(), st_.last_pcg_iters))
        progress(f"{level}: run {q}: {n0} particles, {st_.last_pcg_iters} PCG iterations")
        d.close()
        del d
    p, (ud, vd, wd), diag, iters = runs[0]
    err = np.abs(p - p_ref).max() / np.abs(p_ref).max()
    assert err <= 1e-6, err
    fl = cells == FLUID
    s = src if src is not None else 0.0
    din = np.abs((divergence(u, v, w) - s)[fl]).max()
    dout = np.abs((divergence(ud, vd, wd) - s)[fl]).max()
    progress(f"{level}: pressure error {err:.2e}, divergence residual {dout:.3e} of {din:.3e}")
    assert dout <= 1e-6 * din
    for other in runs[1:]:
        assert bits_equal(other[0], p)
        for a, b in zip(other[1], runs[0][1]):
            assert bits_equal(a, b)
        assert other[2] == diag and other[3] == iters


@pytest.mark.timeout(600)
def test_c3_full_size_substep(dev):
    """C3: one 256^3 substep of ~121M particles in the precision form."""
    _full_size_check("c3", 1e-12)


@pytest.mark.timeout(900)
def test_c5_full_size_substep(dev):
    """C5: one 512x512x384 substep of ~397M particles, inside 288 GB."""
    out = C.c_uint64()
    from sdr_amd import load_library

    load_library().solver_footprint(512, 512, 384, 400_000_000, A.PRECISION_F32, C.byref(out))
    assert out.value < 288e9 * 0.5
    _full_size_check("c5", 1e-12, n_run=1)


def test_c4_mixed_sizes_alone_and_in_the_mix(dev):
    """C4's grid range: 32^3, 64^3, 96^3 and 128^3 scenarios served together give the
    bits each gives alone."""
    specs = [serving.ScenarioSpec(n, ang, flow, 1.0, 0.0, 500 + q, 6)
             for q, (n, ang, flow) in enumerate([(32, 10.0, 3.0), (64, -20.0, 5.0), (96, 0.0, 1.5), (128, 25.0, 6.0)])]
    mix = serving.serve(specs, max_slots=4)
    assert all(r.status == A.SDR_OK for r in mix)
    assert serving.max_in_flight(mix) >= 2
    for spec, r in zip(specs, mix):
        alone = serving.serve([spec], max_slots=1)[0]
        assert alone.status == A.SDR_OK
        assert (alone.particles, alone.checksum, alone.pcg_iters) == (r.particles, r.checksum, r.pcg_iters)
