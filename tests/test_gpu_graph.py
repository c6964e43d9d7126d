"""PCG iterations replayed from captured graphs (small grids) against the eager launches.

The graphs hold the launches of the eager loop in the same order, with the sweep tags
read from a device word, so every replay must give the eager path's bits: particles,
faces and pressure after several substeps, in both solve forms (the reference rule in
f64, whose plain loop also feeds the oracle parity tests, and the f32 reliable-update
solve of the serving scenarios), with the program's own eager work and host
synchronisations between the replays."""
import os

import numpy as np
import pytest

import oracle
from sdr_amd import _abi as A
from sdr_amd import scenes

from helpers import bits_equal

pytestmark = pytest.mark.gpu


def _run(sc, pos, graph: bool, substeps: int, **kw):
    old = os.environ.get("SDR_PCG_GRAPH")
    os.environ["SDR_PCG_GRAPH"] = "1" if graph else "0"
    try:
        d = scenes.make_solver(sc, **kw)  # the setting is read when the solver is made
    finally:
        if old is None:
            del os.environ["SDR_PCG_GRAPH"]
        else:
            os.environ["SDR_PCG_GRAPH"] = old
    d.set_particles(pos)
    iters = []
    for _ in range(substeps):
        d.substep(sc.dt)
        iters.append(d.state().last_pcg_iters)
        d.particles()  # eager work and a host synchronisation between the replays
    pd, vd = d.particles()
    return pd, vd, d.grid(), d.pressure(), iters


@pytest.mark.parametrize("form", ["f64_reference", "f32_reliable", "f64_fast"])
def test_graph_replay_matches_eager(dev, orc, form):
    sc = scenes.cavity_scene("c1")
    o = scenes.make_solver(sc, lib=orc)
    pos = scenes.fill_particles(o, sc, oracle.cavity_mask(o))
    kw = {"f64_reference": {"dot_mode": A.DOT_EXACT},
          "f64_fast": {"dot_mode": A.DOT_FAST},
          "f32_reliable": {"precision": A.PRECISION_F32, "solve_mode": A.SOLVE_RELATIVE, "refine_tol": 1e-8}}[form]
    a = _run(sc, pos, True, 6, **kw)
    b = _run(sc, pos, False, 6, **kw)
    assert a[4] == b[4] and min(a[4]) > 8  # whole 8-iteration blocks were replayed
    assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1])
    for x, y in zip(a[2], b[2]):
        assert bits_equal(x, y)
    assert bits_equal(a[3], b[3])
