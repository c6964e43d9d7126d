"""Serving many irrigation scenarios on one GPU (the native loop is `sdr_serve`).

The reference simulates one scene per process (src/main.rs:215-386); there is no
serving counterpart.  This module builds the serving workload and summarises a run:

* `serving_workload` draws scenarios from a seed: Poisson arrivals (mean `rate` per
  second), cubic grids n^3 with n uniform in [n_min, n_max], a needle tilted by an
  angle uniform in +-30 degrees, flow uniform in 1-6 ml/s and dose uniform in 0.5-2 ml.
  A dose of 0.5-2 ml at 1-6 ml/s is 0.08-2 s of injection, i.e. up to ~1e5 substeps per
  scenario at these grid spacings; a request here is the first `substeps` substeps of
  that injection, with `substeps` = 8 + 32 * (dose/flow - 1/12) / (2 - 1/12) rounded
  (8-40: longer injections ask for more work).
* `serve` hands them to the native loop (admission by estimated HBM footprint under a
  budget and a number of slots; one solver, stream and host thread per admitted
  scenario) and returns one record per scenario.
* `summarize` gives p50/p99 latency (finish - arrival, queueing included) and the
  aggregate particle-substeps/s over the makespan.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from .scenes import MM, ML_PER_S, _cavity_fn, _socket_outlet
from .solver import FluidParams, Needle, Sdf, SolveParams


@dataclass
class ScenarioSpec:
    """One request of the serving workload."""

    n: int                 # grid cells per axis
    angle_deg: float       # needle tilt about x, from +z
    flow_ml_s: float
    dose_ml: float
    arrival_s: float
    seed: int
    substeps: int
    dt: float = 2e-5
    precision: int = A.PRECISION_F32
    rel_tol: float = 1e-6  # relative residual of the pressure solve (f64-refined)


def substeps_for(dose_ml: float, flow_ml_s: float) -> int:
    d = dose_ml / flow_ml_s
    return int(round(8 + 32 * (d - 1.0 / 12.0) / (2.0 - 1.0 / 12.0)))


def serving_workload(count: int = 64, seed: int = 2026, rate: float = 8.0, n_min: int = 32,
                     n_max: int = 128) -> list[ScenarioSpec]:
    """The seeded request stream (Poisson arrivals at `rate` per second)."""
    g = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(count):
        t += float(g.exponential(1.0 / rate))
        n = int(g.integers(n_min, n_max + 1))
        ang = float(g.uniform(-30.0, 30.0))
        flow = float(g.uniform(1.0, 6.0))
        dose = float(g.uniform(0.5, 2.0))
        out.append(ScenarioSpec(n, ang, flow, dose, t, seed * 1000 + q, substeps_for(dose, flow)))
    return out


def scenario_dt(spec: ScenarioSpec) -> float:
    """The fixed substep of a request: spec.dt, at most one cell of jet travel (the CFL
    bound of solver.rs:262-281)."""
    needle = Needle(tip=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 1.0), radius=0.25 * MM,
                    flow_rate=spec.flow_ml_s * ML_PER_S)
    return min(spec.dt, (32 * MM / spec.n) / needle.jet_speed())


class _Built:
    """Host inputs of one scenario (kept alive while the native loop runs)."""

    def __init__(self, spec: ScenarioSpec):
        n = spec.n
        # a 32 mm box-filling rounded cavity with a socket channel, at dx = 32 mm / n
        dx = 32 * MM / n
        half = n * dx / 2.0
        origin = (-half, -half, -half)
        semi = (0.495 * n * dx,) * 3
        self.sdf = Sdf.from_fn(origin, dx, (n + 1, n + 1, n + 1),
                               _cavity_fn(semi, 2 * MM, 1 * MM, power=12.0))
        a = math.radians(spec.angle_deg)
        axis = (0.0, math.sin(a), math.cos(a))
        tip = (0.0, -0.25 * half, -0.5 * half)
        needle = Needle(tip=tip, axis=axis, radius=0.25 * MM, flow_rate=spec.flow_ml_s * ML_PER_S)
        self.dt = scenario_dt(spec)
        outlet = _socket_outlet(origin[1], 2 * MM, dx)
        c = A.Scenario()
        c.solid = self.sdf.c()
        c.sim_dx = dx
        c.fluid = FluidParams().c()
        c.needle = needle.c()
        c.opts = A.Options(spec.precision, A.DOT_FAST, 0, spec.rel_tol, A.SOLVE_RELATIVE, 0)
        c.solve = A.SolveParamsC(2000, 1e-6)
        c.has_outlet = 1
        c.outlet = outlet.c()
        c.fill_per_cell = 8
        c.fill_y_lo = origin[1] + 2 * dx
        c.fill_y_hi = 0.0
        c.fill_seed = spec.seed
        c.dt = self.dt
        c.substeps = spec.substeps
        c.arrival_s = spec.arrival_s
        self.c = c


@dataclass
class ScenarioRecord:
    spec: ScenarioSpec
    status: int
    slot: int
    arrival_s: float
    admit_s: float
    finish_s: float
    footprint_bytes: int
    particles: int
    particle_substeps: int
    pcg_iters: int
    checksum: int
    device_allocs: int = 0

    @property
    def latency_s(self) -> float:
        return self.finish_s - self.arrival_s


def serve(specs: list[ScenarioSpec], budget_bytes: float = 200e9, max_slots: int = 8,
          lib: A.Lib | None = None, resident_pools: bool = True) -> list[ScenarioRecord]:
    """Run the requests through the native serving loop (blocks until all finish);
    resident_pools keeps each slot's device memory across the requests it serves."""
    if lib is None:
        from . import load_library

        lib = load_library()
    built = [_Built(s) for s in specs]
    arr = (A.Scenario * len(specs))(*[b.c for b in built])
    res = (A.ScenarioResult * len(specs))()
    cfg = A.ServeConfig(float(budget_bytes), int(max_slots), 1 if resident_pools else 0)
    lib.serve(arr, len(specs), C.byref(cfg), res)
    return [ScenarioRecord(s, r.status, r.slot, r.arrival_s, r.admit_s, r.finish_s, r.footprint_bytes,
                           r.particles, r.particle_substeps, r.pcg_iters, r.checksum, r.device_allocs)
            for s, r in zip(specs, res)]


def summarize(recs: list[ScenarioRecord]) -> dict:
    ok = [r for r in recs if r.status == A.SDR_OK]
    lat = np.array([r.latency_s for r in ok]) if ok else np.zeros(1)
    t0 = min(r.arrival_s for r in recs)
    t1 = max(r.finish_s for r in recs)
    ps = sum(r.particle_substeps for r in ok)
    return {
        "scenarios": len(recs),
        "completed": len(ok),
        "rejected": sum(1 for r in recs if r.status == A.SDR_E_OOM),
        "failed": sum(1 for r in recs if r.status not in (A.SDR_OK, A.SDR_E_OOM)),
        "p50_latency_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
        "p99_latency_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
        "mean_latency_ms": round(float(lat.mean()) * 1e3, 2),
        "makespan_s": round(t1 - t0, 3),
        "particle_substeps": int(ps),
        "aggregate_particle_substeps_per_s": round(ps / max(t1 - t0, 1e-9), 1),
        "max_in_flight": max_in_flight(recs),
        "device_allocs": int(sum(r.device_allocs for r in recs)),
    }


def max_in_flight(recs: list[ScenarioRecord]) -> int:
    ev = []
    for r in recs:
        if r.status == A.SDR_OK:
            ev.append((r.admit_s, 1))
            ev.append((r.finish_s, -1))
    ev.sort(key=lambda e: (e[0], e[1]))
    cur = best = 0
    for _, d in ev:
        cur += d
        best = max(best, cur)
    return best
