#!/usr/bin/env python3
"""Benchmark of the FLIP/PIC substep on one MI355X.

A "step" is one `Solver::substep(dt)` (emit, classify, P2G, forces, MIC(0)-PCG
projection, extrapolation, G2P, advection, culling, redistribution) over the 256^3
cavity workload: a box-filling rounded cavity with a socket channel, ~130M particles
seeded 8 per cell, fixed dt.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line with particle-substeps/s (whole job), the pressure-solve time,
the HBM-roofline fraction of the dominant kernel (HIP events on the solver's stream),
and the CPU oracle measured on the reference's own CPU configuration (200 substeps of
C1, with the device run on the same workload beside it) plus a composed estimate of
the headline workload from a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "particle-substeps/s + pressure-solve ms at 256³ vs FP64 GPU; p99 scenario latency"


def region_bytes(name: str, s: int, nc: int, nsh: int, np_: int, nsdf: int) -> float | None:
    """Algorithmic HBM bytes of one call of a timed region (the traffic the mathematics
    needs, independent of the implementation): s = scalar size, nc = cells, np_ = particles,
    flags = the u16 coupling flags of a cell."""
    flags = 2
    table = {
        "mic_fwd": nc * 3 * s,                  # read r, precon (Fluid = precon != 0); write q
        "mic_bwd": nc * (3 * s + flags),        # read q, precon, flags; write z
        "mic_bwd_dot": nc * 3 * s,              # fast mode: read q, precon; write z; dot(q, q) = dot(z, r)
        "mic_fwd_axpy": nc * (8 * s + flags),   # read r, As, p, s, precon, flags; write r, p, q
        "update_s": nc * 3 * s,                 # s = z + b s
        "mic_build": nc * (s + flags),          # read flags; write precon
        "spmv_dot": nc * (2 * s + flags),       # read s, flags; write As (dot(s, As) in the same pass)
        "spmv_update": nc * (4 * s + flags),    # read z, old s, flags; write new s, As (and dot(s, As))
        "axpy_norm": nc * 6 * s,                # p += a s; r -= a As; max|r|
        "dot_update": nc * 5 * s,               # dot(z, r); s = z + b s
        "p2g": np_ * 6 * s + 3 * nc * s,        # read x, v; write u, v, w
        "g2p": np_ * 9 * s + 6 * nc * s,        # read x, v; write v; read u,v,w and saved
        "advect": np_ * 6 * s + 3 * nc * s + nsdf * s,  # read/write x; read u,v,w; SDF
        "g2p_advect": np_ * 12 * s + 6 * nc * s + nsdf * s,  # read/write x, v; read u,v,w, saved; SDF
    }
    return table.get(name)


def host_cpu() -> dict:
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def c1_run(lib, substeps: int, **kw):
    """The reference's own CPU configuration (36x42x34 cells at 1 mm, 8 jittered
    particles per cavity cell, reference rule tol 1e-6 within 200 iterations, f64) for
    `substeps` substeps on `lib` (None = the device library); returns (seconds,
    particle-substeps, mean PCG iterations)."""
    import oracle
    from sdr_amd import scenes

    sc = scenes.cavity_scene("c1")
    o = scenes.make_solver(sc, lib=oracle.load())
    pos = scenes.fill_particles(o, sc, oracle.cavity_mask(o))
    o.close()
    s = scenes.make_solver(sc, lib=lib, **kw)
    s.set_particles(pos)
    ps, its = 0, []
    t = time.perf_counter()
    for _ in range(substeps):
        ps += s.particle_count()
        s.substep(sc.dt)
        its.append(s.state().last_pcg_iters)
    if lib is None:
        s.synchronize()
    el = time.perf_counter() - t
    s.close()
    return el, ps, float(np.mean(its))


def cpu_baseline_c1(threads: int, substeps: int = 200, single_substeps: int = 60) -> dict:
    """The CPU oracle measured on the reference's own CPU configuration (C1, 200
    substeps) with `threads` threads and with one thread (on the first
    `single_substeps`), next to the device f64 path on the same run (same scene, fill,
    dt, solve rule and iteration cap)."""
    import oracle
    from sdr_amd import _abi as A

    lib = oracle.load()
    oracle.set_threads(threads)
    el, ps, it = c1_run(lib, substeps)
    print(f"[cpu] c1 {substeps} substeps, {threads} threads: {el:.2f} s", file=sys.stderr, flush=True)
    oracle.set_threads(1)
    el1, ps1, _ = c1_run(lib, single_substeps)
    print(f"[cpu] c1 {single_substeps} substeps, 1 thread: {el1:.2f} s", file=sys.stderr, flush=True)
    oracle.set_threads(threads)
    c1_run(None, 5, dot_mode=A.DOT_EXACT)  # warm-up (code objects)
    gel, gps, git = c1_run(None, substeps, dot_mode=A.DOT_EXACT)
    print(f"[gpu] c1 {substeps} substeps (f64, exact): {gel:.3f} s", file=sys.stderr, flush=True)
    out = {"value": round(ps / el, 1), "unit": "particle-substeps/s", "cores": threads, "kind": "port",
           "sample": (f"C++ f64 oracle (reference operand order, the reference's rayon loops on OpenMP) "
                      f"on the reference's own CPU configuration C1 (36x42x34 cells, ~60k particles, "
                      f"tol 1e-6 / 200 iterations): {substeps} substeps measured ({el:.1f} s, "
                      f"{it:.1f} PCG iterations per substep); one thread on the first {single_substeps}"),
           "ms_per_substep": round(el / substeps * 1e3, 3),
           "single_thread_value": round(ps1 / el1, 1),
           "single_thread_ms_per_substep": round(el1 / single_substeps * 1e3, 3),
           "gpu_same_run": {"value": round(gps / gel, 1), "ms_per_substep": round(gel / substeps * 1e3, 3),
                            "precision": "f64, reference dot tree (bit-identical to the oracle)",
                            "pcg_iters_mean": round(git, 2)},
           "speedup_same_run": round((gps / gel) / (ps / el), 1),
           "speedup_same_run_single_thread": round((gps / gel) / (ps1 / el1), 1)}
    out.update(host_cpu())
    return out


def cpu_baseline(scene, sdf, gpu_iters: float, budget_s: float, threads: int):
    """The CPU oracle (f64, reference operand order and threading) on a bounded sample
    of the headline workload, an extra beside the measured C1 baseline: every stage on
    the full 256^3 grid, the particle stages on a strided 1/`stride` subset of the
    particles, the PCG at 2 and 4 iterations.  The per-substep time is composed as grid
    stages + PCG per-iteration x the GPU's iteration count + particle stages x stride."""
    import oracle
    from sdr_amd import scenes

    oracle.set_threads(threads)
    lib = oracle.load()
    s = scenes.make_solver(scene, lib=lib)
    mask = oracle.cavity_mask(s)
    kk, jj, ii = np.nonzero(mask)
    cy = s.origin[1] + (jj + 0.5) * s.dx
    n_full = int(((cy >= scene.fill_lo) & (cy <= scene.fill_y)).sum()) * scene.fill_ppc
    stride = max(1, n_full // 1_000_000)
    sub = np.ascontiguousarray(scenes.fill_particles(s, scene, mask, stride=stride))
    s.set_particles(sub)
    dt = scene.dt
    times = {}
    stages = ["emit", "classify", "p2g", "save", "forces"]
    for st in stages:
        t = time.perf_counter()
        s.run_stage(st, dt)
        times[st] = time.perf_counter() - t
        print(f"[cpu] {st}: {times[st]:.3f}s", file=sys.stderr, flush=True)
    # PCG cost: two capped solves on the same post-P2G state.
    from sdr_amd.solver import SolveParams
    u, v, w = s.grid()
    proj = {}
    for m in (2, 4):  # capped solves: the per-iteration slope and the fixed part
        s.set_grid(u, v, w)
        s.solve = SolveParams(m, 0.0)
        t = time.perf_counter()
        s.run_stage("project", dt)
        proj[m] = time.perf_counter() - t
        print(f"[cpu] project ({m} iterations): {proj[m]:.3f}s", file=sys.stderr, flush=True)
    per_iter = max((proj[4] - proj[2]) / 2.0, 0.0)
    fixed_proj = max(proj[2] - 2 * per_iter, 0.0)
    for st in ["extrapolate", "g2p", "advect", "bounds", "redistribute"]:
        t = time.perf_counter()
        s.run_stage(st, dt)
        times[st] = time.perf_counter() - t
        print(f"[cpu] {st}: {times[st]:.3f}s", file=sys.stderr, flush=True)
    particle_stages = ["p2g", "g2p", "advect", "bounds", "redistribute", "classify"]
    t_particles = sum(times[k] for k in particle_stages) * stride
    t_grid = sum(v for k, v in times.items() if k not in particle_stages)
    t_sub = t_grid + t_particles + fixed_proj + per_iter * gpu_iters
    sample_s = sum(times.values()) + sum(proj.values())
    s.close()
    return {
        "value": n_full / t_sub,
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"C++ f64 oracle, 1 substep of the same 256^3 workload: all grid stages on "
                   f"the full grid, particle stages on every {stride}th particle "
                   f"({sub.shape[0]} of {n_full}) scaled x{stride}, PCG timed at 2 and 4 "
                   f"iterations and scaled to the GPU's {gpu_iters:.1f} iterations; "
                   f"{sample_s:.1f} s of CPU work, composed CPU substep {t_sub:.1f} s"),
        "ms_per_substep": t_sub * 1e3,
        "pcg_ms_per_iter": per_iter * 1e3,
    }


def run_path(scene, prec, dot_mode, solve_mode, tol, max_iters, steps, warmup, label, regions=True,
             breakdown_steps=0):
    """Seed the workload, run warmup + timed substeps; returns (solver, elapsed, iters,
    particle-substeps, sweep timing, breakdown).

    The timed substeps carry HIP events around one call in 16 of each MIC(0) sweep
    region (the roofline kernel; an event pair delays the next kernel by a few
    microseconds, so events around every region cost ~4% of a substep and around every
    sweep ~1.7%).  `breakdown_steps` more
    substeps follow, untimed, with events around every region: their per-region times
    and PCG iterations are the breakdown."""
    from sdr_amd import _abi as A
    from sdr_amd import scenes
    from sdr_amd.solver import SolveParams

    t0 = time.time()
    s = scenes.make_solver(scene, precision=prec, dot_mode=dot_mode, solve_mode=solve_mode,
                           refine_tol=tol)
    if solve_mode == A.SOLVE_RELATIVE:
        s.solve = SolveParams(max_iters, scene.solve.tol)
    n0 = s.seed_fill(scene.fill_ppc, scene.fill_lo, scene.fill_y, scene.fill_seed)
    print(f"[bench] {label}: grid {s.nx}x{s.ny}x{s.nz} particles {n0}; setup "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dt = scene.dt
    for k in range(warmup):
        t = time.perf_counter()
        s.substep(dt)
        s.synchronize()
        print(f"[bench] {label} warmup {k}: {time.perf_counter() - t:.3f}s iters {s.last_pcg_iters}",
              file=sys.stderr, flush=True)
    s.enable_timing(regions, sweeps_only=True)
    s.synchronize()
    particle_steps = 0
    iters = []
    t_start = time.perf_counter()
    for k in range(steps):
        particle_steps += s.particle_count()
        t = time.perf_counter()
        s.substep(dt)
        iters.append(s.last_pcg_iters)
        stt = s.state()
        print(f"[bench] {label} step {k}: {time.perf_counter() - t:.3f}s pcg iters {iters[-1]} "
              f"redistribution: active {stt.redist_active} moved {stt.redist_moved} rounds "
              f"{stt.redist_rounds}", file=sys.stderr, flush=True)
    s.synchronize()
    elapsed = time.perf_counter() - t_start
    timing = s.timing()
    s.enable_timing(False)
    breakdown = None
    if breakdown_steps > 0:
        s.enable_timing(True)
        bi = []
        for k in range(breakdown_steps):
            s.substep(dt)
            bi.append(s.last_pcg_iters)
        s.synchronize()
        breakdown = {"steps": breakdown_steps, "pcg_iters": bi, "regions": s.timing()}
        s.enable_timing(False)
        print(f"[bench] {label} breakdown: {breakdown_steps} more substeps, pcg iters {bi}",
              file=sys.stderr, flush=True)
    return s, elapsed, iters, particle_steps, timing, breakdown


SOLVE_REGIONS = ["rhs", "pcg_setup", "mic_build", "mic_fwd", "mic_fwd_dot", "mic_fwd_axpy", "mic_bwd", "mic_bwd_dot",
                 "spmv_dot",
                 "spmv_update",
                 "axpy_norm", "dot_update", "update_s", "gradient", "refine"]


def solve_error(s, scene, rel_ref, rel_tol, f32):
    """Pressure and divergence error of the solver's projection of one substep against an
    f64 solve of the same inputs (the f64 device path, which reproduces the CPU oracle
    bit for bit in reference mode) run to a relative residual of `rel_ref`."""
    from sdr_amd import _abi as A
    from sdr_amd.solver import SolveParams, project_capturing_with_source

    dt = scene.dt
    for st in ("emit", "classify", "p2g", "save", "forces"):
        s.run_stage(st, dt)
    u, v, w = s.grid()
    cells = s.cells()
    # The solver's projection applies the volume-control source of its particle counts
    # (solver.rs:392-423); the reference solve gets the same source explicitly.
    # cell_of (solver.rs:897-908) in the solver's own arithmetic: with f32 storage a
    # particle within rounding distance of a cell face counts where the device puts it
    pos, _ = s.particles()
    ft = np.float32 if f32 else np.float64
    g = (pos.astype(ft) - np.asarray(s.origin).astype(ft)) / ft(s.dx)
    ok = np.all(g >= 0, axis=1)
    ijk = g.astype(np.int64)
    ok &= (ijk[:, 0] < s.nx) & (ijk[:, 1] < s.ny) & (ijk[:, 2] < s.nz)
    cnt = np.bincount(((ijk[ok, 2] * s.ny + ijk[ok, 1]) * s.nx + ijk[ok, 0]),
                      minlength=s.nx * s.ny * s.nz).reshape(s.nz, s.ny, s.nx)
    del pos, g, ijk, ok
    fp = scene.fluid
    rest = float(max(fp.particles_per_cell, 1))
    unit = s.dx / dt
    over = cnt / rest - 1.0
    src = np.where((cells == 1) & (over > 0.0),
                   np.minimum(fp.volume_stiffness * over * unit, 0.25 * unit), 0.0)
    if fp.volume_stiffness <= 0.0:
        src = None
    t = time.perf_counter()
    uo, vo, wo = u.copy(), v.copy(), w.copy()
    rep, p_ref = project_capturing_with_source(
        s.nx, s.ny, s.nz, s.dx, s.origin, uo, vo, wo, cells, scene.fluid.density, dt,
        SolveParams(20000, 0.0), src, dot_mode=A.DOT_FAST, solve_mode=A.SOLVE_RELATIVE,
        refine_tol=rel_ref)
    t_ref = time.perf_counter() - t
    # the plain f64 path stopped by the same rule as the solver under test
    u8, v8, w8 = u.copy(), v.copy(), w.copy()
    rep8, p8 = project_capturing_with_source(
        s.nx, s.ny, s.nz, s.dx, s.origin, u8, v8, w8, cells, scene.fluid.density, dt,
        SolveParams(20000, 0.0), src, dot_mode=A.DOT_FAST, solve_mode=A.SOLVE_RELATIVE,
        refine_tol=rel_tol)
    s.run_stage("project", dt)
    p = s.pressure()
    ud, vd, wd = s.grid()
    fl = cells == 1

    def div(a, b, c):
        d = (a[:, :, 1:] - a[:, :, :-1]) + (b[:, 1:, :] - b[:, :-1, :]) + (c[1:] - c[:-1])
        return np.abs(d[fl]).max()

    d_in = div(u, v, w)
    return {
        "reference": f"f64 device solve of the same substep to relative residual {rel_ref:g} "
                     f"({rep.iters} iterations, {t_ref:.1f} s)",
        "pressure_rel_err": float(np.abs(p - p_ref).max() / max(np.abs(p_ref).max(), 1e-300)),
        "divergence_in": float(d_in),
        "divergence_out": float(div(ud, vd, wd)),
        "divergence_out_reference": float(div(uo, vo, wo)),
        "face_rel_err": float(max(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
                                  for a, b in ((ud, uo), (vd, vo), (wd, wo)))),
        "fp64_same_rule_pressure_rel_err": float(np.abs(p8 - p_ref).max() /
                                                 max(np.abs(p_ref).max(), 1e-300)),
        "fp64_same_rule_iters": int(rep8.iters),
    }


def copy_calibration(nbytes: int) -> None:
    """Calibration copies for rocprofv3 --pmc:
      * cold: a device-to-device copy (HIP runtime copy kernel) of `nbytes` from a
        buffer that a later 2 x nbytes fill pushed out of the 256 MB Infinity Cache, so
        FETCH_SIZE reads the whole source from HBM;
      * warm: a 256 MiB copy by a plain elementwise kernel (one 16-byte load and one
        16-byte store per lane) from a buffer written just before, which the Infinity
        Cache still holds: FETCH_SIZE reads (almost) nothing, so the counter measures HBM
        reads behind that cache, not the bytes a kernel requests; its WRITE_SIZE reads
        the bytes written."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b, e = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(e), C.c_size_t(2 * nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(e, 2, C.c_size_t(2 * nbytes)) == 0  # evicts a from the Infinity Cache
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    for q in (a, b, e):
        hip.hipFree(q)
    import torch

    x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y


def pmc_traffic(kernel_region: str):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc
    summary (reads from the TCC_EA0_RDREQ 32/64/128-byte request counts, writes from
    WRITE_SIZE; both read 1.00x on a cold 1 GiB copy), if it covers this region; None
    otherwise."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    r = d.get("regions", {}).get(kernel_region)
    keys = ("bytes_per_launch", "read_bytes", "write_bytes", "algorithmic_read_bytes", "algorithmic_write_bytes")
    if not r or any(k not in r for k in keys):
        return None
    return {"bytes_per_launch": r["bytes_per_launch"], "read_bytes": r["read_bytes"],
            "write_bytes": r["write_bytes"], "algorithmic_read_bytes": r["algorithmic_read_bytes"],
            "algorithmic_write_bytes": r["algorithmic_write_bytes"], "source": "profiles/pmc_traffic.json"}


def mesh_sdf_run(subdivisions: int = 6, n: int = 96, cpu_nodes: int = 2048) -> dict:
    """The mesh -> SDF build (`Sdf::from_mesh`, sdf.rs:47-86) of a closed icosphere of
    20 * 4^s triangles on an n^3 node grid, on the device, with the CPU oracle timed on
    `cpu_nodes` nodes of the same grid (the build is O(nodes x triangles)).  FP64-bound:
    about 120 flops per node-triangle pair (point-triangle distance and solid angle)."""
    import oracle
    from sdr_amd import scenes
    from sdr_amd.solver import Aabb, Sdf

    v, f = scenes.icosphere(1.0, subdivisions)
    dx = 2.6 / (n - 1)
    box = Aabb((-1.3, -1.3, -1.3), (-1.3 + (n - 1) * dx,) * 3)
    Sdf.from_mesh(v, f, dx, bounds=box)  # warm-up (code object load)
    t = time.perf_counter()
    s = Sdf.from_mesh(v, f, dx, bounds=box)
    gpu_s = time.perf_counter() - t
    nn = s.nx * s.ny * s.nz
    pairs = nn * f.shape[0]
    # the oracle on a row of the same grid
    lib = oracle.load()
    row = Aabb((-1.3, 0.0, 0.0), (-1.3 + (cpu_nodes - 1) * (2.6 / (cpu_nodes - 1)), 0.0, 0.0))
    t = time.perf_counter()
    Sdf.from_mesh(v, f, 2.6 / (cpu_nodes - 1), bounds=row, lib=lib)
    cpu_s = time.perf_counter() - t
    flops = 120.0 * pairs
    out = {"triangles": int(f.shape[0]), "nodes": int(nn), "gpu_s": round(gpu_s, 4),
           "node_triangle_pairs_per_s": round(pairs / gpu_s, 1),
           "roofline": {"bound": "fp64", "achieved": round(flops / gpu_s / 1e12, 3), "peak": 78.6,
                        "unit": "TFLOP/s", "frac": round(flops / gpu_s / 1e12 / 78.6, 4),
                        "flops_per_pair": 120},
           "cpu_oracle_pairs_per_s": round(cpu_nodes * f.shape[0] / cpu_s, 1),
           "cpu_threads": oracle.get_threads(),
           "speedup_vs_cpu": round((pairs / gpu_s) / (cpu_nodes * f.shape[0] / cpu_s), 1),
           "workload": f"icosphere of {f.shape[0]} triangles, {s.nx}^3 nodes (host-timed call incl. "
                       f"copies); oracle on {cpu_nodes} nodes"}
    print(f"[bench] mesh_sdf: {out}", file=sys.stderr, flush=True)
    return out


def surface_run(s, ms_per_step: float, reps: int = 3) -> dict:
    """Particle surfacing of the run's final state on the device: the particle distance
    field (balls of radius dx / cbrt(ppc) on nodes at the MAC cell corners) and naive
    surface nets (src/surface_nets.rs:48-221), as a run surfacing every 10 substeps does.
    The first call cell-sorts the particles (the next substep's classify re-sorts them
    anyway); later calls reuse the order."""
    t = time.perf_counter()
    info = s.surface()
    first_s = time.perf_counter() - t
    t = time.perf_counter()
    for _ in range(reps):
        info = s.surface()
    per_s = (time.perf_counter() - t) / reps
    nodes = int(info["nx"] * info["ny"] * info["nz"])
    out = {"ms_first_call": round(first_s * 1e3, 3), "ms": round(per_s * 1e3, 3), "nodes": nodes,
           "vertices": int(info["vertices"]), "triangles": int(info["triangles"]),
           "area_m2": info["area"], "volume_ml": info["volume"] * 1e6,
           "nodes_per_s": round(nodes / per_s, 1),
           "overhead_every_10_substeps": round(first_s * 1e3 / (10.0 * ms_per_step), 5),
           "workload": f"{info['nx']}x{info['ny']}x{info['nz']} nodes, r = {info['radius'] * 1e3:.4f} mm, "
                       "host-timed call incl. the count read-backs"}
    print(f"[bench] surface: {out}", file=sys.stderr, flush=True)
    return out


def serving_run(count: int, slots: int) -> dict:
    """The serving form: `count` irrigation scenarios (seeded Poisson arrivals at 8/s,
    grids 32^3-128^3, needle angle +-30 deg, 1-6 ml/s, 0.5-2 ml) through the native
    serving loop under a 200 GB footprint budget; p50/p99 latency is finish - arrival."""
    from sdr_amd import serving

    specs = serving.serving_workload(count)
    t = time.perf_counter()
    recs = serving.serve(specs, budget_bytes=200e9, max_slots=slots)
    out = serving.summarize(recs)
    # the full dose of a request is dose / flow seconds of injection at its fixed dt
    full = [s.dose_ml / s.flow_ml_s / serving.scenario_dt(s) for s in specs]
    out.update({"slots": slots, "budget_gb": 200, "wall_s": round(time.perf_counter() - t, 2),
                "resident_pools": True,
                "truncated": True,
                "substeps_run": int(sum(s.substeps for s in specs)),
                "substeps_full_dose": int(sum(full)),
                "workload": f"{count} scenarios, Poisson 8/s, n^3 with n in [32,128], f32 + f64 "
                            f"refinement to 1e-6; each request runs the first 8-40 substeps of its "
                            f"injection (dose/flow scaled), not its full dose of "
                            f"{int(min(full))}-{int(max(full))} substeps"})
    print(f"[bench] serving: {out}", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c3a", "c5"])
    ap.add_argument("--precision", default="f32", choices=["f64", "f32"])
    ap.add_argument("--dot", default="fast", choices=["exact", "fast"])
    ap.add_argument("--solve", default="relative", choices=["relative", "reference"],
                    help="relative: stop at --tol * max|b| (f32: by f64 refinement); "
                         "reference: the reference's absolute tol / 200 iterations")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iters", type=int, default=2000)
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain f64 comparison run")
    ap.add_argument("--no-error", action="store_true", help="skip the error-vs-f64 solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-sample", action="store_true",
                    help="skip the composed CPU estimate of the headline workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the 64-scenario serving run")
    ap.add_argument("--no-regions", action="store_true",
                    help="time the substeps without the per-region HIP events (no roofline)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 256 MiB device-to-device copy (the reference kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    ap.add_argument("--no-mesh", action="store_true", help="skip the mesh -> SDF build")
    ap.add_argument("--no-surface", action="store_true", help="skip the particle surfacing")
    ap.add_argument("--serve-count", type=int, default=64)
    ap.add_argument("--serve-slots", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the substep runs on one "
              "device; scenarios would run one per GPU, and a single large grid would split "
              "into k-slabs with halo exchange). Use --gpus 1.", file=sys.stderr)
        return 2

    import sdr_amd
    from sdr_amd import _abi as A
    from sdr_amd import scenes

    ndev, arch = sdr_amd.device_info()
    if args.copy_calibration:
        copy_calibration(1 << 30)
    if ndev < 1:
        print("no HIP device", file=sys.stderr)
        return 3
    scene = scenes.cavity_scene(args.config)
    prec = A.PRECISION_F32 if args.precision == "f32" else A.PRECISION_F64
    dot_mode = A.DOT_FAST if args.dot == "fast" else A.DOT_EXACT
    smode = A.SOLVE_RELATIVE if args.solve == "relative" else A.SOLVE_REFERENCE
    s, elapsed, iters, particle_steps, timing, bd = run_path(
        scene, prec, dot_mode, smode, args.tol, args.max_iters, args.steps, args.warmup,
        args.precision, regions=not args.no_regions, breakdown_steps=0 if args.no_regions else args.steps)
    ms_per_step = elapsed * 1e3 / args.steps
    value = particle_steps / elapsed
    sz = 4 if prec == A.PRECISION_F32 else 8
    nc = s.nx * s.ny * s.nz
    nsdf = scene.sdf.nx * scene.sdf.ny * scene.sdf.nz
    npart = particle_steps / max(args.steps, 1)
    bd_regions = bd["regions"] if bd else {}
    bd_steps = bd["steps"] if bd else 1
    pressure_ms = sum(bd_regions.get(r, (0.0, 0))[0] for r in SOLVE_REGIONS) / bd_steps
    # Dominant kernel region by total time; roofline from its average call duration.
    best, best_t = None, -1.0
    for name, (ms, calls) in timing.items():
        if region_bytes(name, sz, nc, 0, 0, 0) is None:
            continue
        if ms > best_t:
            best, best_t = name, ms
    roof = None
    if best is not None:
        ms, calls = timing[best]
        per_call_s = ms / calls / 1e3
        by = region_bytes(best, sz, nc, nc, npart, nsdf)
        ach = by / per_call_s / 1e9
        # the committed counter passes cover the default 256^3 workload only
        tr = pmc_traffic(best) if args.config == "c3" else None
        roof = {"bound": "hbm", "kernel": best, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": tr["bytes_per_launch"] if tr else None, "traffic_detail": tr,
                "bytes_per_call": int(by), "avg_call_ms": round(per_call_s * 1e3, 4)}
    surf = None
    if not args.no_surface:
        try:
            surf = surface_run(s, ms_per_step)
        except Exception as e:  # the throughput number stands on its own
            surf = {"error": repr(e)}
    err = None
    if not args.no_error and smode == A.SOLVE_RELATIVE:
        try:
            err = solve_error(s, scene, 1e-12, args.tol, prec == A.PRECISION_F32)
        except Exception as e:  # the throughput number stands on its own
            err = {"error": repr(e)}
    s.close()
    fp64 = None
    if not args.no_fp64 and prec == A.PRECISION_F32:
        s64, el64, it64, ps64, tm64, bd64 = run_path(scene, A.PRECISION_F64, dot_mode, smode, args.tol,
                                                     args.max_iters, args.steps, args.warmup, "f64",
                                                     breakdown_steps=args.steps)
        s64.close()
        fp64 = {"ms_per_step": round(el64 * 1e3 / args.steps, 3),
                "value": round(ps64 / el64, 1),
                "pressure_solve_ms": round(sum(bd64["regions"].get(r, (0.0, 0))[0] for r in SOLVE_REGIONS)
                                           / bd64["steps"], 3),
                "pcg_iters": [int(i) for i in it64],
                "breakdown_pcg_iters": [int(i) for i in bd64["pcg_iters"]],
                "speedup_of_f32": round((ps64 / el64) and value / (ps64 / el64), 3)}
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline_c1(threads)
        except Exception as e:  # the GPU number stands on its own
            cpu = {"error": repr(e)}
        if args.config == "c3" and not args.no_cpu_sample:
            try:
                cpu["c3_composed_sample"] = cpu_baseline(scene, scene.sdf, float(np.mean(iters)), 20.0, threads)
            except Exception as e:
                cpu["c3_composed_sample"] = {"error": repr(e)}
    serve = None
    if not args.no_serve:
        try:
            serve = serving_run(args.serve_count, args.serve_slots)
        except Exception as e:  # the throughput number stands on its own
            serve = {"error": repr(e)}
    mesh = None
    if not args.no_mesh:
        try:
            mesh = mesh_sdf_run()
        except Exception as e:  # the throughput number stands on its own
            mesh = {"error": repr(e)}
    prec_txt = ("f32 storage + f64 dots/norms/refinement" if prec == A.PRECISION_F32 else "f64")
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == A.PRECISION_F32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: 256^3 MAC grid rounded cavity + socket, "
                               f"~{npart / 1e6:.0f}M particles 8/cell, fixed dt {scene.dt:g} s"
                   if args.config == "c3" else
                   (f"c5: 512x512x384 MAC grid noisy ellipsoid + two sockets, ~{npart / 1e6:.0f}M "
                    f"particles 8/cell, fixed dt {scene.dt:g} s" if args.config == "c5" else args.config),
                   "grid": [s.nx, s.ny, s.nz], "particles": int(npart), "dt": scene.dt,
                   "precision": prec_txt, "dot": args.dot,
                   "solve": (f"relative {args.tol:g} (max {args.max_iters} iterations)"
                             if smode == A.SOLVE_RELATIVE else
                             f"reference tol {scene.solve.tol:g} / {scene.solve.max_iters}")},
        "pressure_solve_ms": round(pressure_ms, 3),
        "pcg_iters": [int(i) for i in iters],
        "breakdown_pcg_iters": [int(i) for i in bd["pcg_iters"]] if bd else None,
        "fp64_gpu": fp64,
        "error_vs_fp64": err,
        "regions_ms_per_step": {k: round(v[0] / bd_steps, 4) for k, v in sorted(bd_regions.items())},
        "roofline": roof,
        "cpu_baseline": cpu,
        "serving": serve,
        "mesh_sdf": mesh,
        "surface": surf,
    }
    if cpu and "value" in cpu:
        # like for like: the device on the CPU run's own workload (C1); the headline
        # workload's CPU figure is the composed estimate in cpu_baseline.c3_composed_sample
        out["speedup_vs_cpu"] = cpu["speedup_same_run"]
        c3 = cpu.get("c3_composed_sample") or {}
        if "value" in c3:
            out["speedup_vs_cpu_c3_composed"] = round(value / c3["value"], 1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
