"""The serving loop (sdr_serve): workload generation and the footprint estimate on the
CPU; admission, overlap and 'the same result alone and in the mix' on the GPU."""
import ctypes as C

import numpy as np
import pytest

import sdr_amd
from sdr_amd import _abi as A
from sdr_amd import serving


def test_workload_is_seeded_and_in_range():
    a = serving.serving_workload(64, seed=7)
    b = serving.serving_workload(64, seed=7)
    assert a == b
    assert serving.serving_workload(64, seed=8) != a
    arr = np.array([s.arrival_s for s in a])
    assert np.all(np.diff(arr) > 0)
    # Poisson arrivals at 8/s: 64 arrivals take about 8 s
    assert 4.0 < arr[-1] < 14.0
    for s in a:
        assert 32 <= s.n <= 128
        assert -30.0 <= s.angle_deg <= 30.0
        assert 1.0 <= s.flow_ml_s <= 6.0 and 0.5 <= s.dose_ml <= 2.0
        assert 8 <= s.substeps <= 40
        assert s.substeps == serving.substeps_for(s.dose_ml, s.flow_ml_s)


def test_substeps_grow_with_injection_time():
    assert serving.substeps_for(0.5, 6.0) == 8
    assert serving.substeps_for(2.0, 1.0) == 40
    assert serving.substeps_for(1.0, 1.0) > serving.substeps_for(1.0, 2.0)


def test_footprint_estimate_scales_with_cells_particles_and_precision():
    lib = sdr_amd.load_library()

    def fp(n, parts, prec):
        out = C.c_uint64()
        lib.solver_footprint(n, n, n, parts, prec, C.byref(out))
        return out.value

    assert fp(64, 10**6, A.PRECISION_F32) < fp(128, 10**6, A.PRECISION_F32)
    assert fp(64, 10**6, A.PRECISION_F32) < fp(64, 10**7, A.PRECISION_F32)
    assert fp(64, 10**6, A.PRECISION_F32) < fp(64, 10**6, A.PRECISION_F64)
    # the largest serving scenario (128^3, ~16M particles) is a few GB, far below 200 GB
    assert 1e9 < fp(128, 16 * 10**6, A.PRECISION_F32) < 10e9


def _small_mix():
    specs = []
    for q, (n, ang, flow, arr) in enumerate([(16, 10.0, 3.0, 0.0), (20, -20.0, 5.0, 0.0),
                                             (24, 0.0, 1.5, 0.0), (18, 25.0, 4.0, 0.01),
                                             (22, -5.0, 2.0, 0.02), (16, 15.0, 6.0, 0.03)]):
        specs.append(serving.ScenarioSpec(n, ang, flow, 1.0, arr, 100 + q, 3))
    return specs


@pytest.mark.gpu
def test_each_result_is_the_same_alone_and_in_the_mix(dev):
    """Six scenarios served three at a time give, bit for bit, the particles each gives
    when served alone (no float atomics, no state shared between scenarios)."""
    specs = _small_mix()
    mix = serving.serve(specs, max_slots=3)
    assert all(r.status == A.SDR_OK for r in mix), [r.status for r in mix]
    assert serving.max_in_flight(mix) >= 2
    for spec, r in zip(specs, mix):
        alone = serving.serve([spec], max_slots=1)[0]
        assert alone.status == A.SDR_OK
        assert alone.particles == r.particles > 0
        assert alone.checksum == r.checksum
        assert alone.particle_substeps == r.particle_substeps
        assert alone.pcg_iters == r.pcg_iters


@pytest.mark.gpu
def test_admission_respects_slots_budget_and_rejects_oversized(dev):
    specs = _small_mix()
    recs = serving.serve(specs, max_slots=2)
    assert serving.max_in_flight(recs) <= 2
    # a budget that holds one scenario at a time serialises them in arrival order
    big = max(r.footprint_bytes for r in recs)
    one = serving.serve(specs, budget_bytes=1.5 * big, max_slots=8)
    assert all(r.status == A.SDR_OK for r in one)
    assert serving.max_in_flight(one) == 1
    order = sorted(one, key=lambda r: r.admit_s)
    assert [r.spec.arrival_s for r in order] == sorted(r.spec.arrival_s for r in order)
    for a, b in zip(order, order[1:]):
        assert b.admit_s >= a.finish_s
    # a scenario that can never fit is rejected, the others still run
    small = min(r.footprint_bytes for r in recs)
    rej = serving.serve(specs, budget_bytes=0.99 * big if big > small else big, max_slots=8)
    assert any(r.status == A.SDR_E_OOM for r in rej) or big == small
    assert all(r.status in (A.SDR_OK, A.SDR_E_OOM) for r in rej)
    assert sum(r.status == A.SDR_OK for r in rej) >= 1
    s = serving.summarize(rej)
    assert s["completed"] + s["rejected"] == len(specs)


@pytest.mark.gpu
def test_resident_pools_reuse_memory_and_keep_results(dev):
    """With resident pools a slot's later requests take their buffers from the blocks its
    earlier ones released (few or no new device allocations), and every result is the
    same as with a fresh allocation per request."""
    specs = _small_mix()
    pooled = serving.serve(specs, max_slots=2, resident_pools=True)
    fresh = serving.serve(specs, max_slots=2, resident_pools=False)
    assert all(r.status == A.SDR_OK for r in pooled + fresh)
    for a, b in zip(pooled, fresh):
        assert (a.checksum, a.particles, a.pcg_iters) == (b.checksum, b.particles, b.pcg_iters)
        assert b.device_allocs == 0  # no pool: nothing is counted
    by_slot = {}
    for r in sorted(pooled, key=lambda r: r.admit_s):
        by_slot.setdefault(r.slot, []).append(r)
    firsts = [rs[0].device_allocs for rs in by_slot.values()]
    laters = [r.device_allocs for rs in by_slot.values() for r in rs[1:]]
    assert min(firsts) > 20 and laters
    assert max(laters) < min(firsts) / 2
