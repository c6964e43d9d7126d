"""The device's particle sort by cell against numpy.

After `classify` the device keeps its particles in slots ordered by clamped cell
(grid.rs:52-88: every particle scatters through it), and inside a cell by reference index
in exact mode, by previous slot otherwise (a stable sort).  The sort is a counting sort:
a histogram and a scatter with one atomic per run of equal cells in a wavefront, then
each cell's segment put in order — by one thread in registers up to 16 entries, by a
workgroup (bitonic runs of 2048 in LDS, then merge passes) above.  The segment sizes here
cover every one of those paths: 1, 8, 16, 17, 300, 2048, 2049, 5000 and 12000 particles
in one cell, in shuffled reference order, beside ordinary cells."""
import numpy as np
import pytest

from sdr_amd import _abi as A
from sdr_amd import scenes

pytestmark = pytest.mark.gpu
SIZES = (1, 8, 16, 17, 300, 2048, 2049, 5000, 12000)


def clamped_cells(s, pos):
    """The clamped cell of each position, in f64 as the device computes it."""
    loc = (pos - np.asarray(s.origin)) / s.dx
    ijk = np.where(loc > 0, np.floor(loc), 0).astype(np.int64)
    ijk = np.minimum(ijk, np.array([s.nx - 1, s.ny - 1, s.nz - 1]))
    return (ijk[:, 2] * s.ny + ijk[:, 1]) * s.nx + ijk[:, 0]


def slot_ranks(s):
    import ctypes as C

    out = np.zeros(s.particle_count(), np.uint32)
    s.lib.solver_get_slot_ranks(s._h, out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size)
    return out


def clumped(s, seed):
    """Particles well inside their cells (|offset| <= 0.4 dx): one clump of each size in
    SIZES on cells spread over the grid, 0-12 particles in every other cell, shuffled."""
    g = np.random.default_rng(seed)
    nc = s.nx * s.ny * s.nz
    cnt = g.integers(0, 13, size=nc)
    big = g.choice(nc, size=len(SIZES), replace=False)
    cnt[big] = SIZES
    cell = np.repeat(np.arange(nc), cnt)
    ijk = np.stack([cell % s.nx, (cell // s.nx) % s.ny, cell // (s.nx * s.ny)], 1).astype(np.float64)
    pos = np.asarray(s.origin) + (ijk + 0.5 + g.uniform(-0.4, 0.4, ijk.shape)) * s.dx
    perm = g.permutation(pos.shape[0])
    return pos[perm], cell[perm]


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_cell_sort_orders_segments_of_every_size(dev, mode):
    dot = A.DOT_EXACT if mode == "exact" else A.DOT_FAST
    s = scenes.make_solver(scenes.box_scene(), dot_mode=dot)
    pos, cell = clumped(s, 11)
    assert np.array_equal(clamped_cells(s, pos), cell)
    s.set_particles(pos, np.zeros_like(pos))
    s.run_stage("classify")
    got = slot_ranks(s)
    # set_particles numbers the particles in input order: rank = index = previous slot
    want = np.lexsort((np.arange(pos.shape[0]), cell)).astype(np.uint32)
    assert np.array_equal(got, want)


def test_fast_sort_is_stable_in_the_previous_slot_order(dev):
    """Fast mode keeps the previous slot order inside a cell, which after redistribution
    (moved particles keep their slots in the cell they left) is not reference order; exact
    mode goes by reference index.  Both against numpy's stable orders."""
    out = {}
    for mode, dot in (("exact", A.DOT_EXACT), ("fast", A.DOT_FAST)):
        s = scenes.make_solver(scenes.box_scene(), dot_mode=dot)
        pos, cell = clumped(s, 12)
        s.set_particles(pos, np.zeros_like(pos))
        s.run_stage("classify")
        before = slot_ranks(s)
        s.redistribute()
        assert s.state().redist_moved > 0
        p2, _ = s.particles()  # reference order
        c2 = clamped_cells(s, p2)
        s.run_stage("classify")
        got = slot_ranks(s)
        if mode == "exact":
            want = np.lexsort((np.arange(p2.shape[0]), c2))
        else:
            prev_slot = np.empty(p2.shape[0], np.int64)
            prev_slot[before] = np.arange(p2.shape[0])
            want = np.lexsort((prev_slot, c2))
        assert np.array_equal(got, want.astype(np.uint32)), mode
        out[mode] = got
    assert not np.array_equal(out["exact"], out["fast"])
