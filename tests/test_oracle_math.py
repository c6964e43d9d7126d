"""Pins of the CPU oracle against what the mathematics fixes.

Where the reference has no test for a function, the oracle is checked against closed
forms, published test vectors, scipy/dense linear algebra and independent numpy
restatements on tiny inputs, each chosen so a plausible slip (dropped term, wrong
index or sign, transposed operand, off-by-one chunk) fails.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from sdr_amd import scenes
from sdr_amd.solver import AIR, FLUID, SOLID, SolveParams, project_capturing_with_source

from helpers import bits_equal, classify_box, fill_divergent

RNG = np.random.default_rng(1234)


def random_cells(n, p_fluid=0.7, p_solid=0.15, seed=0):
    g = np.random.default_rng(seed)
    r = g.random((n, n, n))
    c = np.where(r < p_fluid, FLUID, np.where(r < p_fluid + p_solid, SOLID, AIR)).astype(np.uint8)
    return c


def dense_laplacian(cells):
    """The 7-point Laplacian of pressure.rs:171-223 built independently: for fluid rows,
    diag = number of non-solid neighbours (outside = solid), -1 to fluid neighbours."""
    nz, ny, nx = cells.shape
    n = cells.size
    A = np.zeros((n, n))
    idx = lambda i, j, k: (k * ny + j) * nx + i
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if cells[k, j, i] != FLUID:
                    continue
                c = idx(i, j, k)
                for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    a, b, e = i + di, j + dj, k + dk
                    if 0 <= a < nx and 0 <= b < ny and 0 <= e < nz:
                        kind = cells[e, b, a]
                        if kind != SOLID:
                            A[c, c] += 1.0
                        if kind == FLUID:
                            A[c, idx(a, b, e)] = -1.0
    return A


# ---------------------------------------------------------------- Laplacian + SpMV
def test_laplacian_matches_independent_assembly_and_is_symmetric():
    cells = random_cells(6, seed=3)
    d, px, py, pz = oracle.laplacian(cells)
    A = dense_laplacian(cells)
    assert np.array_equal(np.diag(A), d.reshape(-1))
    assert np.array_equal(A, A.T)
    x = RNG.standard_normal(cells.shape) * (cells == FLUID)
    y = oracle.apply_a(cells, x)
    assert np.allclose(y.reshape(-1), A @ x.reshape(-1), rtol=1e-13, atol=1e-13)
    assert set(np.unique(A)) <= {-1.0, 0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0}


# ---------------------------------------------------------------- MIC(0)
def _factor(cells, pc):
    """M = L L^T with L = D + L_A D^-1, D = diag(1/precon) over the fluid unknowns: the
    matrix the forward/backward solves of pressure.rs:517-570 invert."""
    A = dense_laplacian(cells)
    f = (cells == FLUID).reshape(-1)
    A = A[np.ix_(f, f)]
    p = pc.reshape(-1)[f]
    D = np.diag(1.0 / p)
    L = D + np.tril(A, -1) @ np.diag(p)
    return A, L @ L.T


def test_ic0_identity_at_tau_zero():
    """With the modification switched off the factor is IC(0): diag(M) = diag(A) and M
    equals A on A's sparsity pattern."""
    cells = random_cells(6, seed=5)
    pc, _ = oracle.mic0(cells, tau=0.0)
    A, M = _factor(cells, pc)
    assert np.allclose(np.diag(M), np.diag(A), rtol=1e-12)
    mask = A != 0
    assert np.allclose(M[mask], A[mask], rtol=1e-12, atol=1e-12)


def test_mic_row_sum_identity_at_tau_one():
    """Full modification (tau = 1) moves the dropped fill-in onto the diagonal, so M and
    A have equal row sums wherever the safety floor (e < sigma*diag) did not fire."""
    cells = random_cells(6, seed=7)
    pc, _ = oracle.mic0(cells, tau=1.0)
    A, M = _factor(cells, pc)
    f = (cells == FLUID).reshape(-1)
    e = 1.0 / pc.reshape(-1)[f] ** 2
    no_floor = ~np.isclose(e, np.diag(A), rtol=0, atol=1e-12)
    rs_a, rs_m = A.sum(1), M.sum(1)
    assert no_floor.sum() > 10
    assert np.allclose(rs_m[no_floor], rs_a[no_floor], rtol=1e-10, atol=1e-10)


def test_mic0_apply_inverts_the_factor():
    cells = random_cells(6, seed=9)
    r = RNG.standard_normal(cells.shape) * (cells == FLUID)
    pc, z = oracle.mic0(cells, r)
    _, M = _factor(cells, pc)
    f = (cells == FLUID).reshape(-1)
    assert np.allclose(M @ z.reshape(-1)[f], r.reshape(-1)[f], rtol=1e-11, atol=1e-11)
    assert np.all(z.reshape(-1)[~f] == 0.0)


# ---------------------------------------------------------------- dot (pressure.rs:588-603)
def test_dot_uses_4096_element_chunks():
    """1e16 absorbs +1.0 within its own chunk but a fresh chunk sums the ones exactly, so
    the result pins the chunk length: (1e16) + 4096 only for CHUNK = 4096."""
    a = np.ones(2 * 4096)
    a[0] = 1e16
    b = np.ones_like(a)
    assert oracle.dot(a, b) == 1e16 + 4096.0
    a2 = np.ones(4096 + 7)
    assert oracle.dot(a2, a2) == 4103.0
    x = RNG.standard_normal(50_000)
    assert abs(oracle.dot(x, x) - float(np.dot(x, x))) < 1e-9 * float(np.dot(x, x))


# ---------------------------------------------------------------- PCG + projection
def test_pcg_matches_scipy_direct_solve():
    """On a tiny grid the projection's pressure solves A p = b (b = -(rho dx/dt)(div - s))
    to the PCG tolerance: compare with scipy's sparse direct solver."""
    n = 7
    cells = random_cells(n, p_fluid=0.75, p_solid=0.1, seed=11)
    u, v, w = fill_divergent(n, n, n)
    src = RNG.random((n, n, n)) * 0.01
    u0, v0, w0 = u.copy(), v.copy(), w.copy()
    rho, dt, dx = 1000.0, 1e-3, 0.01
    rep, p = project_capturing_with_source(n, n, n, dx, (0, 0, 0), u, v, w, cells, rho, dt,
                                           SolveParams(500, 1e-12), src, lib=oracle.load())
    # independent RHS: zero faces touching solid / the boundary, then the divergence
    solid = np.pad(cells == SOLID, 1, constant_values=True)
    su = solid[1:-1, 1:-1, :-1] | solid[1:-1, 1:-1, 1:]
    sv = solid[1:-1, :-1, 1:-1] | solid[1:-1, 1:, 1:-1]
    sw = solid[:-1, 1:-1, 1:-1] | solid[1:, 1:-1, 1:-1]
    u0[su] = 0.0
    v0[sv] = 0.0
    w0[sw] = 0.0
    div = (u0[:, :, 1:] - u0[:, :, :-1]) + (v0[:, 1:, :] - v0[:, :-1, :]) + (w0[1:] - w0[:-1])
    b = -(rho * dx / dt) * (div - src)
    f = (cells == FLUID).reshape(-1)
    A = dense_laplacian(cells)[np.ix_(f, f)]
    x = spla.spsolve(sp.csr_matrix(A), b.reshape(-1)[f])
    assert rep.iters < 500
    assert np.allclose(p.reshape(-1)[f], x, rtol=1e-9, atol=1e-9 * np.abs(x).max())
    assert np.all(p.reshape(-1)[~f] == 0.0)
    # gradient step (pressure.rs:285-332) on one interior x-face between fluid cells
    coef = dt / (rho * dx)
    fi = np.argwhere((cells[:, :, 1:] == FLUID) & (cells[:, :, :-1] == FLUID))[0]
    k, j, i = fi[0], fi[1], fi[2] + 1
    assert u[k, j, i] == u0[k, j, i] - coef * (p[k, j, i] - p[k, j, i - 1])


def test_zero_rhs_returns_zero_iterations():
    n = 8
    cells = classify_box(n, n, n)
    u = np.zeros((n, n, n + 1))
    v = np.zeros((n, n + 1, n))
    w = np.zeros((n + 1, n, n))
    rep, p = project_capturing_with_source(n, n, n, 0.01, (0, 0, 0), u, v, w, cells, 1000.0, 1e-3,
                                           lib=oracle.load())
    assert rep.iters == 0 and rep.residual == 0.0 and not p.any()


# ---------------------------------------------------------------- extrapolate
def numpy_extrapolate(field, valid, iters):
    """Vectorised restatement of the snapshot flood fill (solver.rs:960-1009)."""
    f = field.copy()
    known = valid.astype(bool).copy()
    nz, ny, nx = f.shape
    for _ in range(iters):
        snap, ks = f.copy(), known.copy()
        s = np.zeros_like(f)
        cnt = np.zeros(f.shape, np.int64)
        for ax, sh in ((2, -1), (2, 1), (1, -1), (1, 1), (0, -1), (0, 1)):
            nb = np.roll(snap, -sh, axis=ax)
            kn = np.roll(ks, -sh, axis=ax)
            edge = [slice(None)] * 3
            edge[ax] = slice(-1, None) if sh == 1 else slice(0, 1)
            kn[tuple(edge)] = False
            s = s + np.where(kn, nb, 0.0)
            cnt = cnt + kn
        upd = (~ks) & (cnt > 0)
        f[upd] = s[upd] / cnt[upd]
        known |= upd
    return f


def test_extrapolate_matches_vectorised_restatement():
    g = np.random.default_rng(21)
    field = g.standard_normal((6, 7, 8))
    valid = (g.random((6, 7, 8)) < 0.15).astype(np.uint8)
    got = oracle.extrapolate(field, valid, 4)
    want = numpy_extrapolate(field, valid, 4)
    assert bits_equal(got, want)


# ---------------------------------------------------------------- trilinear
def test_trilinear_reproduces_linear_fields():
    g = np.random.default_rng(5)
    nz, ny, nx = 5, 6, 7
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    a = 0.3 + 1.5 * i - 2.0 * j + 0.25 * k
    for _ in range(50):
        gx, gy, gz = g.random() * (nx - 1), g.random() * (ny - 1), g.random() * (nz - 1)
        got = oracle.trilinear_sample(a.astype(np.float64), gx, gy, gz)
        assert abs(got - (0.3 + 1.5 * gx - 2.0 * gy + 0.25 * gz)) < 1e-12
    # clamped outside the node range
    assert abs(oracle.trilinear_sample(a.astype(np.float64), -3.0, 0.0, 0.0) - 0.3) < 1e-15
    assert abs(oracle.trilinear_sample(a.astype(np.float64), 99.0, 0.0, 0.0) - (0.3 + 1.5 * 6)) < 1e-12


# ---------------------------------------------------------------- RNG / hashing
def test_xorshift_published_sequence():
    """Marsaglia's xorshift64 (13, 7, 17): first state from seed 1 is 1082269761."""
    u, state = oracle.rng_units(1, 1)
    assert state == 1082269761
    assert u[0] == (1082269761 >> 11) / float(1 << 53)
    # seed | 1 (solver.rs:123-125): an even seed behaves as seed + 1
    assert oracle.rng_units(0x5deece66c, 3)[1] == oracle.rng_units(0x5deece66d, 3)[1]


def test_hash_jitter_splitmix64_vector():
    """splitmix64 from state 0 first outputs 0xE220A8397B1DCDAF (its published vector);
    hash_jitter(0).x mixes key 0 and maps the top 53 bits to [-0.3, 0.3)."""
    z = 0xE220A8397B1DCDAF
    want = ((z >> 11) / float(1 << 53) - 0.5) * 0.6
    assert oracle.hash_jitter(0)[0] == want
    js = np.array([oracle.hash_jitter(k) for k in range(2000)])
    assert js.min() >= -0.3 and js.max() < 0.3


# ---------------------------------------------------------------- scene plumbing
def test_cavity_dims_match_the_workload():
    s = scenes.make_solver(scenes.cavity_scene("c1"), lib=oracle.load())
    assert (s.nx, s.ny, s.nz) == (36, 42, 34)
    # box_solver: bounds (-0.016 + 15*0.002) - (-0.016) = 0.030, ceil(0.030/0.002) = 15
    s2 = scenes.make_solver(scenes.box_scene(), lib=oracle.load())
    assert (s2.nx, s2.ny, s2.nz) == (15, 15, 15)


# ---------------------------------------------------------------- G2P / advect closed forms
def _box_oracle(orc, flip_ratio):
    sc = scenes.box_scene()
    sc.fluid.flip_ratio = flip_ratio
    return scenes.make_solver(sc, lib=orc)


def _interior_points(s, n, seed):
    """Points at least 2 cells inside the 12^3-cell cavity of the box scene."""
    g = np.random.default_rng(seed)
    lo = np.asarray(s.origin) + 4.0 * s.dx
    return lo + g.random((n, 3)) * (7.0 * s.dx)


def _linear_faces(s, a):
    """Face velocities of the linear field u = a0 + a1 x + a2 y + a3 z (per component),
    sampled at the face positions of the MAC grid (grid.rs:128-175)."""
    o, dx = np.asarray(s.origin), s.dx
    out = []
    for comp, shape in enumerate(s.face_shapes()):
        kk, jj, ii = np.meshgrid(*(np.arange(m, dtype=np.float64) for m in shape), indexing="ij")
        pos = [ii, jj, kk]
        x = [o[d] + (pos[d] + (0.0 if d == comp else 0.5)) * dx for d in range(3)]
        c = a[comp]
        out.append(c[0] + c[1] * x[0] + c[2] * x[1] + c[3] * x[2])
    return out


def test_g2p_pure_flip_on_an_unchanged_grid_keeps_velocities(orc):
    """solver.rs:650-677 with flip = 1 and grid == saved grid: every interpolated change
    is exactly 0, so flip_v = v and v*1 + pic*0 = v, bit for bit."""
    s = _box_oracle(orc, 1.0)
    g = np.random.default_rng(7)
    pos = _interior_points(s, 500, 8)
    vel = g.standard_normal(pos.shape)
    s.set_particles(pos, vel)
    s.set_grid(*[g.standard_normal(sh) for sh in s.face_shapes()])
    s.run_stage("save")
    s.run_stage("g2p")
    assert bits_equal(s.particles()[1], vel)


def test_g2p_pure_pic_interpolates_a_linear_field(orc):
    """flip = 0: the particle takes the PIC sample of the grid, which for a linear face
    field is the field at the particle (trilinear interpolation is exact on linears)."""
    s = _box_oracle(orc, 0.0)
    a = [(0.3, 2.0, -1.0, 0.5), (-0.2, 0.7, 1.5, -2.5), (1.1, -0.4, 0.9, 3.0)]
    pos = _interior_points(s, 500, 9)
    s.set_particles(pos, np.full(pos.shape, 9.0))
    s.set_grid(*_linear_faces(s, a))
    s.run_stage("save")
    s.run_stage("g2p")
    v = s.particles()[1]
    for comp in range(3):
        c = a[comp]
        want = c[0] + c[1] * pos[:, 0] + c[2] * pos[:, 1] + c[3] * pos[:, 2]
        assert np.allclose(v[:, comp], want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("speed_cells", [0.3, 3.0])
def test_advect_uniform_flow_is_a_clamped_translation(orc, speed_cells):
    """solver.rs:693-711: in a uniform field both RK2 samples equal the field, so the
    step is U*dt, clamped to length dx when longer; interior points need no projection."""
    s = _box_oracle(orc, 0.95)
    dt = 1e-3
    U = np.array([0.6, -0.8, 0.0]) * speed_cells * s.dx / dt
    pos = _interior_points(s, 300, 10)
    s.set_particles(pos, np.zeros_like(pos))
    s.set_grid(*[np.full(sh, U[c]) for c, sh in enumerate(s.face_shapes())])
    s.run_stage("advect", dt)
    step = U * dt
    if np.linalg.norm(step) > s.dx:
        step = step * (s.dx / np.linalg.norm(step))
    assert np.allclose(s.particles()[0], pos + step, rtol=0, atol=1e-12 * s.dx / 1e-3)
