"""Pins of the oracle's particle surfacing: naive surface nets (src/surface_nets.rs:48-221)
and the particle distance field.

Checks independent of the restatement:
  * the reference's own test `meshes_a_sphere` (surface_nets.rs:228-248);
  * closed forms: one negative node gives 8 vertices at (5/6, 5/6, 5/6)-offsets and 12
    triangles; a plane gives (nx-2)(ny-2) unit quads at the plane's height, facing +z;
  * topology: every edge of a closed net is shared by two triangles (an even number
    where a cell joins two sheets), almost all of them wound oppositely, and
    V - E + F = 2 for a sphere;
  * the particle field against a brute-force numpy minimum over all particles."""
import numpy as np
import pytest

from sdr_amd import scenes
from sdr_amd.surface import TriMesh, sample_nodes, surface_nets


def edge_stats(t: np.ndarray, manifold: bool = True):
    """(closed, misoriented): closed when every undirected edge lies in exactly two
    triangles (manifold) or in an even number of them (a surface-nets cell with two
    sheets joins them through one vertex, so its edges carry four); misoriented = the
    fraction of edges that two triangles run the same way.  The reference orients each
    quad by its first triangle's normal alone (surface_nets.rs:210-218), which can flip
    a quad on a strongly curved patch."""
    d = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]).astype(np.int64)
    u = np.sort(d, axis=1)
    ukey = u[:, 0] * (1 << 32) + u[:, 1]
    _, inv, cu = np.unique(ukey, return_inverse=True, return_counts=True)
    closed = bool((cu == 2).all()) if manifold else bool((cu % 2 == 0).all())
    # on the two-triangle edges: both run it the same way when the forward count is 0 or 2
    fwd = np.bincount(inv, weights=(d[:, 0] < d[:, 1]).astype(np.float64), minlength=len(cu))
    two = cu == 2
    return closed, float((fwd[two] != 1).sum()) / max(int(two.sum()), 1)


def edges_twice_opposite(t: np.ndarray) -> bool:
    closed, mis = edge_stats(t)
    return closed and mis == 0.0


def test_meshes_a_sphere(orc):  # surface_nets.rs:228-248
    n, dx = 41, 0.1
    origin = (-2.0, -2.0, -2.0)
    mesh, info = surface_nets(origin, dx, (n, n, n), lambda p: float(np.sqrt(p @ p)) - 1.0, lib=orc)
    assert not mesh.is_empty()
    vol, area = mesh.signed_volume(), mesh.surface_area()
    assert vol > 0.0
    assert abs(vol - 4.0 / 3.0 * np.pi) <= 0.15
    assert abs(area - 4.0 * np.pi) <= 1.2
    assert info["volume"] == pytest.approx(vol, rel=1e-12)
    assert info["area"] == pytest.approx(area, rel=1e-12)
    closed, mis = edge_stats(mesh.indices)
    assert closed and mis < 0.01
    nv, nf = len(mesh.vertices), mesh.triangle_count()
    assert nv - 3 * nf // 2 + nf == 2  # Euler characteristic of a sphere
    # every vertex lies within its cell, and near the sphere
    r = np.linalg.norm(mesh.vertices, axis=1)
    assert np.abs(r - 1.0).max() < dx


def test_single_negative_node_closed_form(orc):
    vals = np.ones((5, 5, 5))
    vals[2, 2, 2] = -1.0
    mesh, info = surface_nets((0.0, 0.0, 0.0), 1.0, (5, 5, 5), vals, lib=orc)
    assert len(mesh.vertices) == 8 and mesh.triangle_count() == 12
    # each of the 8 cells around node (2,2,2) crosses its 3 edges at that node at t = 1/2:
    # the vertex is the node offset by 1/6 towards the cell's centre
    want = sorted(tuple(2.0 + s / 6.0 for s in (sx, sy, sz))
                  for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1))
    got = sorted(tuple(v) for v in mesh.vertices)
    assert np.allclose(got, want, rtol=0, atol=1e-15)
    assert edges_twice_opposite(mesh.indices)
    assert mesh.signed_volume() > 0.0  # outward: the field rises away from the node


def test_plane_closed_form(orc):
    nx, ny, nz, c = 7, 6, 5, 1.37
    vals = np.array([[[k - c for _ in range(nx)] for _ in range(ny)] for k in range(nz)], float)
    mesh, info = surface_nets((0.0, 0.0, 0.0), 1.0, (nx, ny, nz), vals, lib=orc)
    assert len(mesh.vertices) == (nx - 1) * (ny - 1)
    assert mesh.triangle_count() == 2 * (nx - 2) * (ny - 2)
    assert np.allclose(mesh.vertices[:, 2], c, rtol=0, atol=1e-15)
    assert np.allclose(mesh.vertices[:, 0] % 1.0, 0.5) and np.allclose(mesh.vertices[:, 1] % 1.0, 0.5)
    assert info["area"] == pytest.approx((nx - 2) * (ny - 2), rel=1e-14)
    a, b, cc = (mesh.vertices[mesh.indices[:, q]] for q in range(3))
    assert (np.cross(b - a, cc - a)[:, 2] > 0).all()  # the field rises along +z


def test_open_surface_at_the_max_faces(orc):
    """A field straddling on the grid's max faces: no quad uses a ring cell outside the
    grid (the reference would index past its row there)."""
    vals = np.fromfunction(lambda k, j, i: (i + j + k) - 5.5, (4, 4, 4))
    mesh, _ = surface_nets((0.0, 0.0, 0.0), 1.0, (4, 4, 4), vals, lib=orc)
    assert mesh.triangle_count() > 0
    assert int(mesh.indices.max()) < len(mesh.vertices)
    assert (mesh.vertices >= 0).all() and (mesh.vertices <= 3).all()


def test_degenerate_inputs(orc):
    from sdr_amd import _abi as A

    with pytest.raises(ValueError):
        surface_nets((0, 0, 0), 1.0, (1, 4, 4), np.ones(16), lib=orc)
    mesh, info = surface_nets((0, 0, 0), 1.0, (3, 3, 3), np.ones(27), lib=orc)
    assert mesh.is_empty() and info["vertices"] == 0 and info["area"] == 0.0
    assert A.SDR_E_INVALID == -1


def _box_solver(orc, n_particles, seed):
    sc = scenes.box_scene()
    s = scenes.make_solver(sc, lib=orc, fill=False)
    rng = np.random.default_rng(seed)
    lo = np.asarray(s.origin) + 2.0 * s.dx
    hi = lo + np.array([s.nx - 4, s.ny - 4, s.nz - 4]) * s.dx
    pos = rng.uniform(lo, hi, (n_particles, 3))
    s.set_particles(pos, np.zeros_like(pos))
    return s, pos


@pytest.mark.parametrize("radius_cells,spacing_cells", [(0.0, 0.0), (0.35, 1.0), (0.6, 0.5), (0.5, 1.5)])
def test_particle_field_matches_brute_force(orc, radius_cells, spacing_cells):
    s, pos = _box_solver(orc, 400, seed=11)
    info = s.surface(radius=radius_cells * s.dx, spacing=spacing_cells * s.dx)
    r = info["radius"]
    assert r == (radius_cells * s.dx if radius_cells else s.dx / np.cbrt(s.fluid.particles_per_cell))
    phi = s.surface_field()
    h = info["spacing"]
    o = np.asarray(s.origin)
    k, j, i = np.meshgrid(np.arange(info["nz"]), np.arange(info["ny"]), np.arange(info["nx"]), indexing="ij")
    nodes = np.stack([o[0] + i * h, o[1] + j * h, o[2] + k * h], axis=-1).reshape(-1, 3)
    best = np.full(len(nodes), np.inf)
    for p in pos:
        e = nodes - p
        best = np.minimum(best, e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1] + e[:, 2] * e[:, 2])
    want = np.minimum(np.sqrt(best), 2.0 * r) - r
    assert np.array_equal(phi.reshape(-1), want)
    # the surface is the nets of that field
    mesh, ninfo = surface_nets(info["origin"], h, (info["nx"], info["ny"], info["nz"]), want, lib=orc)
    v, t = s.surface_mesh()
    assert np.array_equal(v, mesh.vertices) and np.array_equal(t, mesh.indices)
    assert (ninfo["vertices"], ninfo["triangles"]) == (info["vertices"], info["triangles"])


def test_particle_surface_is_closed_and_bounds_the_balls(orc):
    s, pos = _box_solver(orc, 1500, seed=5)
    info = s.surface()
    v, t = s.surface_mesh()
    closed, mis = edge_stats(t, manifold=False)
    assert info["triangles"] > 0 and closed and mis < 0.05
    assert info["volume"] > 0.0
    # the union of balls holds at most n (4/3) pi r^3 of volume
    assert info["volume"] <= len(pos) * 4.0 / 3.0 * np.pi * info["radius"] ** 3 * 1.05
    # vertices lie within one node spacing of some particle's sphere
    d = np.min(np.linalg.norm(v[:, None, :] - pos[None, :, :], axis=2), axis=1)
    assert np.abs(d - info["radius"]).max() < info["spacing"]


def test_sample_nodes_order():
    vals = sample_nodes((1.0, 2.0, 3.0), 0.5, (3, 2, 2), lambda p: p[0] + 10 * p[1] + 100 * p[2])
    assert vals.shape == (2, 2, 3)
    assert vals[1, 1, 2] == pytest.approx(2.0 + 10 * 2.5 + 100 * 3.5)
    assert TriMesh(np.zeros((0, 3)), np.zeros((0, 3), np.uint32)).is_empty()
