"""The oracle's mesh -> SDF build (`Sdf::from_mesh`, src/sdf.rs:47-86, 187-272) against
the reference's known answers and closed forms: the box winding test (sdf.rs:298-335),
the point-triangle distances (sdf.rs:338-348), the exact signed distance of a box on a
grid (every closest-point region: faces, edges, vertices, inside and out), a subdivided
icosahedron against the sphere it approximates, and the committed real antrum mesh
against its documented volume when the reference checkout is present."""
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle
from sdr_amd import scenes
from sdr_amd.solver import Aabb, Sdf


def _at(lib, v, f, p):
    s = Sdf.from_mesh(v, f, 1.0, bounds=Aabb(tuple(p), tuple(p)), lib=lib)
    assert (s.nx, s.ny, s.nz) == (1, 1, 1)
    return float(s.data[0])


def test_box_winding_known_answers(orc):
    v, f = scenes.box_mesh()
    assert abs(_at(orc, v, f, (0.0, 0.0, 0.0)) - (-1.0)) < 1e-9   # sdf.rs:328-330
    d = _at(orc, v, f, (3.0, 0.0, 0.0))                              # sdf.rs:331-334
    assert d > 0.0 and abs(d - 2.0) < 1e-9


def test_point_triangle_distance_known_answers(orc):
    """sdf.rs:338-348 through a one-triangle (open) mesh: |phi| = sqrt(d2)."""
    v = np.array([(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)])
    f = np.array([(0, 1, 2)], np.uint32)
    assert abs(abs(_at(orc, v, f, (0.25, 0.25, 2.0))) - 2.0) < 1e-12
    assert abs(abs(_at(orc, v, f, (-1.0, -1.0, 0.0))) - np.sqrt(2.0)) < 1e-12


def _box_sdf(p, h):
    q = np.abs(p) - h
    out = np.sqrt((np.maximum(q, 0.0) ** 2).sum(-1))
    inside = np.all(q < 0.0, axis=-1)
    return np.where(inside, q.max(-1), out)


def test_box_mesh_grid_matches_closed_form(orc):
    h = 0.7
    v, f = scenes.box_mesh(h)
    s = Sdf.from_mesh(v, f, 0.13, padding=0.5, lib=orc)
    k, j, i = np.meshgrid(np.arange(s.nz), np.arange(s.ny), np.arange(s.nx), indexing="ij")
    p = np.stack([s.origin[0] + i * s.dx, s.origin[1] + j * s.dx, s.origin[2] + k * s.dx], -1)
    want = _box_sdf(p, h).reshape(-1)
    got = s.data
    off = np.abs(want) > 1e-9  # nodes on the surface have no defined sign
    assert off.sum() > 0.9 * off.size
    assert np.allclose(got[off], want[off], rtol=0, atol=1e-12)
    assert (got < 0).sum() > 100 and (got > 0).sum() > 100


def test_icosphere_sign_and_distance(orc):
    v, f = scenes.icosphere(1.0, 3)
    s = Sdf.from_mesh(v, f, 0.1, padding=0.35, lib=orc)
    k, j, i = np.meshgrid(np.arange(s.nz), np.arange(s.ny), np.arange(s.nx), indexing="ij")
    p = np.stack([s.origin[0] + i * s.dx, s.origin[1] + j * s.dx, s.origin[2] + k * s.dx], -1).reshape(-1, 3)
    r = np.linalg.norm(p, axis=1)
    got = s.data
    assert np.all(got[r < 0.97] < 0.0) and np.all(got[r > 1.02] > 0.0)
    far = np.abs(r - 1.0) > 0.05
    assert np.allclose(np.abs(got[far]), np.abs(r[far] - 1.0), atol=0.01)


STL = Path("/root/reference/assets/maxillary_sinus.stl")


def read_stl(path):
    """Binary STL (mesh.rs:383-423): 80-byte header, count, 50-byte records of a normal
    and three f32 vertices."""
    b = path.read_bytes()
    n = struct.unpack_from("<I", b, 80)[0]
    rec = np.frombuffer(b, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]), count=n,
                        offset=84)
    v = rec["v"].reshape(-1, 3).astype(np.float64)
    return v, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


@pytest.mark.skipif(not STL.exists(), reason="the reference checkout (assets/) is not present")
def test_real_antrum_volume(orc):
    """The committed antrum mesh (assets/README.md: 13 804 triangles, signed volume
    +2.228 ml) in metres: the inside nodes of its SDF at 1.5 mm hold about that volume."""
    v, f = read_stl(STL)
    assert f.shape[0] == 13804
    v = v * 1e-3
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0
    assert abs(vol * 1e6 - 2.228) < 0.002
    dx = 1.5e-3
    s = Sdf.from_mesh(v, f, dx, padding=3 * dx, lib=orc)
    inside = (s.data < 0.0).sum() * dx ** 3
    assert abs(inside / vol - 1.0) < 0.1, inside / vol
