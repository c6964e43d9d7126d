"""Triangle meshes and naive surface nets over the C ABI.

`TriMesh` mirrors the parts of the reference's `TriMesh` (src/mesh.rs:16-456) that a
surface needs: `surface_area`, `signed_volume`, `save_obj`, `save_stl`.
`surface_nets(origin, dx, dims, field)` is the reference's polygoniser
(src/surface_nets.rs:48-195) run on the device; `field` is a callable of a point (as in
the reference) or the sampled node values.  `Solver.surface` surfaces a solver's
resident particles the same way.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _abi as A


def _lib(lib):
    if lib is not None:
        return lib
    from . import load_library

    return load_library()


@dataclass
class TriMesh:
    vertices: np.ndarray  # (nv, 3) f64
    indices: np.ndarray   # (nt, 3) u32

    def triangle_count(self) -> int:
        return int(self.indices.shape[0])

    def is_empty(self) -> bool:
        return self.triangle_count() == 0

    def surface_area(self) -> float:
        """mesh.rs:52-60: the sum of triangle areas, in triangle order."""
        a, b, c = (self.vertices[self.indices[:, q]] for q in range(3))
        areas = 0.5 * np.sqrt(np.sum(np.cross(b - a, c - a) ** 2, axis=1))
        return float(sum(areas.tolist()))

    def signed_volume(self) -> float:
        """mesh.rs:63-70: signed tetrahedra to the origin, positive for outward winding."""
        a, b, c = (self.vertices[self.indices[:, q]] for q in range(3))
        vols = np.sum(a * np.cross(b, c), axis=1) / 6.0
        return float(sum(vols.tolist()))

    def save_obj(self, path, *, lib: A.Lib | None = None) -> None:
        """`TriMesh::save_obj` (mesh.rs:356-379)."""
        v, t = self._c()
        _lib(lib).write_mesh_obj(os.fsencode(path), A.dptr(v), v.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 t.shape[0])

    def save_stl(self, path, *, lib: A.Lib | None = None) -> None:
        """`TriMesh::save_stl` (mesh.rs:427-456)."""
        v, t = self._c()
        _lib(lib).write_mesh_stl(os.fsencode(path), A.dptr(v), v.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 t.shape[0])

    def _c(self):
        v = np.ascontiguousarray(self.vertices, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(self.indices, np.uint32).reshape(-1, 3)
        return v, t


def sample_nodes(origin, dx: float, dims, field) -> np.ndarray:
    """The node values of surface_nets.rs:61-72: field(origin + (i, j, k) dx), x fastest."""
    nx, ny, nz = dims
    o = np.asarray(origin, np.float64)
    out = np.empty((nz, ny, nx))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                out[k, j, i] = field(o + np.array([i * dx, j * dx, k * dx]))
    return out


def surface_nets(origin, dx: float, dims, field, *, device: int = 0, lib: A.Lib | None = None):
    """`surface_nets(origin, dx, dims, field)` (src/surface_nets.rs:48-195).

    Returns (TriMesh, info) where info carries the counts, area and signed volume the
    device summed.  Fewer than 2 nodes on an axis raises (the reference asserts)."""
    nx, ny, nz = (int(d) for d in dims)
    if nx < 2 or ny < 2 or nz < 2:
        raise ValueError("surface_nets needs at least 2 nodes per axis")
    vals = field if isinstance(field, np.ndarray) else sample_nodes(origin, dx, dims, field)
    vals = np.ascontiguousarray(vals, np.float64).reshape(-1)
    if vals.size != nx * ny * nz:
        raise ValueError("field must hold nx*ny*nz node values")
    lb = _lib(lib)
    info = A.SurfaceInfo()
    o = A.Vec3(*map(float, origin))
    lb.surface_nets(A.dptr(vals), nx, ny, nz, o, dx, device, C.byref(info), None, 0, None, 0)
    v = np.zeros((int(info.vertices), 3))
    t = np.zeros((int(info.triangles), 3), np.uint32)
    lb.surface_nets(A.dptr(vals), nx, ny, nz, o, dx, device, C.byref(info), A.dptr(v), v.shape[0],
                    t.ctypes.data_as(C.POINTER(C.c_uint32)), t.shape[0])
    return TriMesh(v, t), A.struct_dict(info)
