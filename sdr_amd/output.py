"""Frame output over the C ABI: the reference's `output` module (src/output.rs:23-122)
and an asynchronous frame writer for resident solvers.

`write_points_vtk`, `read_points_vtk`, `write_points_xyz` and `frame_path` keep the
reference's names, formats and errors (a missing file or a malformed POINTS block
raises).  `FrameWriter` has no reference counterpart: it copies a solver's particles
off the device on a stream of its own and writes them from a host thread, so the
substeps queued after `submit` overlap the download and the write.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi as A


def _lib(lib):
    if lib is not None:
        return lib
    from . import load_library

    return load_library()


def _points(positions) -> np.ndarray:
    p = np.ascontiguousarray(positions, np.float64)
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError("positions must be an (n, 3) array")
    return p


def write_points_vtk(path, positions, *, lib: A.Lib | None = None) -> None:
    """`write_points_vtk` (src/output.rs:23-47): ASCII legacy VTK of VTK_VERTEX cells."""
    p = _points(positions)
    _lib(lib).write_points_vtk(os.fsencode(path), A.dptr(p), p.shape[0])


def read_points_vtk(path, *, lib: A.Lib | None = None) -> np.ndarray:
    """`read_points_vtk` (src/output.rs:56-95): the POINTS block, (n, 3) f64."""
    lb = _lib(lib)
    n = C.c_uint64()
    lb.read_points_vtk(os.fsencode(path), None, 0, C.byref(n))
    out = np.zeros((int(n.value), 3))
    lb.read_points_vtk(os.fsencode(path), A.dptr(out), out.shape[0], C.byref(n))
    return out


def write_points_xyz(path, positions, *, lib: A.Lib | None = None) -> None:
    """`write_points_xyz` (src/output.rs:99-112): raw little-endian f32 triples."""
    p = _points(positions)
    _lib(lib).write_points_xyz(os.fsencode(path), A.dptr(p), p.shape[0])


def frame_path(dir, prefix: str, frame: int, ext: str, *, lib: A.Lib | None = None) -> str:
    """`frame_path` (src/output.rs:115-122): ``dir/prefix_NNNNNN.ext``."""
    buf = C.create_string_buffer(4096)
    _lib(lib).frame_path(os.fsencode(dir), prefix.encode(), frame, ext.encode(), buf, len(buf))
    return os.fsdecode(buf.value)


class FrameWriter:
    """Asynchronous frame files ``frame_path(dir, prefix, frame, "vtk" | "xyz")``.

    `slots` staging buffers (1-8) bound the frames in flight; `submit` blocks only while
    all are busy.  A write error is raised by the next `submit` or `flush`."""

    def __init__(self, dir, prefix: str = "particles", fmt: str = "vtk", slots: int = 2, *,
                 lib: A.Lib | None = None):
        self.lib = _lib(lib)
        code = {"vtk": A.FORMAT_VTK, "xyz": A.FORMAT_XYZ}.get(fmt)
        if code is None:
            raise ValueError("fmt must be 'vtk' or 'xyz'")
        self._h = C.c_void_p()
        self.lib.frame_writer_new(os.fsencode(dir), prefix.encode(), code, slots, C.byref(self._h))

    def submit(self, solver, frame: int) -> None:
        self.lib.frame_writer_submit(self._h, solver._h, frame)

    def flush(self) -> int:
        n = C.c_uint64()
        self.lib.frame_writer_flush(self._h, C.byref(n))
        return int(n.value)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            h, self._h = self._h, C.c_void_p()
            self.lib.frame_writer_free(h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
