"""The no-alloc projection (sdr_project_device_noalloc) against the allocating one
(sdr_project_capturing_with_source) through the C ABI: the same kernels on the same
inputs, so pressures, faces and the report agree bit for bit, on two calls in a row
(the second on a smaller grid, over working memory the first call left behind), in
both dot modes; a grid past the fixed working memory is an error."""
import ctypes as C

import numpy as np
import pytest

from sdr_amd import _abi as A
from sdr_amd.solver import SolveParams, project_capturing_with_source

pytestmark = pytest.mark.gpu

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _HIP.hipFree.argtypes = [C.c_void_p]
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _HIP


class DevArray:
    """A device copy of a numpy array (the caller's memory)."""

    def __init__(self, a: np.ndarray):
        self.a = np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(self.a.nbytes, 256)) == 0
        assert hip().hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0  # host to device

    def get(self) -> np.ndarray:
        out = np.empty_like(self.a)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.p.value:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def _case(shape, seed):
    nx, ny, nz = shape
    g = np.random.default_rng(seed)
    r = g.random((nz, ny, nx))
    cells = np.where(r < 0.8, 1, np.where(r < 0.9, 0, 2)).astype(np.uint8)
    faces = (g.standard_normal((nz, ny, nx + 1)), g.standard_normal((nz, ny + 1, nx)),
             g.standard_normal((nz + 1, ny, nx)))
    src = np.where(cells == 1, g.random((nz, ny, nx)) * 0.05, 0.0)
    return cells, faces, src


def _noalloc(lib, shape, cells, faces, src, params, dot, solve_mode):
    nx, ny, nz = shape
    bufs = [DevArray(f) for f in faces] + [DevArray(cells), DevArray(src), DevArray(np.zeros((nz, ny, nx)))]
    d = A.Dims(nx, ny, nz, 0.01, A.vec3((0.0, 0.0, 0.0)))
    rep = A.SolveReportC()
    opts = A.Options(A.PRECISION_F64, dot, 0, 1e-9, solve_mode, 0)
    try:
        lib.project_device_noalloc(C.byref(d), bufs[0].p, bufs[1].p, bufs[2].p, bufs[3].p, 1000.0, 1e-3,
                                   C.byref(A.SolveParamsC(params.max_iters, params.tol)), bufs[4].p,
                                   bufs[5].p, C.byref(rep), C.byref(opts), None)
        return rep, bufs[5].get(), [b.get() for b in bufs[:3]]
    finally:
        for b in bufs:
            b.free()


@pytest.mark.parametrize("dot,solve_mode", [(A.DOT_EXACT, A.SOLVE_REFERENCE), (A.DOT_FAST, A.SOLVE_RELATIVE)])
def test_noalloc_projection_matches_the_allocating_entry_point(dev, dot, solve_mode):
    params = SolveParams(80, 1e-9)
    for shape, seed in (((70, 44, 52), 1), ((24, 20, 28), 2)):  # the second call is smaller
        nx, ny, nz = shape
        cells, faces, src = _case(shape, seed)
        rep, p, got = _noalloc(dev, shape, cells, faces, src, params, dot, solve_mode)
        u, v, w = (f.copy() for f in faces)
        rref, pref = project_capturing_with_source(nx, ny, nz, 0.01, (0, 0, 0), u, v, w, cells, 1000.0, 1e-3,
                                                   params, src, lib=dev, dot_mode=dot, solve_mode=solve_mode,
                                                   refine_tol=1e-9)
        assert rep.iters == rref.iters > 0 and rep.residual == rref.residual
        assert np.array_equal(p.view(np.uint64), pref.view(np.uint64))
        for a, b in zip(got, (u, v, w)):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_noalloc_rejects_grids_past_its_working_memory(dev):
    shape = (120, 120, 120)
    small = DevArray(np.zeros(64))
    d = A.Dims(*shape, 0.01, A.vec3((0.0, 0.0, 0.0)))
    try:
        with pytest.raises(A.AbiError) as e:
            dev.project_device_noalloc(C.byref(d), small.p, small.p, small.p, small.p, 1000.0, 1e-3, None, None,
                                       None, None, None, None)
        assert e.value.code == A.SDR_E_INVALID and "working memory" in str(e.value)
    finally:
        small.free()
