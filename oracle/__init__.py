"""The CPU oracle (test infrastructure only).

A faithful f64 restatement of the reference FLIP/PIC substep (oracle.cpp), bound with
the same ctypes declarations as the device library.  Only tests/, the smoke check and
the CPU-baseline leg of bench.py use it; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from sdr_amd import _abi

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_LIB = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def load() -> _abi.Lib:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        _LIB = _abi.Lib(LIB_PATH, "orc_", device=False)
        d = _LIB.dll
        d.orc_set_threads.argtypes = [C.c_int32]
        d.orc_get_threads.restype = C.c_int
        d.orc_trilinear_sample.argtypes = [_abi.c_double_p, C.c_uint64, C.c_uint64, C.c_uint64,
                                           C.c_double, C.c_double, C.c_double, _abi.c_double_p]
        d.orc_trilinear_scatter.argtypes = [_abi.c_double_p, _abi.c_double_p, C.c_uint64,
                                            C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                            C.c_double, C.c_double]
        d.orc_sdf_sample.argtypes = [C.POINTER(_abi.SdfDesc), _abi.c_double_p, _abi.c_double_p,
                                     _abi.c_double_p]
        d.orc_hash_jitter.argtypes = [C.c_uint64, _abi.c_double_p]
        d.orc_rng_units.argtypes = [C.c_uint64, C.c_uint64, _abi.c_double_p, _abi.c_u64_p]
        d.orc_laplacian.argtypes = [C.c_uint64] * 3 + [_abi.c_u8_p] + [_abi.c_double_p] * 4
        d.orc_mic0.argtypes = [C.c_uint64] * 3 + [_abi.c_u8_p, _abi.c_double_p, C.c_double,
                                                  _abi.c_double_p, _abi.c_double_p]
        d.orc_apply_a.argtypes = [C.c_uint64] * 3 + [_abi.c_u8_p, _abi.c_double_p, _abi.c_double_p]
        d.orc_dot.argtypes = [_abi.c_double_p, _abi.c_double_p, C.c_uint64, _abi.c_double_p]
        d.orc_extrapolate.argtypes = [_abi.c_double_p, _abi.c_u8_p, C.c_uint64, C.c_uint64,
                                      C.c_uint64, C.c_int32]
        d.orc_solver_get_mask.argtypes = [C.c_void_p, _abi.c_u8_p]
    return _LIB


def set_threads(n: int) -> None:
    load().dll.orc_set_threads(int(n))


def get_threads() -> int:
    return int(load().dll.orc_get_threads())


# ---- pieces for pinning the oracle against closed forms and brute force
def trilinear_sample(arr: np.ndarray, gx, gy, gz) -> float:
    nz, ny, nx = arr.shape
    a = np.ascontiguousarray(arr, np.float64)
    out = C.c_double()
    load().dll.orc_trilinear_sample(_abi.dptr(a), nx, ny, nz, gx, gy, gz, C.byref(out))
    return out.value


def trilinear_scatter(arr: np.ndarray, wsum: np.ndarray, gx, gy, gz, val) -> None:
    nz, ny, nx = arr.shape
    load().dll.orc_trilinear_scatter(_abi.dptr(arr), _abi.dptr(wsum), nx, ny, nz, gx, gy, gz, val)


def sdf_sample(sdf, p):
    sd = sdf.c()
    pp = np.asarray(p, np.float64)
    phi = C.c_double()
    g = np.zeros(3)
    load().dll.orc_sdf_sample(C.byref(sd), _abi.dptr(pp), C.byref(phi), _abi.dptr(g))
    return phi.value, g


def hash_jitter(key: int) -> np.ndarray:
    out = np.zeros(3)
    load().dll.orc_hash_jitter(key, _abi.dptr(out))
    return out


def rng_units(seed: int, n: int):
    out = np.zeros(n)
    st = C.c_uint64()
    load().dll.orc_rng_units(seed, n, _abi.dptr(out), C.byref(st))
    return out, st.value


def laplacian(cells: np.ndarray):
    nz, ny, nx = cells.shape
    c = np.ascontiguousarray(cells, np.uint8)
    d, px, py, pz = (np.zeros(cells.shape) for _ in range(4))
    load().dll.orc_laplacian(nx, ny, nz, _abi.u8ptr(c), _abi.dptr(d), _abi.dptr(px),
                             _abi.dptr(py), _abi.dptr(pz))
    return d, px, py, pz


def mic0(cells: np.ndarray, r: np.ndarray | None = None, tau: float = -1.0):
    nz, ny, nx = cells.shape
    c = np.ascontiguousarray(cells, np.uint8)
    pc = np.zeros(cells.shape)
    z = np.zeros(cells.shape) if r is not None else None
    rr = None if r is None else np.ascontiguousarray(r, np.float64)
    load().dll.orc_mic0(nx, ny, nz, _abi.u8ptr(c), _abi.dptr(rr), tau, _abi.dptr(pc),
                        _abi.dptr(z))
    return pc, z


def apply_a(cells: np.ndarray, x: np.ndarray) -> np.ndarray:
    nz, ny, nx = cells.shape
    c = np.ascontiguousarray(cells, np.uint8)
    xx = np.ascontiguousarray(x, np.float64)
    out = np.zeros(cells.shape)
    load().dll.orc_apply_a(nx, ny, nz, _abi.u8ptr(c), _abi.dptr(xx), _abi.dptr(out))
    return out


def dot(a: np.ndarray, b: np.ndarray) -> float:
    aa = np.ascontiguousarray(a, np.float64).reshape(-1)
    bb = np.ascontiguousarray(b, np.float64).reshape(-1)
    out = C.c_double()
    load().dll.orc_dot(_abi.dptr(aa), _abi.dptr(bb), aa.size, C.byref(out))
    return out.value


def extrapolate(field: np.ndarray, valid: np.ndarray, iters: int = 4) -> np.ndarray:
    nz, ny, nx = field.shape
    f = np.ascontiguousarray(field, np.float64).copy()
    v = np.ascontiguousarray(valid, np.uint8)
    load().dll.orc_extrapolate(_abi.dptr(f), _abi.u8ptr(v), nx, ny, nz, iters)
    return f


def cavity_mask(solver) -> np.ndarray:
    m = np.zeros((solver.nz, solver.ny, solver.nx), np.uint8)
    load().dll.orc_solver_get_mask(solver._h, _abi.u8ptr(m))
    return m
