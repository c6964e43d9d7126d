"""sdr_amd — MI355X-native FLIP/PIC substep of the `sdr` sinus-irrigation simulator.

The product is the gfx950 HIP library ``libsdr_amd.so`` built in-tree
(``python -m sdr_amd.build``) and its C ABI (include/sdr_amd.h).  This package is a
thin ctypes mirror of the reference `Solver` / `pressure` API on top of it.  There is
no CPU fallback: every entry point raises if the library is missing.
"""
from __future__ import annotations

from pathlib import Path

from . import _abi
from ._abi import (DOT_EXACT, DOT_FAST, PRECISION_F32, PRECISION_F64, STAGES, AbiError)
from .solver import (AIR, FLUID, SOLID, Aabb, FluidParams, Needle, Sdf, SolveParams, SolveReport,
                     Solver, project_capturing_with_source)

LIB_PATH = Path(__file__).resolve().parent / "libsdr_amd.so"
_LIB = None


def load_library() -> _abi.Lib:
    """The device library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m sdr_amd.build`")
        _LIB = _abi.Lib(LIB_PATH, "sdr_", device=True)
    return _LIB


def device_info():
    import ctypes as C

    lib = load_library()
    n = C.c_int32()
    arch = C.create_string_buffer(64)
    lib.device_info(C.byref(n), arch, 64)
    return int(n.value), arch.value.decode()


__all__ = [
    "AIR", "FLUID", "SOLID", "Aabb", "FluidParams", "Needle", "Sdf", "SolveParams", "SolveReport",
    "Solver", "project_capturing_with_source", "load_library", "device_info", "AbiError",
    "PRECISION_F32", "PRECISION_F64", "DOT_EXACT", "DOT_FAST", "STAGES",
]
