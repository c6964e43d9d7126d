"""Build libsdr_amd.so (gfx950 HIP kernels + host solver + C ABI) in-tree.

Usage: ``python -m sdr_amd.build`` (or ``sdr_amd.build.build()``).  Objects go to
``sdr_amd/_build/``, the shared library to ``sdr_amd/libsdr_amd.so``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libsdr_amd.so"
BUILD = PKG / "_build"
ARCH = os.environ.get("SDR_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the f64 path reproduces the reference operand order bit for bit.
# packed-fp32-ops off: packed FP32 VALU forms are kept out of generated code.
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
SOURCES = ["kernels.hip", "mic_pipe.hip", "redist.hip", "p2g.hip", "sort.hip", "metrics.hip", "mesh_sdf.hip",
           "surface.hip", "output.cpp", "pool.cpp",
           "solver.cpp",
           "serve.cpp"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or Path(cand).exists()):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> Path:
    obj = BUILD / (src.replace(".", "_") + ".o")
    cmd = [_hipcc(), *COMMON, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}", file=sys.stderr)
    return OUT


if __name__ == "__main__":
    build()
