"""`MetricsCollector` (src/metrics.rs:116-446) over a solver's resident state.

The collector, its wall/washed flags and every per-frame reduction live with the
library the solver was made by (the device library, or the CPU oracle in tests); a
frame is recorded without downloading particles or cells.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A


class MetricsCollector:
    """`MetricsCollector::new(solver, cavity_volume)`; `record(solver, frame)` returns the
    frame's `FrameMetrics` as a dict, `summary()` the run's `Summary`."""

    def __init__(self, solver, cavity_volume: float):
        self.solver = solver
        self.lib = solver.lib
        self._h = C.c_void_p()
        self.lib.metrics_new(solver._h, float(cavity_volume), C.byref(self._h))
        self.history: list[dict] = []

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.metrics_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def wall_count(self) -> int:
        n = C.c_uint64()
        self.lib.metrics_wall_count(self._h, C.byref(n))
        return int(n.value)

    def record(self, solver, frame: int) -> dict:
        m = A.FrameMetrics()
        self.lib.metrics_record(self._h, solver._h, int(frame), C.byref(m))
        d = A.struct_dict(m)
        self.history.append(d)
        return d

    def summary(self) -> dict:
        s = A.Summary()
        self.lib.metrics_summary(self._h, C.byref(s))
        return A.struct_dict(s)

    def cells(self):
        """(washed, is_wall) flags per cell, shaped (nz, ny, nx)."""
        s = self.solver
        washed = np.zeros((s.nz, s.ny, s.nx), np.uint8)
        wall = np.zeros_like(washed)
        self.lib.metrics_get_cells(self._h, A.u8ptr(washed), A.u8ptr(wall))
        return washed, wall

    # metrics.rs:385-431 (write_json / write_csv) over the recorded frames
    CSV_FIELDS = ["frame", "time_s", "particles", "fluid_volume_ml", "drained_volume_ml",
                  "fill_fraction", "wall_coverage", "mean_pressure_pa", "max_pressure_pa",
                  "max_pressure_mmhg", "wall_pressure_pa", "peak_wall_pressure_pa",
                  "peak_wall_pressure_mmhg", "membrane_artifacts", "max_speed_m_s", "kinetic_energy",
                  "pcg_iters"]

    def write_json(self, path) -> None:
        import json

        with open(path, "w") as f:
            json.dump([{k: h[k] for k in self.CSV_FIELDS} for h in self.history], f, indent=2)

    def write_csv(self, path) -> None:
        with open(path, "w") as f:
            f.write(",".join(self.CSV_FIELDS) + "\n")
            for h in self.history:
                f.write(",".join(repr(h[k]) if isinstance(h[k], float) else str(h[k])
                                 for k in self.CSV_FIELDS) + "\n")
