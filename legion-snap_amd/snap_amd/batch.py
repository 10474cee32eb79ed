"""Batches of independent problems solved together on the HIP device (the
snap_hip_batch_* entry points): one launch per sweep stage for all
problems, per-problem iteration control on the device, a hipGraph-captured
outer iteration, and a tile shape autotuned on the batch."""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import numpy as np

from . import _lib
from .problems import Problem
from .solver import _PREC, MIXED_FP64_POLISH, _ptr, problem_struct


class Batch:
    def __init__(self, problems: List[Problem], precision: str = "fp64", tile=(0, 0, 0),
                 waves_per_tile: int = 0, angles_per_lane: int = 0, use_graph: bool = True,
                 fp64_polish: Optional[float] = None, sweep_kernel: str = "auto",
                 groups_per_wave: int = 0):
        """fp64_polish (mixed): a problem's next inner sweep runs in FP64 once
        its previous inner of the same outer changed its flux by at most
        fp64_polish * epsi; None = the solver's default, 0 = FP32 throughout.
        sweep_kernel: "auto" (autotuned), "tiled" or "wave_tile" (one wavefront
        per 4 x 4 x TZ tile; tile[2] sets TZ).
        groups_per_wave: 0 = two groups per wavefront on one-wavefront tiles
        where that form applies (every problem's x and y extents multiples of
        4), 1 = one."""
        self.problems = [q.contiguous() for q in problems]
        for q in self.problems:
            q.validate()
        self.lib = _lib.load()
        arr = (_lib.Problem * len(self.problems))()
        for i, q in enumerate(self.problems):
            arr[i] = problem_struct(q)
        o = _lib.BatchOptions()
        o.precision = _PREC[precision]
        o.tile[0], o.tile[1], o.tile[2] = tile
        o.waves_per_tile = waves_per_tile
        o.angles_per_lane = angles_per_lane
        o.use_graph = int(bool(use_graph))
        if fp64_polish is None:
            fp64_polish = MIXED_FP64_POLISH if precision == "mixed" else 0.0
        o.fp64_polish = float(fp64_polish)
        o.sweep_kernel = {"auto": 0, "tiled": 1, "wave_tile": 2}[sweep_kernel]
        o.groups_per_wave = groups_per_wave
        h = ctypes.c_void_p()
        _lib.check(self.lib.snap_hip_batch_create(arr, len(self.problems), ctypes.byref(o),
                                                  ctypes.byref(h)), "snap_hip_batch_create")
        self._h = h

    def run(self) -> Tuple[dict, List[dict]]:
        st = _lib.BatchStats()
        per = (_lib.ProblemStats * len(self.problems))()
        _lib.check(self.lib.snap_hip_batch_run(self._h, ctypes.byref(st), per), "snap_hip_batch_run")
        stats = {k: getattr(st, k) for k, _ in st._fields_}
        stats["tile"] = tuple(st.tile)
        return stats, [{k: getattr(x, k) for k, _ in x._fields_} for x in per]

    def flux(self, i: int) -> Tuple[np.ndarray, Optional[np.ndarray]]:
        p = self.problems[i]
        f0 = np.zeros((p.ng, p.nz, p.ny, p.nx))
        fm = np.zeros((p.ng, p.nz, p.ny, p.nx, p.cmom - 1)) if p.cmom > 1 else None
        _lib.check(self.lib.snap_hip_batch_get_flux(self._h, i, _ptr(f0), _ptr(fm)),
                   "snap_hip_batch_get_flux")
        return f0, fm

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.snap_hip_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()
