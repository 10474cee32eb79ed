"""Host-side front end of the HIP solve (Snap::transport_solve, snap.cc:225-568).

``Solver`` owns one problem resident in HBM; ``run()`` performs the
outer/inner loop on the GPU and returns the iteration counts and timings.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import _lib
from .problems import Problem

FP64 = "fp64"              # FP64 sweep, fused multiply-adds
FP64_EXACT = "fp64_exact"  # FP64 sweep in the reference CPU operation order
MIXED = "mixed"            # FP32 sweep, FP64 moments, sources and tests
_PREC = {FP64: _lib.PRECISION_FP64, MIXED: _lib.PRECISION_MIXED,
         FP64_EXACT: _lib.PRECISION_FP64_EXACT}


def _ptr(a: Optional[np.ndarray], typ=ctypes.c_double):
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(typ))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(typ))


def problem_struct(p: Problem) -> "_lib.Problem":
    """The C-ABI view of a contiguous Problem (its arrays must outlive the
    call that reads them)."""
    s = _lib.Problem()
    s.nx, s.ny, s.nz = p.nx, p.ny, p.nz
    s.lx, s.ly, s.lz = p.lx, p.ly, p.lz
    s.ng, s.nang, s.nmom, s.nmat = p.ng, p.nang, p.nmom, p.nmat
    s.mat = _ptr(p.mat, ctypes.c_int32)
    s.sigt = _ptr(p.sigt)
    s.slgg = _ptr(p.slgg)
    s.qi = _ptr(p.qi)
    s.sigt_cell = _ptr(p.sigt_cell)
    s.epsi, s.iitm, s.oitm, s.fixup = p.epsi, p.iitm, p.oitm, int(bool(p.fixup))
    s.timedep, s.tf, s.nsteps, s.mms = int(bool(p.timedep)), p.tf, p.nsteps, int(bool(p.mms))
    return s


# FP64 polish thresholds (in units of epsi).  FP32 sweeps throughout: with
# the flux change contracting by about 0.2 per inner iteration, the last one
# or two inners of a converging outer run in FP64 (5 of 23 sweeps on the
# 64^3 P3 workload), which brings the FP32 sweeps' error from about 4e-6 to
# below 2e-7 against the FP64 solve.  With defect correction (an option) the
# FP32 sweeps carry only the source defect, so a lower threshold suffices on
# that workload: 2 polish sweeps in all, 1.7e-7, 1.54x the FP64 grind; on
# small problems where a fifth of the cell sweeps run the fixup its
# correction sweeps (which skip the fixup) leave up to 1e-5.
MIXED_FP64_POLISH = 25.0
MIXED_REFINE_FP64_POLISH = 10.0


class Solver:
    """One problem on the current HIP device.

    precision: "fp64" (FP64 sweeps with fused multiply-adds), "fp64_exact"
    (FP64 sweeps in the reference CPU operation order, bit-equal to the CPU
    arithmetic) or "mixed" (FP32 sweeps, FP64 moments, sources and tests).
    refine (mixed only): defect correction per outer iteration -- the first
    inner sweep runs in FP64 and is kept as the base; later inners sweep the
    FP64 source defect in FP32 without the fixup and add it to the base (an
    FP64 polish sweep becomes the new base).  Off by default (None); steady
    problems only.
    wide_sweep: one workgroup per cell and sweep, with angles across up to
    1024 lanes -- the form every solve past 128 angles per octant takes.
    sweep_kernel: "auto", "tiled" (multi-wave workgroups per tile) or
    "wave_tile" (one wavefront per 4 x 4 x TZ tile; steady, up to 64 angles).
    fp64_polish (mixed only): an inner runs its sweeps in FP64 when the
    previous inner of its outer changed the scalar flux by at most
    fp64_polish * epsi (or by at most 1e-5, the FP32 sweeps' own rounding
    level, when epsi asks for less), so the outers end on FP64 sweeps; the
    default for mixed precision is MIXED_REFINE_FP64_POLISH with defect
    correction and MIXED_FP64_POLISH without; 0 sweeps every inner after the
    first in FP32.
    groups_per_wave (one-wavefront tiles): 0 = two groups per wavefront
    where that form applies (fused sweeps of up to 32 angles, an even group
    count, x and y extents multiples of 4; bit-equal results), 1 = one.
    """

    def __init__(self, problem: Problem, precision: str = FP64, tile=(0, 0, 0),
                 waves_per_tile: int = 0, first_sweep_fp64: bool = False,
                 dinv: str = "default", angles_per_lane: int = 0,
                 refine: Optional[bool] = None,
                 outer_ordered: bool = False, wide_sweep: bool = False,
                 sweep_kernel: str = "auto", fp64_polish: Optional[float] = None,
                 groups_per_wave: int = 0):
        problem.validate()
        self.problem = p = problem.contiguous()
        self.lib = _lib.load()
        s = problem_struct(p)
        o = _lib.SolveOptions()
        o.precision = _PREC[precision]
        o.tile[0], o.tile[1], o.tile[2] = tile
        o.waves_per_tile = waves_per_tile
        o.first_sweep_fp64 = int(bool(first_sweep_fp64))
        o.dinv_mode = {"default": _lib.DINV_DEFAULT, "stream": _lib.DINV_STREAM,
                       "recompute": _lib.DINV_RECOMPUTE}[dinv]
        o.angles_per_lane = angles_per_lane
        if refine is None:
            refine = False
        o.refine = int(bool(refine))
        o.outer_ordered = int(bool(outer_ordered))
        o.wide_sweep = int(bool(wide_sweep))
        o.sweep_kernel = {"auto": 0, "tiled": 1, "wave_tile": 2}[sweep_kernel]
        if fp64_polish is None:
            fp64_polish = (0.0 if precision != MIXED else
                           MIXED_REFINE_FP64_POLISH if refine else MIXED_FP64_POLISH)
        o.fp64_polish = float(fp64_polish)
        o.groups_per_wave = groups_per_wave
        self.precision = precision
        h = ctypes.c_void_p()
        _lib.check(self.lib.snap_hip_solver_create(ctypes.byref(s), ctypes.byref(o), ctypes.byref(h)),
                   "snap_hip_solver_create")
        self._h = h

    def run(self) -> dict:
        st = _lib.SolveStats()
        _lib.check(self.lib.snap_hip_solver_run(self._h, ctypes.byref(st)), "snap_hip_solver_run")
        return st.as_dict()

    def flux(self) -> Tuple[np.ndarray, Optional[np.ndarray]]:
        p = self.problem
        f0 = np.zeros((p.ng, p.nz, p.ny, p.nx))
        fm = np.zeros((p.ng, p.nz, p.ny, p.nx, p.cmom - 1)) if p.cmom > 1 else None
        _lib.check(self.lib.snap_hip_solver_get_flux(self._h, _ptr(f0), _ptr(fm)),
                   "snap_hip_solver_get_flux")
        return f0, fm

    def time_sweeps(self, count: int, precision: Optional[str] = None) -> dict:
        ms = ctypes.c_double()
        ml = ctypes.c_double()
        nl = ctypes.c_int()
        prec = _PREC[precision or self.precision]
        _lib.check(self.lib.snap_hip_solver_time_sweeps(self._h, count, prec, ctypes.byref(ms),
                                                         ctypes.byref(ml), ctypes.byref(nl)),
                   "snap_hip_solver_time_sweeps")
        return {"ms_per_sweep": ms.value, "ms_per_launch": ml.value, "launches_per_sweep": nl.value}

    def gradient(self, resp: np.ndarray, precision: Optional[str] = None,
                 tol: Optional[float] = None, maxit: int = 2000):
        """dR/dt_xs[g, z, y, x] of R = sum(resp * flux0) at the current flux
        (run() first; nmom = 1).  precision "fp64" or "mixed" (FP32 reverse
        sweeps; the forward recompute and its fixup decisions stay FP64);
        tol defaults to 1e-12 (FP64) or 1e-9 (FP32, whose iteration also ends
        where its rounding stops the change from shrinking, near 1e-9 on the
        48^3 gradient workload).  Returns (grad, R, adjoint iterations)."""
        p = self.problem
        r = np.ascontiguousarray(resp, dtype=np.float64).reshape(p.ng, p.nz, p.ny, p.nx)
        grad = np.zeros_like(r)
        R = ctypes.c_double()
        it = ctypes.c_int()
        prec = _PREC[precision or ("mixed" if self.precision == "mixed" else "fp64")]
        if tol is None:
            tol = 1e-9 if prec == _lib.PRECISION_MIXED else 1e-12
        _lib.check(self.lib.snap_hip_solver_gradient(self._h, _ptr(r), prec, tol, maxit, _ptr(grad),
                                                     ctypes.byref(R), ctypes.byref(it)),
                   "snap_hip_solver_gradient")
        return grad, R.value, it.value

    def mms_compare(self):
        """MMSCompare of the current flux (mms.cc:497-541): (max, min, mean)
        of |phi/ref - 1| over all cells and groups (mms problems only)."""
        out = np.zeros(3)
        _lib.check(self.lib.snap_hip_solver_mms_compare(self._h, _ptr(out)),
                   "snap_hip_solver_mms_compare")
        return float(out[0]), float(out[1]), float(out[2])

    def iteration_log(self) -> list:
        """The convergence tests of the last run as dicts (kind "inner" or
        "outer", time_step, outer, inner, converged, t_us)."""
        n = ctypes.c_int()
        _lib.check(self.lib.snap_hip_solver_get_log(self._h, None, 0, ctypes.byref(n)),
                   "snap_hip_solver_get_log")
        recs = (_lib.IterationRecord * max(1, n.value))()
        _lib.check(self.lib.snap_hip_solver_get_log(self._h, recs, n.value, ctypes.byref(n)),
                   "snap_hip_solver_get_log")
        return [{"kind": "inner" if r.kind == 0 else "outer", "time_step": r.time_step,
                 "outer": r.outer, "inner": r.inner, "converged": bool(r.converged),
                 "t_us": r.t_us} for r in recs[:n.value]]

    def mms_reference(self) -> np.ndarray:
        """The manufactured scalar flux [ng, nz, ny, nx] (mms problems only)."""
        p = self.problem
        ref = np.zeros((p.ng, p.nz, p.ny, p.nx))
        _lib.check(self.lib.snap_hip_solver_get_mms_reference(self._h, _ptr(ref)),
                   "snap_hip_solver_get_mms_reference")
        return ref

    def plan(self) -> dict:
        """The sweep plan the solver chose (tile edges, waves per tile,
        angles per lane, cells per wavefront)."""
        v = (ctypes.c_int32 * 6)()
        _lib.check(self.lib.snap_hip_solver_plan(self._h, v), "snap_hip_solver_plan")
        return {"tile": (v[0], v[1], v[2]), "waves": v[3], "angles_per_lane": v[4],
                "cells_per_wave": v[5]}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.snap_hip_solver_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(problem: Problem, precision: str = FP64, **kw):
    """Solve on the GPU; returns (flux0 [ng,nz,ny,nx], fluxm or None, stats)."""
    with Solver(problem, precision=precision, **kw) as s:
        stats = s.run()
        f0, fm = s.flux()
    return f0, fm, stats
