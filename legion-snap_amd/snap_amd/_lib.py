"""ctypes binding of libsnap_hip.so (the C ABI declared in include/snap_hip.h).

The library is built in-tree by ``legion-snap_amd/build.py``.  Loading fails
loudly when it is missing: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsnap_hip.so")

SNAP_OK = 0
SNAP_ERR_INVALID = -1
SNAP_ERR_UNSUPPORTED = -2
SNAP_ERR_HIP = -3
SNAP_ERR_NO_CONTEXT = -4
SNAP_ERR_NO_MEMORY = -5
PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_FP64_EXACT = 2
DINV_DEFAULT = 0
DINV_STREAM = 1
DINV_RECOMPUTE = 2

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class Accessor1(ctypes.Structure):
    _fields_ = [("base", _vp), ("stride", _i64 * 1)]


class Accessor2(ctypes.Structure):
    _fields_ = [("base", _vp), ("stride", _i64 * 2)]


class Accessor3(ctypes.Structure):
    _fields_ = [("base", _vp), ("stride", _i64 * 3)]


class Rect3(ctypes.Structure):
    _fields_ = [("lo", _i64 * 3), ("hi", _i64 * 3)]


class SweepArgs(ctypes.Structure):
    _fields_ = [
        ("origin", _i64 * 3),
        ("qtot", Accessor3), ("flux", Accessor3), ("fluxm", Accessor3), ("dinv", Accessor3),
        ("time_flux_in", Accessor3), ("time_flux_out", Accessor3), ("t_xs", Accessor3),
        ("ghostx", Accessor2), ("ghosty", Accessor2), ("ghostz", Accessor2), ("qim", Accessor3),
        ("x_range", ctypes.c_int), ("y_range", ctypes.c_int), ("z_range", ctypes.c_int),
        ("corner", ctypes.c_int),
        ("stride_x_positive", ctypes.c_int), ("stride_y_positive", ctypes.c_int),
        ("stride_z_positive", ctypes.c_int),
        ("mms_source", ctypes.c_int), ("num_moments", ctypes.c_int),
        ("hi", ctypes.c_double), ("hj", ctypes.c_double), ("hk", ctypes.c_double),
        ("vdelt", ctypes.c_double),
        ("num_angles", ctypes.c_int), ("fixup", ctypes.c_int),
    ]


class Problem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("ng", ctypes.c_int), ("nang", ctypes.c_int), ("nmom", ctypes.c_int),
        ("nmat", ctypes.c_int),
        ("mat", _ip), ("sigt", _dp), ("slgg", _dp), ("qi", _dp), ("sigt_cell", _dp),
        ("epsi", ctypes.c_double), ("iitm", ctypes.c_int), ("oitm", ctypes.c_int),
        ("fixup", ctypes.c_int), ("timedep", ctypes.c_int), ("tf", ctypes.c_double),
        ("nsteps", ctypes.c_int), ("mms", ctypes.c_int),
    ]


class SolveOptions(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int), ("tile", ctypes.c_int * 3),
        ("waves_per_tile", ctypes.c_int), ("first_sweep_fp64", ctypes.c_int),
        ("dinv_mode", ctypes.c_int), ("angles_per_lane", ctypes.c_int), ("refine", ctypes.c_int),
        ("outer_ordered", ctypes.c_int), ("wide_sweep", ctypes.c_int),
        ("sweep_kernel", ctypes.c_int), ("fp64_polish", ctypes.c_double),
        ("groups_per_wave", ctypes.c_int),
    ]


class BatchOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("tile", ctypes.c_int * 3),
                ("waves_per_tile", ctypes.c_int), ("angles_per_lane", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("fp64_polish", ctypes.c_double),
                ("sweep_kernel", ctypes.c_int), ("groups_per_wave", ctypes.c_int)]


class ProblemStats(ctypes.Structure):
    _fields_ = [("inner_total", ctypes.c_int), ("outer_total", ctypes.c_int),
                ("converged", ctypes.c_int), ("sweeps_fp64", ctypes.c_int),
                ("time_ms", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [("outer_iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("solve_ms", ctypes.c_double), ("units", ctypes.c_double),
                ("grind_ns", ctypes.c_double), ("tile", ctypes.c_int * 3), ("waves", ctypes.c_int),
                ("angles_per_lane", ctypes.c_int), ("autotune_ms", ctypes.c_double),
                ("plan_cached", ctypes.c_int), ("groups_per_wave", ctypes.c_int)]


class IterationRecord(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("time_step", ctypes.c_int), ("outer", ctypes.c_int),
                ("inner", ctypes.c_int), ("converged", ctypes.c_int), ("t_us", ctypes.c_double)]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("inner_total", ctypes.c_int), ("outer_total", ctypes.c_int),
        ("converged", ctypes.c_int), ("sweeps_fp64", ctypes.c_int),
        ("sweeps_fp32", ctypes.c_int),
        ("sweep_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
        ("units_per_sweep", ctypes.c_double), ("grind_ns", ctypes.c_double),
        ("time_steps", ctypes.c_int), ("groups_per_wave", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# every extern "C" symbol of include/snap_hip.h, with its ctypes signature
SIGNATURES = {
    "snap_hip_last_error": ([], ctypes.c_char_p),
    "snap_hip_version": ([], ctypes.c_char_p),
    "snap_hip_quadrature": ([ctypes.c_int, ctypes.c_int] + [_dp] * 5, ctypes.c_int),
    "snap_hip_initialize_gpu_context": (
        [_dp, _dp, _dp, _dp, _dp] + [ctypes.c_int] * 6, ctypes.c_int),
    "snap_hip_run_gpu_sweep": ([ctypes.POINTER(SweepArgs), _vp], ctypes.c_int),
    "snap_hip_run_gpu_sweep_noalloc": ([ctypes.POINTER(SweepArgs), _vp], ctypes.c_int),
    "snap_hip_run_gpu_sweep_groups": ([ctypes.POINTER(SweepArgs), ctypes.c_int, ctypes.c_int, _vp],
                                      ctypes.c_int),
    "snap_hip_run_flux0_outer_source": (
        [Rect3, ctypes.POINTER(Accessor3), ctypes.POINTER(Accessor3), ctypes.POINTER(Accessor2),
         ctypes.POINTER(Accessor3), Accessor3, ctypes.c_int, _vp], ctypes.c_int),
    "snap_hip_run_fluxm_outer_source": (
        [Rect3, ctypes.POINTER(Accessor3), ctypes.POINTER(Accessor2), ctypes.POINTER(Accessor3),
         Accessor3, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp], ctypes.c_int),
    "snap_hip_run_inner_source_single_moment": (
        [Rect3, Accessor3, Accessor3, Accessor3, Accessor3, _vp], ctypes.c_int),
    "snap_hip_run_inner_source_multi_moment": (
        [Rect3] + [Accessor3] * 6 + [ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp],
        ctypes.c_int),
    "snap_hip_run_inner_convergence": (
        [Rect3, _vp, ctypes.POINTER(Accessor3), ctypes.POINTER(Accessor3), ctypes.c_int,
         ctypes.c_double, _vp], ctypes.c_int),
    "snap_hip_run_outer_convergence": (
        [Rect3, _vp, ctypes.POINTER(Accessor3), ctypes.POINTER(Accessor3), ctypes.c_int,
         ctypes.c_double, _vp], ctypes.c_int),
    "snap_hip_run_expand_cross_section": (
        [ctypes.POINTER(Accessor1), Accessor3, ctypes.POINTER(Accessor3), ctypes.c_int, Rect3,
         _vp], ctypes.c_int),
    "snap_hip_run_expand_scattering_cross_section": (
        [ctypes.POINTER(Accessor2), Accessor3, ctypes.POINTER(Accessor3), ctypes.c_int, Rect3,
         ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "snap_hip_run_geometry_param": (
        [ctypes.POINTER(Accessor3), ctypes.POINTER(Accessor3), _dp, ctypes.c_int,
         ctypes.c_double, ctypes.c_double, ctypes.c_double, Rect3, ctypes.c_int, _vp],
        ctypes.c_int),
    "snap_hip_solver_create": (
        [ctypes.POINTER(Problem), ctypes.POINTER(SolveOptions), ctypes.POINTER(_vp)],
        ctypes.c_int),
    "snap_hip_solver_run": ([_vp, ctypes.POINTER(SolveStats)], ctypes.c_int),
    "snap_hip_solver_get_flux": ([_vp, _dp, _dp], ctypes.c_int),
    "snap_hip_solver_time_sweeps": (
        [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "snap_hip_solver_plan": ([_vp, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "snap_hip_solver_mms_compare": ([_vp, _dp], ctypes.c_int),
    "snap_hip_solver_get_log": ([_vp, ctypes.POINTER(IterationRecord), ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "snap_hip_mapper_tunables": ([ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int32)],
                                 ctypes.c_int),
    "snap_hip_solver_get_mms_reference": ([_vp, _dp], ctypes.c_int),
    "snap_hip_batch_create": (
        [ctypes.POINTER(Problem), ctypes.c_int, ctypes.POINTER(BatchOptions), ctypes.POINTER(_vp)],
        ctypes.c_int),
    "snap_hip_batch_run": ([_vp, ctypes.POINTER(BatchStats), ctypes.POINTER(ProblemStats)],
                           ctypes.c_int),
    "snap_hip_batch_get_flux": ([_vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
    "snap_hip_batch_destroy": ([_vp], None),
    "snap_hip_solver_gradient": (
        [_vp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp,
         ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "snap_hip_solver_destroy": ([_vp], None),
}


class SnapError(RuntimeError):
    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where} failed with {code}: {message}")
        self.code = code


_lib = None


def load():
    """Load libsnap_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python legion-snap_amd/build.py`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def check(code: int, where: str) -> None:
    if code != SNAP_OK:
        msg = load().snap_hip_last_error().decode(errors="replace")
        raise SnapError(code, where, msg)
