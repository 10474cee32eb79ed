"""The C ABI library: built, loadable, exports every entry point declared in
include/snap_hip.h, and rejects invalid arguments before touching a GPU."""
import ctypes
import os
import re

import pytest

from snap_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "snap_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(snap_hip_\w+)\s*\(", text)))


def test_library_loads():
    lib = _lib.load()
    assert lib.snap_hip_version().decode().startswith("snap_hip")


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 18
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(_lib.SIGNATURES), "ctypes table out of sync with the header"


def test_struct_sizes_match_header_layout():
    # natural alignment of the ABI structs (x86-64 SysV)
    assert ctypes.sizeof(_lib.Accessor3) == 32
    assert ctypes.sizeof(_lib.Accessor2) == 24
    assert ctypes.sizeof(_lib.Rect3) == 48


def test_solver_rejects_null_problem():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.snap_hip_solver_create(None, None, ctypes.byref(h))
    assert rc == _lib.SNAP_ERR_INVALID
    assert b"NULL" in lib.snap_hip_last_error()


def test_sweep_rejects_incomplete_arguments():
    lib = _lib.load()
    a = _lib.SweepArgs()
    a.mms_source = 1  # without a qim accessor
    assert lib.snap_hip_run_gpu_sweep(ctypes.byref(a), None) == _lib.SNAP_ERR_INVALID
    a.mms_source = 0
    a.vdelt = 0.5  # without time-edge flux accessors
    assert lib.snap_hip_run_gpu_sweep(ctypes.byref(a), None) == _lib.SNAP_ERR_INVALID
    a.vdelt = 0.0
    a.corner = 3
    a.stride_x_positive = 0  # corner 3 sweeps +x
    assert lib.snap_hip_run_gpu_sweep(ctypes.byref(a), None) == _lib.SNAP_ERR_INVALID


def test_check_raises_with_message():
    with pytest.raises(_lib.SnapError) as e:
        _lib.check(_lib.SNAP_ERR_INVALID, "x")
    assert "x failed" in str(e.value)


def test_library_quadrature_is_the_oracles(oracle_mod):
    """snap_hip_quadrature (host code, no GPU) against the oracle's
    compute_derived_globals restatement, P0 to P3."""
    import numpy as np
    from snap_amd import tasks
    for nang, nmom in ((10, 1), (36, 4), (7, 2), (1024, 1)):
        a = tasks.quadrature(nang, nmom)
        b = oracle_mod.quadrature(nang, nmom)
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), (nang, nmom)


# ctypes mirror -> C typedef of include/snap_hip.h
MIRRORS = {
    "Accessor1": "snap_accessor1_t", "Accessor2": "snap_accessor2_t",
    "Accessor3": "snap_accessor3_t", "Rect3": "snap_rect3_t", "SweepArgs": "snap_sweep_args_t",
    "Problem": "snap_problem_t", "SolveOptions": "snap_solve_options_t",
    "SolveStats": "snap_solve_stats_t", "IterationRecord": "snap_iteration_record_t",
    "BatchOptions": "snap_batch_options_t", "ProblemStats": "snap_problem_stats_t",
    "BatchStats": "snap_batch_stats_t",
}


def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler
    gives the header's struct (a C program built with gcc prints them)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "snap_hip.h"', "int main(void) {"]
    for py, c in MIRRORS.items():
        lines.append(f'  printf("{py} %zu\\n", sizeof({c}));')
        for f, _ in getattr(_lib, py)._fields_:
            lines.append(f'  printf("{py}.{f} %zu\\n", offsetof({c}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for py in MIRRORS:
        cls = getattr(_lib, py)
        assert int(got[py]) == ctypes.sizeof(cls), py
        for f, _ in cls._fields_:
            assert int(got[f"{py}.{f}"]) == getattr(cls, f).offset, f"{py}.{f}"
