"""GPU solve (libsnap_hip's Legion-free transport_solve) against the CPU oracle.

Bars, from the requirement on this path:
* FP64 in the reference's operation order ("fp64_exact"): bit-equal scalar
  flux and flux moments, identical inner/outer counts;
* FP64 with fused multiply-adds ("fp64"): <= 1e-12 relative, identical counts;
* mixed precision (FP32 sweeps, FP64 moments/sources/tests, the inners that
  end an outer in FP64): max and L2 relative error <= 1e-6 against the FP64
  oracle, identical counts (the measured errors are printed).
"""
import os

import numpy as np
import pytest

from snap_amd import problems
from snap_amd.solver import Solver

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = {
    "config0": lambda: problems.config0(),
    "config1_small": lambda: problems.config1(n=16, ng=8, nang=12, nmom=4),
    "config1_ragged": lambda: problems.config1(n=20, ng=6, nang=36, nmom=2),
    "config3_small": lambda: problems.config3(n=16, ng=4, nang=24),
    "p2_nang10": lambda: problems.config1(n=24, ng=4, nang=10, nmom=3),
    "odd_shape": lambda: problems.config1(n=9, ng=3, nang=7, nmom=1),
    "nang80": lambda: problems.config1(n=10, ng=2, nang=80, nmom=2),
}

_oracle_cache = {}


def oracle_solution(oracle_mod, name):
    if name not in _oracle_cache:
        _oracle_cache[name] = oracle_mod.solve(CASES[name]())
    return _oracle_cache[name]


def gpu_solve(p, precision, **kw):
    with Solver(p, precision=precision, **kw) as s:
        st = s.run()
        f, fm = s.flux()
    return f, fm, st


def rel_errors(a, b):
    d = np.abs(a - b)
    return float(np.max(d / np.maximum(np.abs(b), 1e-300))), float(np.linalg.norm(d) / np.linalg.norm(b))


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_exact_is_bit_equal(gpu, oracle_mod, name):
    fo, fmo, sto = oracle_solution(oracle_mod, name)
    f, fm, st = gpu_solve(CASES[name](), "fp64_exact")
    print(name, "fp64_exact", st)
    assert (st["inner_total"], st["outer_total"], st["converged"]) == \
        (sto["inner_total"], sto["outer_total"], sto["converged"])
    assert np.array_equal(f, fo)
    if fmo is not None:
        assert np.array_equal(fm, fmo)


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_fused_within_1e12(gpu, oracle_mod, name):
    fo, fmo, sto = oracle_solution(oracle_mod, name)
    f, fm, st = gpu_solve(CASES[name](), "fp64")
    mx, l2 = rel_errors(f, fo)
    print(name, "fp64", f"maxrel={mx:.3e} l2rel={l2:.3e}")
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    assert mx <= 1e-12
    if fmo is not None:
        assert np.max(np.abs(fm - fmo)) <= 1e-12 * np.max(np.abs(fmo))


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_precision_error(gpu, oracle_mod, name):
    # the default: FP32 sweeps with FP64 polish sweeps at the end of each outer
    fo, fmo, sto = oracle_solution(oracle_mod, name)
    f, fm, st = gpu_solve(CASES[name](), "mixed")
    mx, l2 = rel_errors(f, fo)
    print(name, "mixed", f"maxrel={mx:.3e} l2rel={l2:.3e}", st)
    assert st["sweeps_fp32"] > 0 and st["sweeps_fp32"] + st["sweeps_fp64"] == st["inner_total"]
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    assert mx <= 1e-6 and l2 <= 1e-6


@pytest.mark.parametrize("name", ["config0", "config1_small", "config1_ragged"])
def test_fp32_sweeps_without_polish(gpu, oracle_mod, name):
    # every sweep in FP32 (fp64_polish = 0): a regression guard for the FP32
    # kernels only -- this form's rounding error (about 4e-6 max on the 64^3
    # workload) does not meet the 1e-6 requirement, which the polished form
    # above is tested against
    fo, _, sto = oracle_solution(oracle_mod, name)
    f, _, st = gpu_solve(CASES[name](), "mixed", refine=False, fp64_polish=0.0)
    mx, l2 = rel_errors(f, fo)
    print(name, "fp32 sweeps", f"maxrel={mx:.3e} l2rel={l2:.3e}")
    assert st["sweeps_fp32"] == st["inner_total"]
    assert mx <= 5e-6 and l2 <= 1e-6


@pytest.mark.parametrize("name", ["config1_small", "config1_ragged", "odd_shape"])
@pytest.mark.parametrize("apl", [1, 2])
def test_angles_per_lane_keep_exact_results(gpu, oracle_mod, name, apl):
    # one or two angles per lane (and the cells per wavefront that follow)
    # change only the lane layout, never the arithmetic
    fo, fmo, _ = oracle_solution(oracle_mod, name)
    with Solver(CASES[name](), precision="fp64_exact", angles_per_lane=apl) as s:
        s.run()
        f, fm = s.flux()
        assert s.plan()["angles_per_lane"] == apl
    assert np.array_equal(f, fo)
    if fmo is not None:
        assert np.array_equal(fm, fmo)


def test_mixed_defect_correction(gpu, oracle_mod):
    # config 0 has few fixed cells: sweeping only the FP64 source defect in
    # FP32 brings the FP32 error from ~1e-6 down by orders of magnitude
    fo, _, sto = oracle_solution(oracle_mod, "config0")
    f, _, st = gpu_solve(CASES["config0"](), "mixed", refine=True, fp64_polish=0.0)
    mx, l2 = rel_errors(f, fo)
    print("config0 mixed refine", f"maxrel={mx:.3e} l2rel={l2:.3e}", st)
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    assert st["sweeps_fp64"] == st["outer_total"]
    assert mx <= 1e-8


@pytest.mark.parametrize("name", ["config1_small", "config1_ragged", "odd_shape"])
def test_outer_source_matrix_cores(gpu, oracle_mod, name):
    # the FP64 matrix-core outer source against the CPU-ordered one: the
    # same products summed in another order
    fo, fmo, sto = oracle_solution(oracle_mod, name)
    f1, fm1, st1 = gpu_solve(CASES[name](), "fp64", outer_ordered=True)
    f2, fm2, st2 = gpu_solve(CASES[name](), "fp64")
    assert (st1["inner_total"], st1["outer_total"]) == (st2["inner_total"], st2["outer_total"])
    mx, _ = rel_errors(f2, f1)
    assert mx <= 1e-13
    if fm1 is not None:
        assert np.max(np.abs(fm2 - fm1)) <= 1e-13 * np.max(np.abs(fm1))


@pytest.mark.parametrize("ng", [40, 56])
def test_outer_source_matrix_cores_many_groups(gpu, ng):
    # P3 with 40 groups stages its outputs in LDS over two passes of 16-group
    # blocks; with 56 groups no block fits beside the fluxes and every lane
    # stores directly.  Both against the CPU-ordered outer source.
    p = problems.config1(n=6, ng=ng, nang=4, nmom=4)
    f1, fm1, st1 = gpu_solve(p, "fp64", outer_ordered=True)
    f2, fm2, st2 = gpu_solve(problems.config1(n=6, ng=ng, nang=4, nmom=4), "fp64")
    assert (st1["inner_total"], st1["outer_total"]) == (st2["inner_total"], st2["outer_total"])
    mx, _ = rel_errors(f2, f1)
    assert mx <= 1e-13
    assert np.max(np.abs(fm2 - fm1)) <= 1e-13 * np.max(np.abs(fm1))


@pytest.mark.parametrize("kw", [{"first_sweep_fp64": True}, {"dinv": "recompute"}])
def test_mixed_options(gpu, oracle_mod, kw):
    # the first inner's sweep in FP64, or dinv formed in the FP32 sweep from
    # t_xs: same iteration counts, errors within the mixed-precision bar
    fo, _, sto = oracle_solution(oracle_mod, "config1_small")
    f, _, st = gpu_solve(CASES["config1_small"](), "mixed", **kw)
    mx, l2 = rel_errors(f, fo)
    print(kw, f"maxrel={mx:.3e} l2rel={l2:.3e}", st)
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    if kw.get("first_sweep_fp64"):
        assert st["sweeps_fp64"] >= 1
    assert mx <= 1e-6 and l2 <= 1e-6


def test_dinv_modes_agree(gpu):
    # forming dinv in the kernel is the same arithmetic as the geometry kernel
    p = CASES["config1_small"]()
    f1, _, _ = gpu_solve(p, "fp64_exact", dinv="stream")
    f2, _, _ = gpu_solve(p, "fp64_exact", dinv="recompute")
    assert np.array_equal(f1, f2)


@pytest.mark.parametrize("tile,waves", [((4, 4, 4), 4), ((8, 4, 2), 2), ((5, 3, 7), 4),
                                        ((16, 16, 1), 8)])
def test_tile_shapes_do_not_change_results(gpu, oracle_mod, tile, waves):
    fo, _, _ = oracle_solution(oracle_mod, "config1_ragged")
    f, _, _ = gpu_solve(CASES["config1_ragged"](), "fp64_exact", tile=tile, waves_per_tile=waves)
    assert np.array_equal(f, fo)


def test_golden_fixtures(gpu):
    g = np.load(os.path.join(GOLDEN, "config0.npz"), allow_pickle=False)
    f, _, st = gpu_solve(problems.config0(), "fp64_exact")
    assert np.array_equal(f, g["flux0"]) and st["inner_total"] == int(g["inner_total"])
    g = np.load(os.path.join(GOLDEN, "config1_small.npz"), allow_pickle=False)
    f, fm, st = gpu_solve(CASES["config1_small"](), "fp64_exact")
    idx = g["cells"]
    assert np.array_equal(f.reshape(f.shape[0], -1)[:, idx], g["flux0_samples"])
    assert np.array_equal(fm.reshape(fm.shape[0], -1, fm.shape[-1])[:, idx, :], g["fluxm_samples"])


def test_solver_rejects_bad_input(gpu):
    from snap_amd import _lib
    p = problems.config0(n=4, ng=2, nang=4)
    p.mat[0, 0, 0] = 5
    with pytest.raises(ValueError):
        Solver(p)
    p = problems.config0(n=4, ng=2, nang=2049)  # more angles than a 1024-lane workgroup holds twice
    with pytest.raises(_lib.SnapError) as e:
        Solver(p)
    assert e.value.code == _lib.SNAP_ERR_UNSUPPORTED
