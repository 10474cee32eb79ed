"""The one-wavefront tile sweep (csrc/snap_sweep_seq.hip) in the solver,
against the CPU oracle: the exact form bit-equal with identical counts, the
fused form within 1e-12, for every moment count it instantiates (P0..P3),
ragged tiles, several tile depths, dinv streamed or formed in the kernel."""
import numpy as np
import pytest

from snap_amd import problems
from snap_amd.solver import Solver

pytestmark = pytest.mark.gpu

CASES = {
    "config0": lambda: problems.config0(),
    "config1_small": lambda: problems.config1(n=16, ng=8, nang=12, nmom=4),
    "config1_ragged": lambda: problems.config1(n=20, ng=6, nang=36, nmom=2),
    "config3_small": lambda: problems.config3(n=16, ng=4, nang=24),
    "p2_nang10": lambda: problems.config1(n=24, ng=4, nang=10, nmom=3),
    "odd_shape": lambda: problems.config1(n=9, ng=3, nang=7, nmom=1),
    "nang64": lambda: problems.config1(n=10, ng=2, nang=64, nmom=4),
}

_cache = {}


def _oracle(oracle_mod, name):
    if name not in _cache:
        _cache[name] = oracle_mod.solve(CASES[name]())
    return _cache[name]


def _solve(name, precision, **kw):
    with Solver(CASES[name](), precision=precision, sweep_kernel="wave_tile", **kw) as s:
        st = s.run()
        f, fm = s.flux()
        plan = s.plan()
    assert plan["waves"] == 1 and tuple(plan["tile"][:2]) == (4, 4), plan
    return f, fm, st


@pytest.mark.parametrize("name", list(CASES))
def test_exact_bit_equal(gpu, oracle_mod, name):
    fo, fmo, sto = _oracle(oracle_mod, name)
    f, fm, st = _solve(name, "fp64_exact")
    assert (st["inner_total"], st["outer_total"], st["converged"]) == \
        (sto["inner_total"], sto["outer_total"], sto["converged"])
    assert np.array_equal(f, fo)
    if fmo is not None:
        assert np.array_equal(fm, fmo)


@pytest.mark.parametrize("name", list(CASES))
def test_fused_within_1e12(gpu, oracle_mod, name):
    fo, fmo, sto = _oracle(oracle_mod, name)
    f, fm, st = _solve(name, "fp64")
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    rel = float(np.max(np.abs(f - fo) / np.abs(fo)))
    print(name, f"fused maxrel={rel:.2e}")
    assert rel <= 1e-12
    if fmo is not None:
        assert np.max(np.abs(fm - fmo)) <= 1e-12 * np.max(np.abs(fmo))


@pytest.mark.parametrize("tz", [1, 3, 5, 16])
@pytest.mark.parametrize("dinv", ["stream", "recompute"])
def test_tile_depths_and_dinv_forms(gpu, oracle_mod, tz, dinv):
    fo, fmo, _ = _oracle(oracle_mod, "config1_ragged")
    f, fm, _ = _solve("config1_ragged", "fp64_exact", tile=(0, 0, tz), dinv=dinv)
    assert np.array_equal(f, fo) and np.array_equal(fm, fmo)


@pytest.mark.parametrize("name", ["config1_small", "config1_ragged", "odd_shape"])
def test_mixed_within_bar(gpu, oracle_mod, name):
    fo, _, sto = _oracle(oracle_mod, name)
    f, _, st = _solve(name, "mixed")
    rel = float(np.max(np.abs(f - fo) / np.abs(fo)))
    l2 = float(np.linalg.norm(f - fo) / np.linalg.norm(fo))
    print(name, f"mixed maxrel={rel:.2e} l2={l2:.2e}")
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    assert rel <= 1e-6 and l2 <= 1e-6


def test_adjoint_with_wave_tile_solver(gpu, oracle_mod):
    # the gradient rebuilds its checkpoint faces on the multi-wave tile plan
    # even when the solver sweeps with one wavefront per tile
    import dataclasses
    p = dataclasses.replace(problems.config3(n=8, ng=3, nang=6, block=2), epsi=1e-13, iitm=200,
                            oitm=400)
    r = np.zeros((p.ng, p.nz, p.ny, p.nx))
    r[:, -2:, -2:, -2:] = 1.0
    fo, _, _ = oracle_mod.solve(p)
    go, _, _ = oracle_mod.adjoint(p, fo, r)
    with Solver(p, precision="fp64", sweep_kernel="wave_tile") as s:
        s.run()
        g, _, _ = s.gradient(r)
    assert np.max(np.abs(g - go)) <= 1e-9 * np.max(np.abs(go))
