"""Time-dependent solves and the manufactured solution on the GPU against the
CPU oracle (sweep.cc:414-534, snap.cc:386-559, mms.cc).

Bars: FP64 in the reference's operation order is bit-equal with identical
iteration counts; fused FP64 within 1e-12 relative; mixed precision (FP32
sweeps) within the bounds below; the mms.in deck passes the reference's MMS
acceptance on the GPU; the drop-in run_gpu_sweep with vdelt != 0 writes the
oracle's time-edge outflow bit for bit.
"""
import dataclasses
import os

import numpy as np
import pytest

from snap_amd import deck, problems, tasks
from snap_amd.solver import Solver

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


def _deck(name, **kw):
    p, extra = deck.read_deck(os.path.join(DECKS, name))
    return dataclasses.replace(p, **kw) if kw else p


def _td(p, nsteps, tf):
    return dataclasses.replace(p, timedep=True, nsteps=nsteps, tf=tf)


CASES = {
    "mms_deck": lambda: _deck("mms.in"),
    "tiny_512_deck": lambda: _deck("tiny_512_1x1x1.in"),
    "mms_steady": lambda: _deck("mms.in", timedep=False, nsteps=1, tf=1.0, nx=8, ny=8, nz=8,
                                mat=np.zeros((8, 8, 8), np.int32), qi=np.zeros((4, 8, 8, 8))),
    "td_config1_p1": lambda: _td(problems.config1(n=12, ng=4, nang=12, nmom=2), 3, 0.3),
    # two steps of dt = 0.5: long enough for the flux to reach every cell (a
    # shorter transient leaves cells near 1e-23, where relative error means little)
    "td_config3": lambda: _td(problems.config3(n=10, ng=3, nang=24), 2, 1.0),
    "td_nang80": lambda: _td(problems.config1(n=8, ng=2, nang=80, nmom=1), 2, 1.0),
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
        mms = (s.mms_compare(), s.mms_reference()) if p.mms else None
    return f, fm, st, mms


@pytest.mark.parametrize("name", list(CASES))
def test_time_dependent_fp64_exact_is_bit_equal(gpu, oracle_mod, name):
    fo, fmo, sto = oracle_solution(oracle_mod, name)
    p = CASES[name]()
    print("case", name, flush=True)
    f, fm, st, mms = gpu_solve(p, "fp64_exact")
    print(name, "fp64_exact", st, sto)
    assert (st["inner_total"], st["outer_total"], st["converged"], st["time_steps"]) == \
        (sto["inner_total"], sto["outer_total"], sto["converged"], sto["time_steps"])
    assert np.array_equal(f, fo)
    if fmo is not None:
        assert np.array_equal(fm, fmo)
    if p.mms:
        ref_o, _, _ = oracle_mod.mms_setup(p)
        assert np.array_equal(mms[1], ref_o)
        assert mms[0] == oracle_mod.mms_compare(fo, ref_o)


@pytest.mark.parametrize("name", list(CASES))
def test_time_dependent_fp64_fused_within_1e12(gpu, oracle_mod, name):
    fo, fmo, sto = oracle_solution(oracle_mod, name)
    print("case", name, flush=True)
    f, fm, st, _ = gpu_solve(CASES[name](), "fp64")
    err = float(np.max(np.abs(f - fo) / np.maximum(np.abs(fo), 1e-300)))
    print(name, "fp64", f"maxrel={err:.3e}")
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    assert err <= 1e-12


@pytest.mark.parametrize("name", ["mms_deck", "tiny_512_deck", "td_config1_p1"])
def test_time_dependent_mixed_precision(gpu, oracle_mod, name):
    fo, _, sto = oracle_solution(oracle_mod, name)
    print("case", name, flush=True)
    f, _, st, _ = gpu_solve(CASES[name](), "mixed")
    d = np.abs(f - fo)
    mx = float(np.max(d / np.maximum(np.abs(fo), 1e-300)))
    l2 = float(np.linalg.norm(d) / np.linalg.norm(fo))
    print(name, "mixed", f"maxrel={mx:.3e} l2rel={l2:.3e}", st["inner_total"], sto["inner_total"])
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    assert l2 <= 1e-6 and mx <= 1e-6


def test_mms_deck_passes_the_reference_acceptance(gpu):
    p, extra = deck.read_deck(os.path.join(DECKS, "mms.in"))
    f, _, st, (cmp, _) = gpu_solve(p, "fp64")
    mx, mn, mean = cmp
    counted = (p.nx * extra["ichunk"]) * (p.ny * extra["npey"]) * (p.nz * extra["npez"])
    print("mms", st, cmp)
    assert st["converged"] and st["time_steps"] == 2
    assert mx <= 0.1 and mean * p.ncells / counted <= 0.001 and mean <= 3e-3


@pytest.mark.parametrize("fixup", [False, True])
def test_dropin_time_dependent_sweep(gpu, oracle_mod, fixup):
    import torch
    nx, ny, nz, ng, nang, nmom = 7, 6, 5, 3, 12, 1
    rng = np.random.default_rng(77 + fixup)
    quad = oracle_mod.quadrature(nang, nmom)
    mu, eta, xi, w, ec = quad
    h = (2.0 / (0.7 / nx), 2.0 / (0.6 / ny), 2.0 / (0.5 / nz))
    tasks.initialize_gpu_context(quad, nang, nmom, (nx, ny, nz))
    txs = rng.uniform(0.5, 2.0, (ng, nz, ny, nx))
    vdelt = np.array([3.0, 1.5, 0.75])
    dinv = np.stack([oracle_mod.geometry_param(txs[g:g + 1], nang, h, mu, eta, xi,
                                               vdelt=np.ascontiguousarray(vdelt[g:g + 1]))[0]
                     for g in range(ng)])
    qtot = rng.uniform(-0.2 if fixup else 0.0, 1.0, (ng, nz, ny, nx, 1))
    tin = rng.uniform(0.0, 1.0, (ng, nz, ny, nx, nang))
    qim = rng.uniform(-0.5, 0.5, (ng, nz, ny, nx, nang))
    dev = torch.device("cuda")
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)  # noqa: E731
    for corner in (0, 3, 6):
        flux = torch.zeros((ng, nz, ny, nx), dtype=torch.float64, device=dev)
        tout = torch.full((ng, nz, ny, nx, nang), 7.0, dtype=torch.float64, device=dev)
        tasks.mini_kba(corner, range(ng), T(qtot), flux, None, T(dinv), T(txs), None, h, nang,
                       nmom, fixup=fixup, vdelt=vdelt, time_flux_in=T(tin), time_flux_out=tout,
                       qim=T(qim))
        torch.cuda.synchronize()
        for g in range(ng):
            pf, _, _, to, _ = oracle_mod.sweep_td(
                nx, ny, nz, nang, 1, corner, h, quad, np.ascontiguousarray(qtot[g]),
                np.ascontiguousarray(dinv[g]), np.ascontiguousarray(txs[g]), fixup, vdelt[g],
                tin[g], qim[g])
            assert np.array_equal(tout[g].cpu().numpy(), to), (corner, g)
            assert np.allclose(flux[g].cpu().numpy(), pf, rtol=1e-14, atol=0), (corner, g)


def test_time_dependent_problem_options_are_checked(gpu):
    from snap_amd import _lib
    p = _td(problems.config0(n=4, ng=2, nang=4), 2, 1.0)
    with pytest.raises(_lib.SnapError):
        Solver(p, precision="mixed", refine=True)
    with pytest.raises(_lib.SnapError):
        Solver(p, precision="fp64", dinv="recompute")
