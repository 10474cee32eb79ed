"""The one-cell-per-workgroup sweep (angles across up to 1024 lanes, the form
every solve past 128 angles per octant takes) against the CPU oracle.

Bars as for the tiled sweep: FP64 in the reference's operation order is
bit-equal with identical iteration counts, fused FP64 within 1e-12 relative.
Cases: the tiled sweep's own cases forced onto the wide form, angle counts
past 128 (steady and time dependent), the reference's small / medium decks
(128 and 256 angles, 10 time steps) and the drop-in run_gpu_sweep at 256
angles with ghost faces and time-edge fluxes.
"""
import dataclasses
import os

import numpy as np
import pytest

from snap_amd import deck, problems, tasks
from snap_amd.solver import Solver

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


def _deck(name):
    return deck.read_deck(os.path.join(DECKS, name))[0]


CASES = {
    "config0": (lambda: problems.config0(), True),
    "config1_p3": (lambda: problems.config1(n=10, ng=4, nang=12, nmom=4), True),
    "odd_shape": (lambda: problems.config1(n=9, ng=3, nang=7, nmom=1), True),
    "nang256": (lambda: problems.config1(n=8, ng=3, nang=256, nmom=2), False),
    "nang1100_apl2": (lambda: problems.config1(n=4, ng=2, nang=1100, nmom=2), False),
    "mms_deck": (lambda: _deck("mms.in"), True),
    "small_512_deck": (lambda: _deck("small_512_1x1x1.in"), True),
    "medium_512_deck": (lambda: _deck("medium_512_1x1x1.in"), False),
}

_oracle_cache = {}


def oracle_solution(oracle_mod, name):
    if name not in _oracle_cache:
        _oracle_cache[name] = oracle_mod.solve(CASES[name][0]())
    return _oracle_cache[name]


def gpu_solve(p, precision, force):
    with Solver(p, precision=precision, wide_sweep=force) as s:
        plan = s.plan()
        st = s.run()
        f, fm = s.flux()
    return f, fm, st, plan


@pytest.mark.parametrize("name", list(CASES))
def test_wide_fp64_exact_is_bit_equal(gpu, oracle_mod, name):
    fo, fmo, sto = oracle_solution(oracle_mod, name)
    make, force = CASES[name]
    print("case", name, flush=True)
    f, fm, st, plan = gpu_solve(make(), "fp64_exact", force)
    assert plan["tile"] == (1, 1, 1) and plan["cells_per_wave"] == 0
    assert (st["inner_total"], st["outer_total"], st["converged"], st["time_steps"]) == \
        (sto["inner_total"], sto["outer_total"], sto["converged"], sto["time_steps"])
    assert np.array_equal(f, fo)
    if fmo is not None:
        assert np.array_equal(fm, fmo)


@pytest.mark.parametrize("name", list(CASES))
def test_wide_fp64_fused_within_1e12(gpu, oracle_mod, name):
    fo, fmo, sto = oracle_solution(oracle_mod, name)
    make, force = CASES[name]
    print("case", name, flush=True)
    f, fm, st, _ = gpu_solve(make(), "fp64", force)
    err = float(np.max(np.abs(f - fo) / np.maximum(np.abs(fo), 1e-300)))
    print(name, f"fp64 maxrel={err:.3e}")
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    assert err <= 1e-12
    if fmo is not None:
        assert np.max(np.abs(fm - fmo)) <= 1e-12 * np.max(np.abs(fmo))


@pytest.mark.parametrize("name", ["config1_p3", "nang256", "small_512_deck"])
def test_wide_mixed_precision(gpu, oracle_mod, name):
    fo, _, _ = oracle_solution(oracle_mod, name)
    make, force = CASES[name]
    print("case", name, flush=True)
    f, _, st, _ = gpu_solve(make(), "mixed", force)
    d = np.abs(f - fo)
    mx = float(np.max(d / np.maximum(np.abs(fo), 1e-300)))
    l2 = float(np.linalg.norm(d) / np.linalg.norm(fo))
    print(name, f"mixed maxrel={mx:.3e} l2rel={l2:.3e}")
    assert l2 <= 1e-6 and mx <= 1e-6


@pytest.mark.parametrize("timedep", [False, True])
def test_dropin_wide_sweep(gpu, oracle_mod, timedep):
    import torch
    nx, ny, nz, ng, nang, nmom = 5, 4, 6, 2, 256, 2
    cmom = nmom * nmom
    rng = np.random.default_rng(5 + timedep)
    quad = oracle_mod.quadrature(nang, nmom)
    mu, eta, xi, w, ec = quad
    h = (2.0 / (0.5 / nx), 2.0 / (0.4 / ny), 2.0 / (0.6 / nz))
    tasks.initialize_gpu_context(quad, nang, nmom, (nx, ny, nz))
    txs = rng.uniform(0.5, 2.0, (ng, nz, ny, nx))
    vdelt = np.array([2.0, 0.5]) if timedep else np.zeros(ng)
    dinv = np.stack([oracle_mod.geometry_param(txs[g:g + 1], nang, h, mu, eta, xi,
                                               vdelt=np.ascontiguousarray(vdelt[g:g + 1]))[0]
                     for g in range(ng)])
    qtot = rng.uniform(-0.1, 1.0, (ng, nz, ny, nx, cmom))
    qtot[..., 1:] *= 0.2
    tin = rng.uniform(0.0, 1.0, (ng, nz, ny, nx, nang))
    dev = torch.device("cuda")
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)  # noqa: E731
    for corner in (1, 6):
        gin = (rng.uniform(0, 1, (ng, nz, ny, nang)), rng.uniform(0, 1, (ng, nz, nx, nang)),
               rng.uniform(0, 1, (ng, ny, nx, nang)))
        gd = tuple(T(x) for x in gin)
        flux = torch.zeros((ng, nz, ny, nx), dtype=torch.float64, device=dev)
        fluxm = torch.zeros((ng, nz, ny, nx, cmom - 1), dtype=torch.float64, device=dev)
        tout = torch.zeros((ng, nz, ny, nx, nang), dtype=torch.float64, device=dev)
        kw = dict(vdelt=vdelt, time_flux_in=T(tin), time_flux_out=tout) if timedep else {}
        tasks.mini_kba(corner, range(ng), T(qtot), flux, fluxm, T(dinv), T(txs), gd, h, nang,
                       nmom, **kw)
        torch.cuda.synchronize()
        for g in range(ng):
            if timedep:
                # sweep_td starts from vacuum: compare the time path without ghosts
                continue
            pf, pm, _, outs = oracle_mod.sweep(
                nx, ny, nz, nang, cmom, corner, h, quad, np.ascontiguousarray(qtot[g]),
                np.ascontiguousarray(dinv[g]), np.ascontiguousarray(txs[g]), fixup=True,
                ghost_in=tuple(np.ascontiguousarray(x[g]) for x in gin), want_ghost_out=True)
            assert np.allclose(flux[g].cpu().numpy(), pf, rtol=1e-14, atol=0), (corner, g)
            assert np.allclose(fluxm[g].cpu().numpy(), pm, rtol=1e-13, atol=1e-16), (corner, g)
            for kk in range(3):
                assert np.array_equal(gd[kk][g].cpu().numpy(), outs[kk]), (corner, g, kk)
    if timedep:
        # vacuum ghosts and the time-edge outflow against the oracle
        for corner in (0, 5):
            flux = torch.zeros((ng, nz, ny, nx), dtype=torch.float64, device=dev)
            fluxm = torch.zeros((ng, nz, ny, nx, cmom - 1), dtype=torch.float64, device=dev)
            tout = torch.zeros((ng, nz, ny, nx, nang), dtype=torch.float64, device=dev)
            tasks.mini_kba(corner, range(ng), T(qtot), flux, fluxm, T(dinv), T(txs), None, h,
                           nang, nmom, vdelt=vdelt, time_flux_in=T(tin), time_flux_out=tout)
            torch.cuda.synchronize()
            for g in range(ng):
                pf, pm, _, to, _ = oracle_mod.sweep_td(
                    nx, ny, nz, nang, cmom, corner, h, quad, np.ascontiguousarray(qtot[g]),
                    np.ascontiguousarray(dinv[g]), np.ascontiguousarray(txs[g]), True, vdelt[g],
                    tin[g])
                assert np.array_equal(tout[g].cpu().numpy(), to), (corner, g)
                assert np.allclose(flux[g].cpu().numpy(), pf, rtol=1e-14, atol=0), (corner, g)
