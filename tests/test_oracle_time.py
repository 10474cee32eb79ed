"""The oracle's time-dependent solve and manufactured solution (sweep.cc:414-534,
snap.cc:386-559, mms.cc), pinned by:

* the reference's own known-answer test: the mms.in deck (16^3 cells, 4
  groups, 16 angles, 2 steps of dt = 1) must pass MMSCompare's thresholds
  (snap.cc:542-559: max |phi/ref - 1| <= 0.1 and mean <= 0.001, the mean taken
  over nx*ichunk * ny*npey * nz*npez cells as snap.cc:546 counts them);
* second-order convergence of the diamond-difference error under mesh
  refinement (the manufactured flux is linear in time, so the time
  differencing adds no error);
* the particle balance of one time-dependent sweep without the fixup: streamed
  leakage + collisions + (v dt)^-1 (psi_t,out - psi_t,in) = source, summed
  over cells and angles;
* the limits: a huge time step is the steady solve, and a long transient of a
  fixed source settles onto the steady solution.
"""
import dataclasses
import os

import numpy as np
import pytest

from snap_amd import deck, problems

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def _mms(n=None):
    p, extra = deck.read_deck(os.path.join(DECKS, "mms.in"))
    if n is not None:
        mat, qi = deck.reference_layouts(n, n, n, p.ng, 0, 3)
        p = dataclasses.replace(p, nx=n, ny=n, nz=n, mat=mat, qi=qi)
    return p, extra


def test_mms_deck_passes_the_reference_acceptance(oracle_mod):
    p, extra = _mms()
    assert p.timedep and p.mms and p.nsteps == 2 and p.tf == 2.0
    flux, _, st = oracle_mod.solve(p)
    assert st["converged"] and st["time_steps"] == 2
    ref, _, _ = oracle_mod.mms_setup(p)
    mx, mn, mean = oracle_mod.mms_compare(flux, ref)
    counted = (p.nx * extra["ichunk"]) * (p.ny * extra["npey"]) * (p.nz * extra["npez"])
    ref_mean = mean * p.ncells / counted
    assert mx <= 0.1 and ref_mean <= 0.001
    # the true mean over the 16^3 cells, a tighter bound than the reference's
    assert mean <= 3e-3 and mn >= 0.0


def test_mms_error_is_second_order(oracle_mod):
    means = []
    for n in (8, 16, 32):
        p, _ = _mms(n)
        flux, _, _ = oracle_mod.solve(p)
        ref, _, _ = oracle_mod.mms_setup(p)
        means.append(oracle_mod.mms_compare(flux, ref)[2])
    assert means[0] / means[1] > 3.0 and means[1] / means[2] > 3.0, means


def test_mms_source_terms(oracle_mod):
    """The manufactured flux is g * (cell average of sin(pi x/L)) per axis,
    and with time dependence the fixed source is phi / v_g with phi scaled to
    tf - dt/2 (mms.cc:130-148, 382-409)."""
    p, _ = _mms(8)
    ref, qim, qi = oracle_mod.mms_setup(p)
    dx = p.lx / p.nx
    e = np.arange(p.nx + 1) * dx
    a = 3.14159265358979 / p.lx
    avg = (np.cos(a * e[:-1]) - np.cos(a * e[1:])) / (a * dx)
    dt = p.tf / p.nsteps
    g = np.arange(1, p.ng + 1)[:, None, None, None]
    expect = g * avg[None, None, None, :] * avg[None, None, :, None] * avg[None, :, None, None]
    np.testing.assert_allclose(ref, expect * (p.tf - 0.5 * dt), rtol=1e-13)
    v = (p.ng - np.arange(p.ng))[:, None, None, None]
    np.testing.assert_allclose(qi, expect / v, rtol=1e-13)
    # the angular source summed over all directions (the streaming terms
    # cancel between opposite octants) is the collision minus scattering term
    mu, eta, xi, w, _ = oracle_mod.quadrature(p.nang, 1)
    total = np.einsum("ogzyxa,a->gzyx", qim, w)
    S = p.slgg[:, 0, :, 0]
    coll = p.sigt[0][:, None, None, None] * expect
    scat = np.einsum("gh,hzyx->gzyx", S, expect)
    np.testing.assert_allclose(total, coll - scat, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("corner", [0, 5, 7])
def test_time_dependent_sweep_balance(oracle_mod, corner):
    rng = np.random.default_rng(11 + corner)
    nx, ny, nz, nang = 5, 4, 3, 6
    mu, eta, xi, w, ec = oracle_mod.quadrature(nang, 1)
    lx, ly, lz = 0.5, 0.4, 0.3
    h = (2.0 * nx / lx, 2.0 * ny / ly, 2.0 * nz / lz)
    nc = nx * ny * nz
    sig = 0.5 + rng.random(nc)
    vdelt = 3.7
    q = rng.random(nc)
    tin = rng.random((nc, nang))
    qim = rng.random((nc, nang))
    dinv = 1.0 / ((((sig[:, None] + vdelt) + h[0] * mu) + h[1] * eta) + h[2] * xi)
    pf, _, _, tout, (gx, gy, gz) = oracle_mod.sweep_td(
        nx, ny, nz, nang, 1, corner, h, (mu, eta, xi, w, ec), q, np.ascontiguousarray(dinv),
        sig, False, vdelt, tin.reshape(nz, ny, nx, nang), qim.reshape(nz, ny, nx, nang))
    pc = pf.ravel()  # sum_a w_a pc_a per cell
    tout = tout.reshape(nc, nang)
    leak = (np.sum(gx * (w * mu)) * h[0] + np.sum(gy * (w * eta)) * h[1]
            + np.sum(gz * (w * xi)) * h[2]) / 2.0
    coll = np.dot(sig, pc)
    dtime = vdelt / 2.0 * np.sum((tout - tin) * w)
    src = np.sum(q) * np.sum(w) + np.sum(qim * w)
    assert abs(leak + coll + dtime - src) <= 1e-12 * src
    # the time-edge closure: psi_t,out = 2 pc - psi_t,in for every cell and angle
    # (weighted sum per cell)
    np.testing.assert_allclose(((tout + tin) * w).sum(axis=1) / 2.0, pc, rtol=1e-12)


def _small_td(nsteps, tf):
    p = problems.config0(n=6, ng=3, nang=4)
    return dataclasses.replace(p, timedep=True, tf=tf, nsteps=nsteps, epsi=1e-8, iitm=50,
                               oitm=400)


def test_huge_time_step_is_the_steady_solve(oracle_mod):
    p = _small_td(1, 1e14)
    f_td, _, st = oracle_mod.solve(p)
    f_ss, _, _ = oracle_mod.solve(dataclasses.replace(p, timedep=False))
    assert st["converged"]
    np.testing.assert_allclose(f_td, f_ss, rtol=1e-9)


def test_long_transient_settles_on_the_steady_solution(oracle_mod):
    # dt = 0.1: the diamond difference in time damps a mode of decay rate L by
    # (1 - L dt/2)/(1 + L dt/2) per step, so dt must be short against the
    # fastest streaming modes (L ~ v * mu * hi ~ 40 here) for them to die out
    p = _small_td(100, 10.0)
    f_td, _, st = oracle_mod.solve(p)
    assert st["time_steps"] == 100 and st["converged"]
    f_ss, _, _ = oracle_mod.solve(dataclasses.replace(p, timedep=False, nsteps=1))
    np.testing.assert_allclose(f_td, f_ss, rtol=1e-8)
    # an early step is still far from it
    f_1, _, _ = oracle_mod.solve(dataclasses.replace(p, nsteps=2, tf=0.2))
    assert np.max(np.abs(f_1 / f_ss - 1.0)) > 1e-2


def test_mms_compare_semantics(oracle_mod):
    """MMSCompare (mms.cc:515-533): |phi/ref - 1|, or |phi| where ref < 1e-12."""
    ref = np.array([2.0, 4.0, 1e-13, 0.0])
    flux = np.array([2.2, 3.0, 0.25, -0.5])
    mx, mn, mean = oracle_mod.mms_compare(flux, ref)
    d = np.array([0.1, 0.25, 0.25, 0.5])
    assert np.isclose(mx, 0.5) and np.isclose(mn, 0.1) and np.isclose(mean, d.mean())
