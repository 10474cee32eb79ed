"""Oracle element-wise task bodies against independent numpy forms:
ExpandCrossSection / ExpandScatteringCrossSection (expxs.cc:85-115, 220-250)
as gathers, CalculateGeometryParam (expxs.cc:358-392) as an inverse,
CalcOuterSource (outer.cc:107-217) as a dense matrix product with the
in-group term removed, CalcInnerSource (inner.cc:106-160), and the
convergence tests (inner.cc:283-324, outer.cc:354-403) on hand-built pairs."""
import numpy as np
import pytest

from snap_amd import problems


def _rand_problem(seed=3, n=(5, 4, 3), ng=5, nmat=3, nmom=3):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    mat = rng.integers(0, nmat, size=(nz, ny, nx)).astype(np.int32)
    sigt = rng.uniform(0.5, 2.0, size=(nmat, ng))
    slgg = rng.uniform(0.0, 0.3, size=(ng, nmat, ng, nmom))
    qi = rng.uniform(0.0, 1.0, size=(ng, nz, ny, nx))
    return mat, sigt, slgg, qi


def test_expand_xs(oracle_mod):
    mat, sigt, slgg, qi = _rand_problem()
    xs = oracle_mod.expand_xs(mat, sigt)
    assert np.array_equal(xs, np.moveaxis(sigt[mat], -1, 0))


def test_expand_scat_xs(oracle_mod):
    mat, sigt, slgg, qi = _rand_problem()
    sxs = oracle_mod.expand_scat_xs(mat, slgg)
    ng = slgg.shape[0]
    for g in range(ng):
        assert np.array_equal(sxs[g], slgg[g, :, g, :][mat])


def test_geometry_param_is_inverse(oracle_mod):
    mat, sigt, slgg, qi = _rand_problem()
    xs = oracle_mod.expand_xs(mat, sigt)
    mu, eta, xi, w, ec = oracle_mod.quadrature(7, 1)
    h = (8.0, 6.0, 4.0)
    dinv = oracle_mod.geometry_param(xs, 7, h, mu, eta, xi)
    den = xs[..., None] + h[0] * mu + h[1] * eta + h[2] * xi
    assert np.allclose(dinv * den, 1.0, rtol=0, atol=4e-16)


def test_outer_source_dense_product(oracle_mod):
    mat, sigt, slgg, qi = _rand_problem(nmom=3)
    ng, nmat, _, nmom = slgg.shape
    cmom = nmom * nmom
    rng = np.random.default_rng(11)
    flux0 = rng.uniform(0, 1, size=qi.shape)
    fluxm = rng.uniform(-1, 1, size=qi.shape + (cmom - 1,))
    q0, qm = oracle_mod.outer_source(mat, qi, slgg, flux0, fluxm)
    # S[m][l][g_to][g_from] with the in-group element removed
    S = np.transpose(slgg, (1, 3, 0, 2)).copy()
    for g in range(ng):
        S[:, :, g, g] = 0.0
    cells_mat = mat.reshape(-1)
    f0 = flux0.reshape(ng, -1)
    ref0 = qi.reshape(ng, -1).copy()
    for c in range(f0.shape[1]):
        ref0[:, c] += S[cells_mat[c], 0] @ f0[:, c]
    assert np.allclose(q0.reshape(ng, -1), ref0, rtol=1e-14, atol=1e-15)
    order = [0, 1, 1, 1, 2, 2, 2, 2, 2]
    fm = fluxm.reshape(ng, -1, cmom - 1)
    refm = np.zeros_like(fm)
    for c in range(fm.shape[1]):
        for m in range(1, cmom):
            refm[:, c, m - 1] = S[cells_mat[c], order[m]] @ fm[:, c, m - 1]
    assert np.allclose(qm.reshape(ng, -1, cmom - 1), refm, rtol=1e-13, atol=1e-15)


def test_inner_source(oracle_mod):
    mat, sigt, slgg, qi = _rand_problem(nmom=2)
    ng = slgg.shape[0]
    sxs = oracle_mod.expand_scat_xs(mat, slgg)
    rng = np.random.default_rng(5)
    flux0 = rng.uniform(0, 1, size=qi.shape)
    q2grp0 = rng.uniform(0, 1, size=qi.shape)
    fluxm = rng.uniform(-1, 1, size=qi.shape + (3,))
    q2grpm = rng.uniform(-1, 1, size=qi.shape + (3,))
    qtot = oracle_mod.inner_source(sxs, flux0, q2grp0, fluxm, q2grpm)
    assert np.allclose(qtot[..., 0], q2grp0 + flux0 * sxs[..., 0], rtol=1e-15, atol=0)
    for m in range(1, 4):
        assert np.allclose(qtot[..., m], q2grpm[..., m - 1] + fluxm[..., m - 1] * sxs[..., 1],
                           rtol=1e-15, atol=1e-16)


@pytest.mark.parametrize("ratio,expect", [(1.0 + 0.99e-4, True), (1.0 + 1.01e-4, False),
                                          (1.0 - 1.01e-4, False)])
def test_convergence_threshold(oracle_mod, ratio, expect):
    prev = np.full(10, 2.5)
    cur = prev.copy()
    cur[7] = prev[7] * ratio
    assert oracle_mod.converged(cur, prev, 1e-4) is expect


def test_convergence_small_previous_uses_absolute(oracle_mod):
    prev = np.array([1.0, 1e-13])
    assert oracle_mod.converged(np.array([1.0, 0.99e-4]), prev, 1e-4)
    assert not oracle_mod.converged(np.array([1.0, 1.01e-4]), prev, 1e-4)


def test_convergence_nan_counts_as_converged(oracle_mod):
    # df > epsi is false for NaN (inner.cc:318-319)
    assert oracle_mod.converged(np.array([np.nan, 1.0]), np.array([1.0, 1.0]), 1e-4)
