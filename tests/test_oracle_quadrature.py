"""Oracle quadrature (compute_derived_globals, snap.cc:1171-1342) against
closed forms: mu_i = (i+1/2)/n, eta_i = 1 - mu_i = mu_{n-1-i},
xi_i = sqrt(2 mu_i (1-mu_i)) (unit directions), sum of the 8*n weights = 1,
and the octant sign pattern of the moment expansion coefficients."""
import numpy as np
import pytest


@pytest.mark.parametrize("nang", [1, 2, 10, 24, 36, 64])
def test_closed_forms(oracle_mod, nang):
    mu, eta, xi, w, ec = oracle_mod.quadrature(nang, 1)
    i = np.arange(nang)
    assert np.allclose(mu, (i + 0.5) / nang, rtol=0, atol=1e-14)
    assert np.allclose(eta, 1.0 - mu, rtol=0, atol=1e-14)
    assert np.allclose(eta, mu[::-1], rtol=0, atol=1e-14)
    assert np.allclose(xi, np.sqrt(2.0 * mu * (1.0 - mu)), rtol=1e-13, atol=0)
    assert np.allclose(mu ** 2 + eta ** 2 + xi ** 2, 1.0, rtol=0, atol=1e-14)
    assert abs(8.0 * w.sum() - 1.0) < 1e-14
    assert np.all(ec[:, 0, :] == 1.0)


def test_single_angle_values(oracle_mod):
    mu, eta, xi, w, ec = oracle_mod.quadrature(1, 1)
    assert mu[0] == 0.5 and eta[0] == 0.5
    assert abs(xi[0] - np.sqrt(0.5)) < 1e-16
    assert w[0] == 0.125


def test_p1_octant_signs(oracle_mod):
    # octant bit0 = +x (is = +1), bit1 = +y (js), bit2 = +z (ks)
    nang = 6
    mu, eta, xi, w, ec = oracle_mod.quadrature(nang, 2)
    assert ec.shape == (8, 4, nang)
    for oct_ in range(8):
        is_ = 1 if oct_ & 1 else -1
        js = 1 if oct_ & 2 else -1
        ks = 1 if oct_ & 4 else -1
        assert np.allclose(ec[oct_, 1], is_ * mu, atol=1e-15)
        assert np.allclose(ec[oct_, 2], is_ * mu * ks * js * xi * eta, atol=1e-15)
        assert np.allclose(ec[oct_, 3], is_ * mu * (xi * eta) ** 2, atol=1e-15)
    # odd in the x direction: the weighted sum over all octants vanishes
    assert abs(np.sum(ec[:, 1, :] * w)) < 1e-15


def test_p3_shape_and_orders(oracle_mod):
    mu, eta, xi, w, ec = oracle_mod.quadrature(36, 4)
    assert ec.shape == (8, 16, 36)
    # moment 4 is the first l=2 moment: (is*mu)^3
    assert np.allclose(ec[7, 4], mu ** 3, rtol=1e-14)
    assert np.allclose(ec[0, 4], -(mu ** 3), rtol=1e-14)
    # moment 9 is the first l=3 moment: (is*mu)^5
    assert np.allclose(ec[7, 9], mu ** 5, rtol=1e-13)


def test_bad_requests(oracle_mod):
    with pytest.raises(ValueError):
        oracle_mod.quadrature(4, 5)
