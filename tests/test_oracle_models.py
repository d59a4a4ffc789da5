"""Pins the oracle's Markov-GP models (bayesnewton/basemodels.py:625-764,
bayesnewton/inference.py:21-325) with closed forms and the dense GP:

* Gaussian likelihood + one inference(lr=1) step: VI, Newton and EP (any power)
  all reach the exact posterior and energy == -log N(y | 0, K + s2 I)
  (pattern of tests/test_vs_exact_marg_lik.py:41-65)
* non-conjugate VI / Newton / EP: after each step the Markov posterior equals
  the dense GP posterior with the same sites, and the VI energy equals the
  dense-algebra energy
* missing data (NaN y) follows inference.py:27-29
"""
import numpy as np
import pytest

from oracle import kernels as K, likelihoods as L, models
from oracle.utils import LOG2PI


def _data(n=40, seed=0):
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(0, 25, n))
    y = np.sin(0.5 * x) + 0.3 * rng.standard_normal(n)
    return x, y


@pytest.mark.parametrize("kind,power", [("VI", 1.0), ("Newton", 1.0), ("EP", 1.0), ("EP", 0.5)])
def test_gaussian_one_step_is_exact(kind, power):
    x, y = _data()
    kern = K.Sum([K.Matern32(0.6, 2.0), K.Matern52(0.8, 0.9)])
    m = models.MarkovGP(kern, L.Gaussian(0.2), x, y, inference=kind, power=power)
    m.inference(lr=1.0)
    dense = models.DenseGP(kern, x, y, np.full(len(x), 0.2))
    mean, cov = dense.posterior()
    np.testing.assert_allclose(m.posterior_mean[:, 0, 0], mean, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(m.posterior_variance[:, 0, 0], np.diag(cov), rtol=1e-8, atol=1e-12)
    E = m.energy()
    if kind == "Newton":
        # Laplace energy with a Gaussian likelihood: -(sum log p(y|m) - KL)
        ll = np.sum(-0.5 * (LOG2PI + np.log(0.2) + (y - mean) ** 2 / 0.2))
        ed = np.sum(-0.5 * (LOG2PI + np.log(0.2) + ((y - mean) ** 2 + np.diag(cov)) / 0.2))
        assert abs(E - (-(ll - (ed - dense.log_marginal())))) < 1e-8 * abs(E)
    else:
        assert abs(E + dense.log_marginal()) < 1e-8 * abs(E)


def _dense_vi_energy(kern, x, y, lik, site_mean, site_var):
    dense = models.DenseGP(kern, x, site_mean, site_var)
    mean, cov = dense.posterior()
    v = np.diag(cov)
    ell, _, _ = lik.variational_expectation(y, mean, v)
    ed = -0.5 * (LOG2PI + np.log(site_var) + ((site_mean - mean) ** 2 + v) / site_var)
    return -(np.sum(ell) - (np.sum(ed) - dense.log_marginal())), mean, v


@pytest.mark.parametrize("lik", [L.Bernoulli(), L.StudentsT(0.3, 3.0), L.Poisson(1.0)],
                         ids=["probit", "studentt", "poisson"])
@pytest.mark.parametrize("kind", ["VI", "Newton", "EP"])
def test_nonconjugate_markov_equals_dense(lik, kind):
    x, y = _data(35, 1)
    if isinstance(lik, L.Bernoulli):
        y = (y > 0).astype(float)
    if isinstance(lik, L.Poisson):
        y = np.round(np.exp(0.5 * y)).astype(float)
    kern = K.Matern52(1.0, 1.5)
    m = models.MarkovGP(kern, lik, x, y, inference=kind)
    for _ in range(3):
        m.inference(lr=0.7)
        dense = models.DenseGP(kern, x, m.site_mean, m.site_cov[:, 0, 0])
        mean, cov = dense.posterior()
        np.testing.assert_allclose(m.posterior_mean[:, 0, 0], mean, rtol=1e-8, atol=1e-10)
        np.testing.assert_allclose(m.posterior_variance[:, 0, 0], np.diag(cov), rtol=1e-7, atol=1e-12)
    if kind == "VI":
        E = m.energy()
        Ed, _, _ = _dense_vi_energy(kern, x, y, lik, m.site_mean[:, 0, 0], m.site_cov[:, 0, 0])
        assert abs(E - Ed) < 1e-8 * abs(E)


def test_vi_sites_are_natural_gradient_step():
    """Sites after inference(lr): nat = (1-lr) nat_old + lr (dE - d2E m, -d2E) (inference.py:21-39, 79-86)."""
    x, y = _data(20, 2)
    lik = L.StudentsT(0.3, 3.0)
    kern = K.Matern32(1.0, 2.0)
    m = models.MarkovGP(kern, lik, x, y)
    m.update_posterior()
    pm, pv = m.posterior_mean[:, 0, 0].copy(), m.posterior_variance[:, 0, 0].copy()
    n1, n2 = m.nat1[:, 0, 0].copy(), m.nat2[:, 0, 0].copy()
    _, dE, d2E = lik.variational_expectation(y, pm, pv)
    d2E = np.where(d2E > 0, -1e-2, d2E)  # ensure_diagonal_positive_precision on -d2E
    lr = 0.3
    m.inference(lr=lr)
    np.testing.assert_allclose(m.nat1[:, 0, 0], (1 - lr) * n1 + lr * (dE - d2E * pm), rtol=1e-12)
    np.testing.assert_allclose(m.nat2[:, 0, 0], (1 - lr) * n2 + lr * (-d2E), rtol=1e-12)


def test_missing_data_sites():
    """NaN y: hessian -> -1e-6, jacobian -> hessian m, so nat1 = 0, nat2 = 1e-6 (inference.py:27-29)."""
    x, y = _data(15, 3)
    y[[2, 9]] = np.nan
    m = models.MarkovGP(K.Matern52(1.0, 1.0), L.Gaussian(0.1), x, y)
    m.inference(lr=1.0)
    assert np.allclose(m.nat1[[2, 9], 0, 0], 0.0) and np.allclose(m.nat2[[2, 9], 0, 0], 1e-6)
    assert np.isfinite(m.energy())
