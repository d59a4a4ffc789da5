"""Pins the oracle's power EP on space-time kernels (inference.py:238-325 with the cavity
of the projected sites, basemodels.py:230-262 and 667-674; the multivariate Gaussian
moment match of likelihoods.py:755-782).

* Several locations per step, Gaussian likelihood: one EP step at lr = 1 makes every
  observed site the exact likelihood (nat1 = y / s2, nat2 = 1 / s2) for any power: the
  joint scale factor times the joint derivatives of log N(y | m, s2 / a + V) is -I / s2.
  A dropped power, a transposed scale factor or a wrong cavity fails it.  The energy
  refuses, as the reference's does (inference.py:313-317).
* One location per step at an inducing point (C = 0), Gaussian likelihood, exact sites:
  the EP energy is -log p(y) of the dense GP for any power (the tilted distributions
  are exact; the pep constants make the powered terms cancel).  A dropped pep constant
  or a wrong 1 / power fails it at power 0.5.
* One location per step at the single inducing point with unit spatial variance (B = 1,
  C = 0): the model is the temporal Markov GP with the temporal kernel, so EP sites,
  posterior and energy equal the temporal oracle's (pinned in test_oracle_models.py)
  for non-Gaussian likelihoods, missing data and power < 1.
"""
import math

import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM


def _grid(g):
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    return np.stack([a.ravel(), b.ravel()], 1)


@pytest.mark.parametrize("power", [1.0, 0.5])
def test_gaussian_ep_sites_are_exact_for_several_locations(power):
    rng = np.random.default_rng(11)
    T_, R_ = 9, 6
    t = np.sort(rng.uniform(0, 10, T_))
    r = rng.uniform(size=(T_, R_, 2))
    y = rng.standard_normal((T_, R_))
    y[2, 3] = np.nan
    s2 = 0.2
    kern = OK.SpatioTemporalKernel(OK.Matern32(1.3, 2.0), OK.Matern32(0.8, 0.4), _grid(2), sparse=True,
                                   conditional="Full")
    o = OM.MarkovGP(kern, OL.Gaussian(s2), t[:, None], y, R=r, inference="EP", power=power)
    o.inference(lr=0.6)
    o.inference(lr=1.0)
    obs = ~np.isnan(y)
    n1 = o.nat1[..., 0]
    n2 = np.diagonal(o.nat2, axis1=1, axis2=2)
    np.testing.assert_allclose(n2[obs], 1.0 / s2, rtol=1e-10)
    np.testing.assert_allclose(n1[obs], y[obs] / s2, rtol=1e-10, atol=1e-10)
    # M != R: no pseudo mask, the missing output is moment-matched at the cavity mean
    assert abs(n2[2, 3] - 1.0 / s2) < 1e-10
    with pytest.raises(NotImplementedError):
        o.energy()


def _dense_neg_log_marginal(kt_var, ls, t, y, s2):
    kt = OK.Matern32(kt_var, ls)
    K = kt(t[:, None], t[:, None]) + s2 * np.eye(len(t))
    return 0.5 * (y @ np.linalg.solve(K, y) + np.linalg.slogdet(K)[1] + len(t) * math.log(2 * math.pi))


@pytest.mark.parametrize("g,tol", [(1, 1e-10), (2, 1e-5)])
@pytest.mark.parametrize("power", [1.0, 0.5])
def test_one_location_ep_energy_is_exact_marginal(g, tol, power):
    rng = np.random.default_rng(12)
    T_ = 14
    z = _grid(g)
    t = np.sort(rng.uniform(0, 10, T_))
    r = np.broadcast_to(z[-1], (T_, 1, 2)).copy()
    y = rng.standard_normal((T_, 1))
    vt, vs, ls, s2 = 1.3, 0.8, 2.0, 0.2
    kern = OK.SpatioTemporalKernel(OK.Matern32(vt, ls), OK.Matern32(vs, 0.4), z, sparse=True, conditional="Full")
    o = OM.MarkovGP(kern, OL.Gaussian(s2), t[:, None], y, R=r, inference="EP", power=power)
    o.inference(lr=1.0)
    exact = _dense_neg_log_marginal(vt * vs, ls, t, y[:, 0], s2)
    # g = 2: the reference's pseudo variance inv(B' nat2 B + 1e-12 I) has eigenvalues near
    # 1e12 when one location projects onto four inducing points, which costs the oracle
    # about six digits of the energy
    assert abs(o.energy() - exact) <= tol * abs(exact)


@pytest.mark.parametrize("lik,power", [("poisson", 1.0), ("probit", 0.5), ("studentt", 1.0)])
def test_one_location_at_inducing_point_equals_temporal_ep(lik, power):
    rng = np.random.default_rng(13)
    T_ = 25
    t = np.sort(rng.uniform(0, 20, T_))
    f = np.sin(0.4 * t)
    if lik == "poisson":
        y, ol = rng.poisson(np.exp(f)).astype(np.float64), OL.Poisson(1.0)
    elif lik == "probit":
        y, ol = (rng.uniform(size=T_) < 0.5 * (1 + np.tanh(f))).astype(np.float64), OL.Bernoulli()
    else:
        y, ol = f + 0.3 * rng.standard_t(3.0, size=T_), OL.StudentsT(0.3, 3.0)
    y[[3, 11]] = np.nan
    vt, ls = 1.4, 3.0
    kern = OK.SpatioTemporalKernel(OK.Matern32(vt, ls), OK.Matern32(1.0, 0.5), _grid(1), sparse=True,
                                   conditional="Full")
    r = np.zeros((T_, 1, 2))
    st = OM.MarkovGP(kern, ol, t[:, None], y[:, None], R=r, inference="EP", power=power)
    tm = OM.MarkovGP(OK.Matern32(vt, ls), ol, t[:, None], y[:, None], inference="EP", power=power)
    for lr in (1.0, 0.5, 0.8):
        st.inference(lr=lr)
        tm.inference(lr=lr)
        np.testing.assert_allclose(st.nat1.reshape(-1), tm.nat1.reshape(-1), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(st.nat2.reshape(-1), tm.nat2.reshape(-1), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(st.posterior_mean.reshape(-1), tm.posterior_mean[:, :1].reshape(-1),
                                   rtol=1e-9, atol=1e-12)
        assert abs(st.energy() - tm.energy()) <= 1e-9 * abs(tm.energy())
