"""Pins the oracle's prediction path (oracle/models.py predict / predict_y /
negative_log_predictive_density, oracle/likelihoods.py predict and log_density):

* the Markov-GP predictive at new times (basemodels.py:766-816 with utils.py:99-215)
  equals the dense GP predictive k*' (K + R)^-1 y, k** - k*' (K + R)^-1 k* for the same
  sites, at test times before, between, on and after the training times;
* Likelihood.predict (cubature.py:438-465) against the closed-form predictive moments
  (probit: (1 - 2e) Phi(m / sqrt(1 + v)) + e; Poisson: lognormal moments; Student-t:
  (m, v + s^2 df / (df - 2))), and log_density against adaptive quadrature."""
import math

import numpy as np
from scipy import integrate
from scipy.special import erf

from oracle import kernels as OK, likelihoods as OL, models as OM


def _dense_predictive(kern, X, y, r, Xs):
    K = kern(X, X)
    Ks = kern(Xs, X)
    Kss = np.diag(kern(Xs, Xs))
    L = np.linalg.cholesky(K + np.diag(r))
    a = np.linalg.solve(L.T, np.linalg.solve(L, y))
    V = np.linalg.solve(L, Ks.T)
    return Ks @ a, Kss - np.sum(V * V, axis=0)


def test_markov_predict_equals_dense_gp_predictive():
    rng = np.random.default_rng(0)
    x = np.sort(rng.uniform(0, 10, 60))
    y = np.sin(x) + 0.1 * rng.standard_normal(60)
    for kern in (OK.Matern52(0.7, 1.3), OK.Sum([OK.Matern52(0.8, 1.5), OK.Matern32(0.4, 4.0)]),
                 OK.Matern12(1.1, 0.6)):
        m = OM.MarkovGP(kern, OL.Gaussian(0.05), x, y, inference="VI")
        m.inference(lr=1.0)  # exact Gaussian sites N(y, 0.05)
        xs = np.concatenate([[-3.0, -0.01], rng.uniform(0, 10, 25), x[[0, 7, 59]], [10.2, 14.0]])
        pm, pv = m.predict(xs[:, None])
        dm, dv = _dense_predictive(kern, x[:, None], y, np.full(60, 0.05), xs[:, None])
        np.testing.assert_allclose(pm, dm, rtol=0, atol=1e-7 * np.max(np.abs(dm)))
        np.testing.assert_allclose(pv, dv, rtol=1e-6, atol=1e-9)


def test_predict_at_training_times_is_the_smoothed_posterior():
    rng = np.random.default_rng(1)
    x = np.sort(rng.uniform(0, 20, 40))
    y = (rng.uniform(size=40) < 0.5).astype(float)
    m = OM.MarkovGP(OK.Matern32(1.0, 2.0), OL.Bernoulli(), x, y, inference="VI")
    for _ in range(3):
        m.inference(lr=1.0)
    m.update_posterior()
    pm, pv = m.predict()
    # exact up to the reference's 1e-8 jitter on Q-+ (utils.py:207)
    np.testing.assert_allclose(pm, m.posterior_mean[:, 0, 0], rtol=0, atol=1e-7)
    np.testing.assert_allclose(pv, m.posterior_variance[:, 0, 0], rtol=0, atol=1e-7)


def test_likelihood_predictive_moments_closed_forms():
    m = np.array([-1.3, 0.0, 0.4, 2.1])
    v = np.array([0.05, 0.7, 1.5, 0.2])
    eps = 1e-3
    Ey, Vy = OL.Bernoulli().predict(m, v)
    p = (1 - 2 * eps) * 0.5 * (1 + erf(m / np.sqrt(1 + v) / math.sqrt(2))) + eps
    # 20-point Gauss-Hermite error on the probit integrand is ~5e-9 here
    np.testing.assert_allclose(Ey, p, rtol=1e-7)
    np.testing.assert_allclose(Vy, p - p ** 2, rtol=1e-7)
    Ey, Vy = OL.Poisson(binsize=2.0).predict(m, v)
    mu = 2.0 * np.exp(m + v / 2)
    np.testing.assert_allclose(Ey, mu, rtol=1e-8)
    np.testing.assert_allclose(Vy, mu + 4.0 * np.exp(2 * m + v) * (np.exp(v) - 1), rtol=1e-7)
    Ey, Vy = OL.StudentsT(scale=0.3, df=4.0).predict(m, v)
    np.testing.assert_allclose(Ey, m, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(Vy, v + 0.09 * 4.0 / 2.0, rtol=1e-12)
    Ey, Vy = OL.Gaussian(0.2).predict(m, v)
    np.testing.assert_allclose(Vy, v + 0.2)


def test_log_density_matches_quadrature():
    """log_density is log sum_i w_i p(y | m + sqrt(v) x_i) (cubature.py:249-276): against
    adaptive quadrature of log int p(y|f) N(f|m, v) df where Gauss-Hermite converges (v
    small against the likelihood's width), to 1e-6 with the reference's 20 nodes."""
    cases = [(OL.Bernoulli(), np.array([1.0, 0.0, 1.0])), (OL.Poisson(binsize=2.0), np.array([0.0, 3.0, 7.0])),
             (OL.StudentsT(scale=0.3, df=4.0), np.array([-0.5, 0.2, 1.7])), (OL.Gaussian(0.2), np.array([0.1, -1.0, 2.0]))]
    m = np.array([-0.4, 0.3, 1.1])
    v = np.array([0.004, 0.01, 0.02])
    for lik, y in cases:
        ld = lik.log_density(y, m, v)
        for i in range(3):
            f = lambda z, i=i: math.exp(float(lik.log_lik(y[i], z))) * math.exp(-0.5 * (z - m[i]) ** 2 / v[i]) / math.sqrt(2 * math.pi * v[i])  # noqa: E731
            Z, _ = integrate.quad(f, m[i] - 12 * math.sqrt(v[i]), m[i] + 12 * math.sqrt(v[i]), epsabs=1e-14, epsrel=1e-12, limit=200)
            assert abs(ld[i] - math.log(Z)) < 1e-6, (type(lik).__name__, i, ld[i], math.log(Z))


def test_spacetime_predict_matches_dense_posterior():
    """Space-time prediction (basemodels.py:766-816 on the spatio-temporal branch): with a
    Gaussian likelihood one VI step gives the exact posterior of the inducing processes
    u(t) (a GP with covariance k_t(t, t') I_M) under y_t = B_t u(t) + noise, so the
    predictive mean and variance at (t*, r*) are b*' E[u(t*) | y] and
    b*' Cov(u(t*) | y) b* + c*, computed here from the dense joint covariance."""
    from oracle import kernels as OK, likelihoods as OL, models as OM
    rng = np.random.default_rng(8)
    g, T_, R_, s2 = 2, 10, 5, 0.2
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    z = np.stack([a.ravel(), b.ravel()], 1)
    M = z.shape[0]
    t = np.sort(rng.uniform(0, 8, T_))
    r = rng.uniform(size=(T_, R_, 2))
    y = rng.standard_normal((T_, R_))
    kt, ks = OK.Matern32(1.3, 2.0), OK.Matern32(0.8, 0.4)
    kern = OK.SpatioTemporalKernel(kt, ks, z)
    o = OM.MarkovGP(kern, OL.Gaussian(s2), t[:, None], y, R=r, inference="VI")
    o.inference(lr=1.0)
    ts = np.array([-1.0, t[0], 0.5 * (t[2] + t[3]), t[6], 9.5])
    rs = rng.uniform(size=(ts.shape[0], 3, 2))
    pm, pv = o.predict(ts[:, None], rs)
    # dense: u at the training times and one test time
    B, _ = kern.spatial_conditional(t[:, None], r)
    Bs, Cs = kern.spatial_conditional(ts[:, None], rs, predict=True)
    Bbig = np.zeros((T_ * R_, T_ * M))
    for k in range(T_):
        Bbig[k * R_:(k + 1) * R_, k * M:(k + 1) * M] = B[k]
    Kt = kt(t[:, None], t[:, None])
    Sy = Bbig @ np.kron(Kt, np.eye(M)) @ Bbig.T + s2 * np.eye(T_ * R_)
    for i, tt in enumerate(ts):
        kx = kt(np.array([[tt]]), t[:, None])
        Cuy = np.kron(kx, np.eye(M)) @ Bbig.T
        mu = Cuy @ np.linalg.solve(Sy, y.reshape(-1))
        Su = kt(np.array([[tt]]), np.array([[tt]]))[0, 0] * np.eye(M) - Cuy @ np.linalg.solve(Sy, Cuy.T)
        np.testing.assert_allclose(pm[i], Bs[i] @ mu, rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(pv[i], np.diagonal(Bs[i] @ Su @ Bs[i].T + Cs[i]), rtol=1e-6, atol=1e-7)


def _decoupled_spacetime(seed, T_):
    from oracle import kernels as OK, likelihoods as OL
    rng = np.random.default_rng(seed)
    gz = np.linspace(0, 1, 3)
    a, b = np.meshgrid(gz, gz)
    z = np.stack([a.ravel(), b.ravel()], 1)
    t = np.arange(T_, dtype=np.float64)
    r = np.broadcast_to(z, (T_,) + z.shape).copy()
    y = rng.poisson(2.0, size=(T_, z.shape[0])).astype(np.float64)
    kt, ks = OK.Matern32(1.2, 4.0), OK.Matern32(0.7, 1e-3)
    return OK.SpatioTemporalKernel(kt, ks, z), OL.Poisson(), t, r, y, z


def test_meanfield_predict_equals_dense_when_space_decouples():
    """Mean-field prediction (basemodels.py:1177-1228): with a spatial kernel that decouples
    the inducing points and data at them, the mean-field model is exact, so its predictions
    equal the dense model's (both forms of the filter)."""
    from oracle import models as OM
    kern, lik, t, r, y, z = _decoupled_spacetime(9, 12)
    ts = np.array([-1.0, 0.0, 2.5, 7.0, 13.0])
    rs = np.broadcast_to(z[:4], (ts.shape[0], 4, 2)).copy()
    for par in (False, True):
        dense = OM.MarkovGP(kern, lik, t[:, None], y, R=r, inference="VI")
        mf = OM.MarkovGP(kern, lik, t[:, None], y, R=r, inference="VI", mean_field=True, parallel=par)
        for o in (dense, mf):
            o.inference(lr=1.0)
        dm, dv = dense.predict(ts[:, None], rs)
        mm, mv = mf.predict(ts[:, None], rs)
        np.testing.assert_allclose(mm, dm, rtol=1e-8, atol=1e-10)
        np.testing.assert_allclose(mv, dv, rtol=1e-8, atol=1e-10)


def test_infinite_horizon_predict_reaches_the_dense_steady_state():
    """Infinite-horizon prediction (temporal_conditional_infinite_horizon, utils.py:142-177):
    with many Riccati iterations, in the middle of a long equispaced series and with the
    sites fixed, its predictive mean and variance equal the dense model's."""
    from oracle import kernels as OK, likelihoods as OL, models as OM
    rng = np.random.default_rng(10)
    T_ = 300
    t = np.arange(T_, dtype=np.float64)
    z = np.array([[0.2, 0.3], [0.7, 0.6]])
    r = np.broadcast_to(z, (T_,) + z.shape).copy()
    y = np.sin(0.1 * t)[:, None] + 0.3 * rng.standard_normal((T_, 2))
    kern = OK.SpatioTemporalKernel(OK.Matern32(1.0, 3.0), OK.Matern32(1.0, 0.5), z)
    dense = OM.MarkovGP(kern, OL.Gaussian(0.1), t[:, None], y, R=r, inference="VI")
    ih = OM.MarkovGP(kern, OL.Gaussian(0.1), t[:, None], y, R=r, inference="VI", infinite_horizon=True,
                     dare_iters=400)
    nat1, nat2 = dense.nat1, dense.nat2
    for o in (dense, ih):
        o.update_nat_params(np.full_like(nat1, 0.0) + y[..., None] / 0.1, np.broadcast_to(np.eye(2) / 0.1, nat2.shape).copy())
    ts = np.array([140.0, 150.5, 160.25])
    rs = np.broadcast_to(z, (3, 2, 2)).copy()
    dm, dv = dense.predict(ts[:, None], rs)
    im, iv = ih.predict(ts[:, None], rs)
    np.testing.assert_allclose(im, dm, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(iv[0], dv[0], rtol=1e-7)
