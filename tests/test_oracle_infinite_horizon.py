"""Pins the oracle's infinite-horizon filter and smoother (ops.py:796-1068 of the
reference; InfiniteHorizonGaussianProcess, basemodels.py:1257-1301).

* The Riccati iteration converges to scipy's DARE solution (filter form: A', H').
* With enough iterations, the filter's means after the initial transient, its
  steady-state covariances and the smoother's means and covariance in the middle of a
  long series equal the exact Kalman filter / RTS smoother's (the steady state they
  approximate).
* Heteroscedastic elements with constant noise equal the homoscedastic form; the
  associative form equals the sequential one.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import kernels as OK, ops as OP


def _st_kernel(M=3, seed=0):
    rng = np.random.default_rng(seed)
    z = rng.uniform(size=(M, 2))
    return OK.SpatioTemporalKernel(OK.Matern32(1.2, 4.0), OK.Matern32(1.0, 0.5), z)


@pytest.mark.parametrize("case", ["scalar", "spacetime"])
def test_dare_matches_scipy(case):
    if case == "scalar":
        k = OK.Matern32(1.0, 3.0)
        R = np.array([[0.1]])
    else:
        k = _st_kernel()
        G = np.random.default_rng(1).standard_normal((3, 3)) * 0.3
        R = G @ G.T + 0.2 * np.eye(3)
    Pinf = k.stationary_covariance()
    A = k.state_transition(1.0)
    Q = Pinf - A @ Pinf @ A.T
    H = k.measurement_model()
    X = OP.dare(A, H, Q, R, Pinf, num_iters=400)
    ref = scipy.linalg.solve_discrete_are(A.T, H.T, Q, R)
    np.testing.assert_allclose(X, ref, rtol=1e-8, atol=1e-11)


def _long(N, seed):
    rng = np.random.default_rng(seed)
    dt = np.r_[0.0, np.ones(N - 1)]
    y = np.sin(0.1 * np.arange(N))[:, None, None] + 0.3 * rng.standard_normal((N, 1, 1))
    return dt, y


def test_filter_and_smoother_reach_the_exact_steady_state():
    k = OK.Matern32(1.0, 3.0)
    N = 400
    dt, y = _long(N, 2)
    R = np.full((N, 1, 1), 0.1)
    ell, (fm, (Pd, cov)) = OP.kalman_filter_infinite_horizon(dt, k, y, R, noise_cov_tied=np.array([[0.1]]),
                                                            dare_iters=400)
    _, (em, eP) = OP.kalman_filter(dt, k, y, R)
    _, (_, pP) = OP.kalman_filter(dt, k, y, R, return_predict=True)
    np.testing.assert_allclose(fm[100:], em[100:], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(cov, eP[-1], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(Pd, pP[-1], rtol=1e-10, atol=1e-13)
    dts = np.r_[dt[1:], 0.0]
    sm, sP, G, dcov = OP.rauch_tung_striebel_smoother_infinite_horizon(dts, k, fm, (Pd, cov), return_full=True,
                                                                       dare_iters=400)
    esm, esP, _ = OP.rauch_tung_striebel_smoother(dts, k, em, eP, return_full=True)
    mid = slice(150, 250)
    np.testing.assert_allclose(sm[mid], esm[mid], rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(dcov, esP[N // 2], rtol=1e-8, atol=1e-12)
    hm, hP, _, _ = OP.rauch_tung_striebel_smoother_infinite_horizon(dts, k, fm, (Pd, cov), dare_iters=400)
    np.testing.assert_allclose(hm[:, 0, 0], sm[:, 0, 0], rtol=1e-12)
    np.testing.assert_allclose(hP[0], dcov[:1, :1], rtol=1e-12)


def test_heteroscedastic_with_tied_noise_equals_homoscedastic():
    k = _st_kernel()
    N = 30
    rng = np.random.default_rng(3)
    dt = np.r_[0.0, np.ones(N - 1)]
    y = rng.standard_normal((N, 3, 1))
    G = rng.standard_normal((3, 3)) * 0.3
    Rt = G @ G.T + 0.2 * np.eye(3)
    R = np.broadcast_to(Rt, (N, 3, 3)).copy()
    a = OP.kalman_filter_infinite_horizon(dt, k, y, R, noise_cov_tied=Rt, heteroscedastic=False)
    b = OP.kalman_filter_infinite_horizon(dt, k, y, R, noise_cov_tied=Rt, heteroscedastic=True)
    assert abs(a[0] - b[0]) <= 1e-12 * abs(a[0])
    np.testing.assert_allclose(a[1][0], b[1][0], rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("heteroscedastic", [False, True])
def test_associative_form_equals_sequential(heteroscedastic):
    k = _st_kernel()
    N = 37
    rng = np.random.default_rng(4)
    dt = np.r_[0.0, np.ones(N - 1)]
    y = rng.standard_normal((N, 3, 1))
    G = rng.standard_normal((N, 3, 3)) * 0.3
    R = G @ np.swapaxes(G, 1, 2) + 0.2 * np.eye(3)
    Rt = np.linalg.inv(np.mean(np.linalg.inv(R), axis=0))
    kw = dict(noise_cov_tied=Rt, heteroscedastic=heteroscedastic, dare_iters=7)
    s = OP.kalman_filter_infinite_horizon(dt, k, y, R, parallel=False, **kw)
    p = OP.kalman_filter_infinite_horizon(dt, k, y, R, parallel=True, **kw)
    assert abs(s[0] - p[0]) <= 1e-12 * abs(s[0])
    np.testing.assert_allclose(p[1][0], s[1][0], rtol=1e-10, atol=1e-12)
    dts = np.r_[dt[1:], 0.0]
    a = OP.rauch_tung_striebel_smoother_infinite_horizon(dts, k, s[1][0], s[1][1], parallel=False)
    b = OP.rauch_tung_striebel_smoother_infinite_horizon(dts, k, s[1][0], s[1][1], parallel=True)
    np.testing.assert_allclose(b[0], a[0], rtol=1e-10, atol=1e-12)
