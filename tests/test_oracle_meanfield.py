"""Pins the oracle's mean-field space-time filter and smoother (ops.py:429-706 of the
reference; MarkovMeanFieldGaussianProcess, basemodels.py:1158-1175).

* Diagonal pseudo-noise: the blocks never couple, so both mean-field filters equal the
  dense Kalman filter of the Kronecker state (whose off-diagonal blocks stay zero).
* Step 0 with coupled noise: the filtered blocks are the diagonal blocks of the exact
  Gaussian conditional of the stationary prior on y_0, computed in information form
  (a different route from the filter's gain form), and the log term is log N(y_0|0, S).
* Parallel form with coupled noise: each block equals an ordinary scalar-observation
  Kalman filter whose per-step noise and observation are r_i = 1/[S^-1]_ii - h Q_i h'
  and y_i = [S^-1 y]_i / [S^-1]_ii with S = diag(h Q h') + R.  The associative scan of
  the reference's element construction must reproduce it.
* Sequential form: permuting the spatial components permutes the results.
* Smoother: per block the ordinary RTS smoother of the temporal kernel.
* Model: with a spatial kernel that decouples the inducing points and data at them,
  the mean-field energy equals the dense model's.
"""
import numpy as np
import pytest
from scipy.stats import multivariate_normal

from oracle import kernels as OK, likelihoods as OL, models as OM, ops as OP


def _problem(seed, M, N, kt=None, coupled=True):
    rng = np.random.default_rng(seed)
    kt = kt or OK.Matern32(1.3, 2.0)
    z = rng.uniform(size=(M, 2))
    kern = OK.SpatioTemporalKernel(kt, OK.Matern32(1.0, 0.4), z)
    dt = np.concatenate([[0.0], rng.uniform(0.1, 1.5, N - 1)])
    y = rng.standard_normal((N, M, 1))
    if coupled:
        G = rng.standard_normal((N, M, M)) * 0.4
        R = G @ np.swapaxes(G, -1, -2) + np.eye(M) * rng.uniform(0.05, 0.5, (N, 1, 1))
    else:
        R = np.zeros((N, M, M))
        R[:, np.arange(M), np.arange(M)] = rng.uniform(0.05, 0.8, (N, M))
    return kern, dt, y, R


@pytest.mark.parametrize("parallel", [False, True])
@pytest.mark.parametrize("temporal", ["m12", "m32"])
def test_diagonal_noise_equals_dense_filter(parallel, temporal):
    kt = OK.Matern12(0.9, 1.1) if temporal == "m12" else OK.Matern32(1.3, 2.0)
    M, N = 4, 17
    kern, dt, y, R = _problem(1, M, N, kt, coupled=False)
    d = kt.state_dim if hasattr(kt, "state_dim") else kt.stationary_covariance().shape[0]
    ell, (fm, fP) = OP.kalman_filter_meanfield(dt, kern, y, R, parallel=parallel)
    ell_d, (fm_d, fP_d) = OP.kalman_filter(dt, kern, y, R, parallel=False)
    assert abs(ell - ell_d) <= 1e-10 * abs(ell_d)
    np.testing.assert_allclose(fm.reshape(N, -1), fm_d[..., 0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(fP, OP.diag_blocks(fP_d, M, d), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(OP.block_diag(fP), fP_d, atol=1e-12)


@pytest.mark.parametrize("parallel", [False, True])
def test_first_step_is_projected_exact_conditional(parallel):
    M, N = 5, 6
    kern, dt, y, R = _problem(2, M, N)
    d = 2
    ell, (fm, fP) = OP.kalman_filter_meanfield(dt, kern, y, R, parallel=parallel)
    Pinf = kern.stationary_covariance()
    H = kern.measurement_model()
    Ri = np.linalg.inv(R[0])
    prec = np.linalg.inv(Pinf) + H.T @ Ri @ H
    cov = np.linalg.inv(prec)
    mean = cov @ H.T @ Ri @ y[0]
    np.testing.assert_allclose(fm[0].reshape(-1), mean[:, 0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(fP[0], OP.diag_blocks(cov, M, d), rtol=1e-9, atol=1e-12)
    # the first log term alone: a one-step problem
    ell1, _ = OP.kalman_filter_meanfield(dt[:1], kern, y[:1], R[:1], parallel=parallel)
    ref = multivariate_normal(np.zeros(M), H @ Pinf @ H.T + R[0]).logpdf(y[0, :, 0])
    assert abs(ell1 - ref) <= 1e-11 * abs(ref)


def test_parallel_form_is_decoupled_scalar_filters():
    M, N = 4, 23
    kern, dt, y, R = _problem(3, M, N)
    kt = kern.temporal_kernel
    ell, (fm, fP) = OP.kalman_filter_meanfield(dt, kern, y, R, parallel=True)
    Pt = kt.stationary_covariance()
    At = kt.state_transition_batch(dt)
    Qt = Pt - At @ Pt @ np.swapaxes(At, -1, -2)
    Qt[0] = Pt
    h = kt.measurement_model()
    hQh = (h @ Qt @ h.T)[:, 0, 0]
    S = R + hQh[:, None, None] * np.eye(M)
    Si = np.linalg.inv(S)
    s = np.diagonal(Si, axis1=1, axis2=2)
    u = (Si @ y)[..., 0]
    r_eff = 1.0 / s - hQh[:, None]
    y_eff = u / s
    assert np.all(r_eff > 0)
    for i in range(M):
        _, (m_i, P_i) = OP.kalman_filter(dt, kt, y_eff[:, i, None, None], r_eff[:, i, None, None])
        np.testing.assert_allclose(fm[:, i], m_i, rtol=1e-8, atol=1e-11)
        np.testing.assert_allclose(fP[:, i], P_i, rtol=1e-8, atol=1e-11)
    # log likelihood: dense S from the predictions of the decoupled filters
    ll = 0.0
    mprev, Pprev = np.zeros((M, 2, 1)), np.broadcast_to(Pt, (M, 2, 2))
    for n in range(N):
        A = At[n]
        Q = Pt - A @ Pt @ A.T
        mp = A @ mprev
        Pp = A @ Pprev @ A.T + Q
        Sd = R[n] + np.diag((h @ Pp @ h.T)[:, 0, 0])
        ll += multivariate_normal((h @ mp)[:, 0, 0], Sd).logpdf(y[n, :, 0])
        mprev, Pprev = fm[n], fP[n]
    assert abs(ell - ll) <= 1e-10 * abs(ll)


def test_sequential_form_is_permutation_equivariant():
    M, N = 5, 12
    kern, dt, y, R = _problem(4, M, N)
    perm = np.array([3, 0, 4, 1, 2])
    ell, (fm, fP) = OP.kalman_filter_meanfield(dt, kern, y, R, parallel=False)
    ell_p, (fm_p, fP_p) = OP.kalman_filter_meanfield(dt, kern, y[:, perm], R[:, perm][:, :, perm],
                                                     parallel=False)
    assert abs(ell - ell_p) <= 1e-11 * abs(ell)
    np.testing.assert_allclose(fm_p, fm[:, perm], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(fP_p, fP[:, perm], rtol=1e-10, atol=1e-13)
    # coupled noise makes the two forms differ (the parallel one is its own approximation)
    ell_par, _ = OP.kalman_filter_meanfield(dt, kern, y, R, parallel=True)
    assert abs(ell - ell_par) > 1e-6 * abs(ell)


@pytest.mark.parametrize("parallel", [False, True])
@pytest.mark.parametrize("return_full", [False, True])
def test_smoother_is_blockwise_rts(parallel, return_full):
    M, N = 4, 19
    kern, dt, y, R = _problem(5, M, N)
    kt = kern.temporal_kernel
    _, (fm, fP) = OP.kalman_filter_meanfield(dt, kern, y, R, parallel=False)
    dts = np.concatenate([dt[1:], [0.0]])
    sm, sP, G = OP.rauch_tung_striebel_smoother_meanfield(dts, kern, fm, fP, return_full=return_full,
                                                          parallel=parallel)
    d = 2
    for i in range(M):
        sm_i, sP_i, g_i = OP.rauch_tung_striebel_smoother(dts, kt, fm[:, i], fP[:, i], return_full=True)
        np.testing.assert_allclose(G[:, i * d:(i + 1) * d, i * d:(i + 1) * d], g_i, rtol=1e-9, atol=1e-12)
        if return_full:
            np.testing.assert_allclose(sm[:, i * d:(i + 1) * d], sm_i, rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(sP[:, i * d:(i + 1) * d, i * d:(i + 1) * d], sP_i, rtol=1e-9, atol=1e-12)
        else:
            np.testing.assert_allclose(sm[:, i, 0], sm_i[:, 0, 0], rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(sP[:, i, i], sP_i[:, 0, 0], rtol=1e-9, atol=1e-12)
    if not return_full:
        off = sP - np.einsum("nii->ni", sP)[..., None] * np.eye(M)
        assert np.all(off == 0.0)


@pytest.mark.parametrize("inference", ["VI", "Newton"])
def test_decoupled_space_gives_dense_model_energy(inference):
    rng = np.random.default_rng(6)
    g, T_ = 3, 9
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    z = np.stack([a.ravel(), b.ravel()], 1)
    t = np.sort(rng.uniform(0, 8, T_))
    r = np.broadcast_to(z, (T_,) + z.shape).copy()
    y = rng.poisson(2.0, size=(T_, z.shape[0])).astype(np.float64)
    kt, ks = OK.Matern32(1.2, 2.5), OK.Matern32(0.7, 1e-3)
    kern = OK.SpatioTemporalKernel(kt, ks, z)
    dense = OM.MarkovGP(kern, OL.Poisson(), t[:, None], y, R=r, inference=inference)
    mf = OM.MarkovGP(kern, OL.Poisson(), t[:, None], y, R=r, inference=inference, mean_field=True)
    mfp = OM.MarkovGP(kern, OL.Poisson(), t[:, None], y, R=r, inference=inference, mean_field=True,
                      parallel=True)
    for _ in range(3):
        for o in (dense, mf, mfp):
            o.inference(lr=0.5)
    e = dense.energy()
    assert abs(mf.energy() - e) <= 1e-9 * abs(e)
    assert abs(mfp.energy() - e) <= 1e-9 * abs(e)
