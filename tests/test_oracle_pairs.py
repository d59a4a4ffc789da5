"""Pins the oracle's filter over pairs of states (ops.py:383-426, the filter of the
reference's SparseMarkovGaussianProcess) against dense Gaussian conditioning: the chain
s_0, s_1, ... with s_0 ~ N(0, Pinf) and s_{k+1} = A_k s_k + q_k, observed in pairs
y_k = (s_k, s_{k+1}) + e_k.  The log likelihood is log N(y | 0, H K H' + R) of the
stacked observations, and output k is the posterior of s_{k+1} given y_0..y_{k+1}."""
import numpy as np
import pytest
from scipy.stats import multivariate_normal

from oracle import kernels as OK, ops as OP


def _case(seed, N, kern):
    rng = np.random.default_rng(seed)
    d = kern.stationary_covariance().shape[0]
    dt = np.concatenate([[1e3], rng.uniform(0.2, 1.5, N - 1)])
    y = rng.standard_normal((N, 2 * d, 1))
    G = rng.standard_normal((N, 2 * d, 2 * d)) * 0.3
    R = G @ np.swapaxes(G, 1, 2) + 0.2 * np.eye(2 * d)
    return dt, y, R


@pytest.mark.parametrize("kern", [OK.Matern32(1.3, 2.0), OK.Matern12(0.8, 1.0),
                                  OK.Sum([OK.Matern52(1.0, 3.0), OK.Matern12(0.5, 0.7)])])
@pytest.mark.parametrize("parallel", [False, True])
def test_pairs_filter_is_dense_conditioning(kern, parallel):
    N = 7
    dt, y, R = _case(1, N, kern)
    ell, (fm, fP) = OP.kalman_filter_pairs(dt, kern, y, R, parallel=parallel)
    Pinf = kern.stationary_covariance()
    d = Pinf.shape[0]
    # joint covariance of s_0..s_N
    A = [kern.state_transition(t) for t in dt]
    Q = [Pinf - a @ Pinf @ a.T for a in A]
    n = N + 1
    K = np.zeros((n * d, n * d))
    M = [np.eye(d)]  # s_k = M_k s_0 + noise; build covariance recursively
    C = [Pinf]
    for k in range(N):
        C.append(A[k] @ C[k] @ A[k].T + Q[k])
    # Cov(s_i, s_j) for i <= j: Phi(j, i) C_i with Phi(j, i) = A_{j-1} ... A_i
    for i in range(n):
        Phi = np.eye(d)
        for j in range(i, n):
            if j > i:
                Phi = A[j - 1] @ Phi
            K[j * d:(j + 1) * d, i * d:(i + 1) * d] = Phi @ C[i]
            K[i * d:(i + 1) * d, j * d:(j + 1) * d] = (Phi @ C[i]).T
    Hs = np.zeros((N * 2 * d, n * d))
    for k in range(N):
        Hs[k * 2 * d:(k + 1) * 2 * d, k * d:(k + 2) * d] = np.eye(2 * d)
    Rs = np.zeros((N * 2 * d, N * 2 * d))
    for k in range(N):
        Rs[k * 2 * d:(k + 1) * 2 * d, k * 2 * d:(k + 1) * 2 * d] = R[k]
    yv = y.reshape(-1)
    ref = multivariate_normal(np.zeros(yv.size), Hs @ K @ Hs.T + Rs).logpdf(yv)
    assert abs(ell - ref) <= 1e-9 * abs(ref)
    for k in range(1, N):
        rows = slice(0, (k + 1) * 2 * d)
        Hk, Rk, yk = Hs[rows], Rs[rows, rows], yv[rows]
        Kk = K[:, k * d:(k + 1) * d]
        Sy = Hk @ K @ Hk.T + Rk
        gain = np.linalg.solve(Sy, Hk @ Kk).T
        mean = gain @ yk
        cov = K[k * d:(k + 1) * d, k * d:(k + 1) * d] - gain @ Hk @ Kk
        np.testing.assert_allclose(fm[k - 1, :, 0], mean, rtol=1e-8, atol=1e-10)
        np.testing.assert_allclose(fP[k - 1], cov, rtol=1e-8, atol=1e-10)
