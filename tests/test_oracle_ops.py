"""Pins the oracle's filter and smoother (bayesnewton/ops.py:154-380) to dense
Gaussian-process algebra, the pattern of the reference's own tests
(tests/test_vs_exact_marg_lik.py:41-65, tests/test_gp_vs_markovgp_reg.py:44-70)
at FP64-level tolerances:

* filter log-likelihood == log N(y | 0, K + R)
* smoother moments == dense GP posterior mean and marginal variance
* associative-scan forms == sequential forms
* the C restatement (oracle/kfrts.c) == the NumPy restatement
"""
import numpy as np
import pytest

from oracle import kernels as K, ops, models, cfast

KERNELS = {
    "m52": (lambda: K.Matern52(1.3, 0.8), ["matern52"], [0.8], [1.3]),
    "m32": (lambda: K.Matern32(0.7, 2.0), ["matern32"], [2.0], [0.7]),
    "m32+m52": (lambda: K.Sum([K.Matern32(0.5, 3.0), K.Matern52(1.0, 1.0)]), ["matern32", "matern52"],
                [3.0, 1.0], [0.5, 1.0]),
}


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(0, 30, n))
    y = np.sin(x) + 0.3 * rng.standard_normal(n)
    r = 0.05 + 0.2 * rng.uniform(size=n)
    return x, y, r, np.concatenate([[0.0], np.diff(x)])


@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("parallel", [False, True])
def test_filter_loglik_is_dense_marginal(name, parallel):
    x, y, r, dt = _data(150, 1)
    k = KERNELS[name][0]()
    ell, _ = ops.kalman_filter(dt, k, y[:, None, None], r[:, None, None], parallel=parallel)
    dense = models.DenseGP(k, x, y, r).log_marginal()
    assert abs(ell - dense) < 1e-9 * abs(dense)


@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("parallel", [False, True])
def test_smoother_is_dense_posterior(name, parallel):
    x, y, r, dt = _data(150, 2)
    k = KERNELS[name][0]()
    _, (fm, fP) = ops.kalman_filter(dt, k, y[:, None, None], r[:, None, None], parallel=parallel)
    dts = np.concatenate([dt[1:], [0.0]])
    sm, sP, _ = ops.rauch_tung_striebel_smoother(dts, k, fm, fP, parallel=parallel)
    mean, cov = models.DenseGP(k, x, y, r).posterior()
    np.testing.assert_allclose(sm[:, 0, 0], mean, rtol=1e-8, atol=1e-9)
    np.testing.assert_allclose(sP[:, 0, 0], np.diag(cov), rtol=1e-8, atol=1e-11)


def test_filtered_marginal_is_dense_prefix_posterior():
    """The filtered mean at step n is the dense posterior given y_0..n (last point)."""
    x, y, r, dt = _data(60, 3)
    k = KERNELS["m32+m52"][0]()
    _, (fm, fP) = ops.kalman_filter(dt, k, y[:, None, None], r[:, None, None])
    H = k.measurement_model()
    for n in [0, 7, 31, 59]:
        mean, cov = models.DenseGP(k, x[:n + 1], y[:n + 1], r[:n + 1]).posterior()
        assert abs((H @ fm[n])[0, 0] - mean[-1]) < 1e-9
        assert abs((H @ fP[n] @ H.T)[0, 0] - cov[-1, -1]) < 1e-9


def test_return_predict_is_one_step_prediction():
    x, y, r, dt = _data(80, 4)
    k = KERNELS["m52"][0]()
    _, (pm, pP) = ops.kalman_filter(dt, k, y[:, None, None], r[:, None, None], return_predict=True)
    H = k.measurement_model()
    for n in [1, 20, 79]:
        # prior predictive of f(x_n) given y_0..n-1
        mean, cov = models.DenseGP(k, np.append(x[:n], x[n]), np.append(y[:n], 0.0),
                                   np.append(r[:n], 1e12)).posterior()
        assert abs((H @ pm[n])[0, 0] - mean[-1]) < 1e-6
        assert abs((H @ pP[n] @ H.T)[0, 0] - cov[-1, -1]) < 1e-6


def test_masked_step_contributes_zero_loglik():
    """mvn_logpdf masking (utils.py:376-396): a masked step adds 0 to the
    log-likelihood but still updates the state with its own noise."""
    x, y, r, dt = _data(100, 5)
    k = KERNELS["m52"][0]()
    mask = np.zeros(100, dtype=bool)
    mask[[3, 50, 99]] = True
    ell_m, _ = ops.kalman_filter(dt, k, y[:, None, None], r[:, None, None], mask=mask[:, None, None])
    # the unmasked terms: loglik increments of the unmasked filter at those steps
    terms = []
    m, P = np.zeros((3, 1)), k.stationary_covariance()
    H = k.measurement_model()
    Pinf = k.stationary_covariance()
    for n in range(100):
        A = k.state_transition(dt[n])
        Q = Pinf - A @ Pinf @ A.T
        m, P = A @ m, A @ P @ A.T + Q
        S = (H @ P @ H.T)[0, 0] + r[n]
        v = y[n] - (H @ m)[0, 0]
        terms.append(-0.5 * (v * v / S + np.log(2 * np.pi * S)))
        Kg = P @ H.T / S
        m, P = m + Kg * v, P - Kg @ H @ P
    assert abs(ell_m - sum(t for t, mk in zip(terms, mask) if not mk)) < 1e-9


@pytest.mark.parametrize("name", list(KERNELS))
def test_c_restatement_matches_numpy(name):
    rng = np.random.default_rng(9)
    lens = [1, 2, 40, 131]
    off = np.concatenate([[0], np.cumsum(lens)])
    N = off[-1]
    dt = np.zeros(N)
    y = rng.standard_normal(N)
    r = 0.05 + 0.1 * rng.uniform(size=N)
    mask = rng.uniform(size=N) < 0.1
    for b in range(len(lens)):
        dt[off[b] + 1:off[b + 1]] = rng.uniform(0.01, 0.5, lens[b] - 1)
    mk, names, ls, var = KERNELS[name]
    pm, pv, ll = cfast.filter_smoother(names, ls, var, off, dt, y, r, mask)
    for b in range(len(lens)):
        s = slice(off[b], off[b + 1])
        k = mk()
        ell, (fm, fP) = ops.kalman_filter(dt[s], k, y[s, None, None], r[s, None, None], mask=mask[s, None, None])
        dts = np.concatenate([dt[s][1:], [0.0]])
        sm, sP, _ = ops.rauch_tung_striebel_smoother(dts, k, fm, fP)
        assert abs(ll[b] - ell) < 1e-10 * max(1.0, abs(ell))
        np.testing.assert_allclose(pm[s], sm[:, 0, 0], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(pv[s], sP[:, 0, 0], rtol=1e-9, atol=1e-13)


def test_c_restatement_per_series_hypers():
    rng = np.random.default_rng(10)
    lens = [30, 45, 12]
    off = np.concatenate([[0], np.cumsum(lens)])
    N = off[-1]
    dt = np.concatenate([np.concatenate([[0.0], rng.uniform(0.01, 0.3, n - 1)]) for n in lens])
    y = rng.standard_normal(N)
    r = np.full(N, 0.1)
    ls = np.array([[0.5, 2.0], [1.0, 1.0], [3.0, 0.7]])
    var = np.array([[1.0, 0.3], [0.5, 0.5], [2.0, 1.0]])
    pm, pv, ll = cfast.filter_smoother(["matern52", "matern32"], ls, var, off, dt, y, r)
    for b in range(3):
        s = slice(off[b], off[b + 1])
        k = K.Sum([K.Matern52(var[b, 0], ls[b, 0]), K.Matern32(var[b, 1], ls[b, 1])])
        x = np.cumsum(dt[s])
        mean, cov = models.DenseGP(k, x, y[s], r[s]).posterior()
        np.testing.assert_allclose(pm[s], mean, rtol=1e-8, atol=1e-10)
        np.testing.assert_allclose(pv[s], np.diag(cov), rtol=1e-8, atol=1e-12)
        assert abs(ll[b] - models.DenseGP(k, x, y[s], r[s]).log_marginal()) < 1e-9 * abs(ll[b])


def test_parallel_operator_is_associative():
    """The reference combine (ops.py:203-219) is associative on well-conditioned elements."""
    x, y, r, dt = _data(3, 6)
    k = KERNELS["m32+m52"][0]()
    Pinf, As, Qs = ops.discretise(k, np.array([0.0, 0.3, 0.4]))
    H = k.measurement_model()
    els = ops.make_associative_filtering_elements(As, Qs, H, y[:, None, None], r[:, None, None],
                                                  np.zeros((5, 1)), Pinf)
    e = [[a[i:i + 1] for a in els] for i in range(3)]
    left = ops.parallel_filtering_operator(ops.parallel_filtering_operator(e[0], e[1]), e[2])
    right = ops.parallel_filtering_operator(e[0], ops.parallel_filtering_operator(e[1], e[2]))
    for a, b in zip(left, right):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-11)
