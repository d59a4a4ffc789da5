"""Pins the oracle's sparse Markov GP (oracle/sparse.py, basemodels.py:928-1152).

The reference's update_posterior smooths the pairs filter's first-state marginals as if
they were filtered marginals, and the data conditional carries the 1e-8 jitter of
compute_conditional_statistics, so the model is an approximation of the GP it induces;
its own tests (tests/test_sparsemarkov.py:49-276) compare it with the full GP at rtol
1e-2 for Z = X.  Pinned here:

* exactly: with a Gaussian likelihood one VI step at lr = 1 makes every transition's
  site sum_n W_n' y_n / s2 and sum_n W_n' W_n / s2 (+ 1e-8 I) over its data points,
  whatever the posterior was (a dropped Hessian-times-mean term, a transposed W or a
  wrong grouping all leave a posterior-dependent residue); transitions without data
  keep their site; Newton gives the same sites as VI for a Gaussian likelihood;
* at the reference's tolerance: Z = X gives the full Markov GP's initial energy and
  posterior, and after one step its posterior, as tests/test_sparsemarkov.py checks;
* the prediction at the data times equals the data conditional's mean / variance;
* the pieces (pairs filter, RTS smoother, conditional statistics) are pinned by
  tests/test_oracle_pairs.py, tests/test_oracle_ops.py and tests/test_oracle_predict.py.
"""
import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM, sparse as OS
from oracle.models import compute_conditional_statistics


def _data(seed, N):
    rng = np.random.default_rng(seed)
    x = np.sort(np.linspace(-10.0, 30.0, N) + 0.5 * rng.standard_normal(N))
    y = np.cos(0.04 * x + 0.33 * np.pi) * np.sin(0.2 * x) + np.sqrt(0.15) * rng.standard_normal(N)
    return x, y


def _W(model):
    P, _ = compute_conditional_statistics(model.X[:, :1], model.Z, model.kernel, model.ind)
    return model.kernel.measurement_model() @ P  # [N, 1, 2d]


@pytest.mark.parametrize("kern", [OK.Matern52(1.5, 4.5), OK.Sum([OK.Matern32(0.7, 2.0), OK.Matern12(0.4, 6.0)])])
@pytest.mark.parametrize("zcase", ["x", "coarse", "gap"])
def test_gaussian_vi_step_gives_projected_exact_sites(kern, zcase):
    x, y = _data(1, 60)
    if zcase == "x":
        z = x
    elif zcase == "coarse":
        z = np.linspace(-12.0, 32.0, 17)
    else:  # a stretch of inducing points with no data between them
        z = np.concatenate([np.linspace(-12.0, 5.0, 12), np.linspace(40.0, 45.0, 4)])
    s2 = 0.3
    m = OS.SparseMarkovGP(kern, OL.Gaussian(s2), x[:, None], y[:, None], z)
    # move the posterior away from its initial value first
    m.inference(lr=0.5)
    old1, old2 = m.nat1.copy(), m.nat2.copy()
    m.inference(lr=1.0)
    W = _W(m)
    d2 = W.shape[-1]
    n1 = np.zeros((m.num_transitions, d2, 1))
    n2 = np.zeros((m.num_transitions, d2, d2))
    has = np.zeros(m.num_transitions, bool)
    for n, i in enumerate(m.ind):
        n1[i] += W[n].T * y[n] / s2
        n2[i] += W[n].T @ W[n] / s2
        has[i] = True
    n1[~has] = old1[~has]
    n2[~has] = old2[~has]
    n2 += 1e-8 * np.eye(d2)
    np.testing.assert_allclose(m.nat1, n1, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(m.nat2, n2, rtol=1e-10, atol=1e-12)
    if zcase == "gap":
        assert not has.all()


def test_newton_equals_vi_for_gaussian_likelihood():
    x, y = _data(2, 40)
    z = np.linspace(-12.0, 32.0, 25)
    a = OS.SparseMarkovGP(OK.Matern32(1.0, 3.0), OL.Gaussian(0.2), x[:, None], y[:, None], z, inference="VI")
    b = OS.SparseMarkovGP(OK.Matern32(1.0, 3.0), OL.Gaussian(0.2), x[:, None], y[:, None], z, inference="Newton")
    for lr in (0.6, 1.0):
        a.inference(lr=lr)
        b.inference(lr=lr)
    np.testing.assert_allclose(a.nat1, b.nat1, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a.nat2, b.nat2, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("var_f,len_f,var_y", [(0.5, 4.5, 0.1), (1.5, 7.5, 0.5), (1.5, 4.5, 0.5)])
def test_z_equal_x_matches_markov_gp_at_reference_tolerance(var_f, len_f, var_y):
    """tests/test_sparsemarkov.py:49-71 (test_initial_loss) and 113-131 (test_inference_step)."""
    x, y = _data(3, 100)
    g = OM.MarkovGP(OK.Matern52(var_f, len_f), OL.Gaussian(var_y), x[:, None], y[:, None], inference="VI")
    s = OS.SparseMarkovGP(OK.Matern52(var_f, len_f), OL.Gaussian(var_y), x[:, None], y[:, None], x)
    g.update_posterior()
    s.update_posterior()
    np.testing.assert_allclose(g.posterior_mean[:, 0, 0], s.posterior_mean[1:, 0, 0], rtol=1e-2, atol=1e-12)
    np.testing.assert_allclose(g.posterior_variance[:, 0, 0], s.posterior_variance[1:, 0, 0], rtol=1e-2)
    assert abs(g.energy() - s.energy()) < 0.15  # decimal=1
    g.inference(lr=1.0)
    s.inference(lr=1.0)
    np.testing.assert_allclose(g.posterior_mean[:, 0, 0], s.posterior_mean[1:, 0, 0], rtol=1e-1, atol=2e-2)
    np.testing.assert_allclose(g.posterior_variance[:, 0, 0], s.posterior_variance[1:, 0, 0], rtol=1e-1)
    assert abs(g.energy() - s.energy()) <= 1e-3 * abs(g.energy())


def test_predict_at_data_times_is_the_data_conditional():
    x, y = _data(4, 50)
    z = np.linspace(-12.0, 32.0, 21)
    s = OS.SparseMarkovGP(OK.Matern52(1.0, 4.0), OL.Bernoulli(), x[:, None], (y > 0).astype(float)[:, None], z)
    for lr in (1.0, 0.5):
        s.inference(lr=lr)
    mf, vf = s.conditional_posterior_to_data()
    pm, pv = s.predict(x)
    np.testing.assert_allclose(pm, mf.reshape(-1), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(pv, vf.reshape(-1), rtol=1e-9, atol=1e-12)
    assert np.isfinite(s.energy())
