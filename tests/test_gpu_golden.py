"""GPU path against the committed golden fixtures (tests/golden/), without running
the oracle: ops.kalman_filter / ops.rauch_tung_striebel_smoother (ops.py:256-285,
357-380) and three MarkovVariationalGP inference steps (inference.py:65-90) with a
Bernoulli-probit likelihood, and three VI steps of the dense, mean-field and
infinite-horizon space-time models with a Poisson likelihood, and the filter over pairs
of states (ops.py:383-426, FP64 throughout: 1e-9 on the log likelihood, 1e-8 on the
moments), and prediction / predict_y / NLPD (basemodels.py:165-193, 766-816) after two
Student-t VI steps.  Tolerances as in the other GPU parity tests: 1e-4
relative on means (to the series' scale) and variances, 1e-5 on log-likelihoods and
energies."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MEAN_TOL, VAR_TOL, LL_TOL = 1e-4, 1e-4, 1e-5


def _load(name):
    z = np.load(os.path.join(GOLDEN, name))
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("parallel", [False, True])
def test_kf_rts_matches_golden(parallel):
    from bayesnewton_amd import kernels as K, ops
    z = _load("kf_rts_sum_m32_m52_n400.npz")
    kern = K.Sum([K.Matern32(variance=0.5, lengthscale=3.0), K.Matern52(variance=1.0, lengthscale=1.0)])
    dt, y, r, mask = z["dt"], z["y"], z["r"], z["mask"]
    ell, (fm, fP) = ops.kalman_filter(dt, kern, y[:, None, None], r[:, None, None], mask=mask[:, None, None],
                                      parallel=parallel)
    assert abs(float(ell) - float(z["loglik"])) < LL_TOL * abs(float(z["loglik"]))
    ofm, ofP = z["filter_mean"], z["filter_cov"]
    assert np.max(np.abs(fm.cpu().numpy() - ofm)) / np.max(np.abs(ofm)) < MEAN_TOL
    assert np.max(np.abs(fP.cpu().numpy() - ofP)) / np.max(np.abs(ofP)) < VAR_TOL
    dts = np.concatenate([dt[1:], [0.0]])
    sm, sP, _ = ops.rauch_tung_striebel_smoother(dts, kern, ofm, ofP, parallel=parallel)
    osm, osv = z["smooth_mean"].reshape(-1), z["smooth_var"].reshape(-1)
    assert np.max(np.abs(sm.cpu().numpy().reshape(-1) - osm)) / np.max(np.abs(osm)) < MEAN_TOL
    assert np.max(np.abs(sP.cpu().numpy().reshape(-1) - osv) / osv) < VAR_TOL


def test_vi_probit_matches_golden():
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    z = _load("vi_probit_n300.npz")
    kern = K.Sum([K.Matern32(variance=0.5, lengthscale=10.0), K.Matern52(variance=0.5, lengthscale=10.0)])
    g = M.MarkovVariationalGP(kern, L.Bernoulli(), z["x"], z["y"], parallel=True)
    for it, lr in enumerate((1.0, 0.7, 0.7)):
        g.inference(lr=lr)
        pm = g.posterior_mean.cpu().numpy().reshape(-1)
        pv = g.posterior_variance.cpu().numpy().reshape(-1)
        om, ov = z["post_mean"][it], z["post_var"][it]
        assert np.max(np.abs(pm - om)) / np.max(np.abs(om)) < MEAN_TOL, it
        assert np.max(np.abs(pv - ov) / ov) < VAR_TOL, it
        e = float(g.energy())
        assert abs(e - z["energy"][it]) <= LL_TOL * abs(z["energy"][it]), (it, e, z["energy"][it])


@pytest.mark.parametrize("variant", ["dense", "meanfield", "ih"])
def test_spacetime_vi_matches_golden(variant):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    z = _load("spacetime_poisson_vi_g3_t24.npz")
    gk = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=5.0), K.Matern32(variance=1.0, lengthscale=0.3),
                                z=z["z"], sparse=True, conditional="Full")
    lik = L.Poisson(binsize=2.0)
    if variant == "dense":
        g = M.MarkovVariationalGP(gk, lik, z["t"], z["y"], R=z["r"])
    elif variant == "meanfield":
        g = M.MarkovVariationalMeanFieldGP(gk, lik, z["t"], z["y"], R=z["r"], parallel=True)
    else:
        g = M.InfiniteHorizonVariationalGP(gk, lik, z["t"], z["y"], R=z["r"])
    print(f"[golden spacetime] variant={variant}", flush=True)
    for it, lr in enumerate((1.0, 0.7, 0.7)):
        g.inference(lr=lr)
        pm = g.posterior_mean.cpu().numpy()
        pv = np.diagonal(g.posterior_variance.cpu().numpy(), axis1=1, axis2=2)
        om, ov = z[f"{variant}_post_mean"][it], z[f"{variant}_post_var"][it]
        assert np.max(np.abs(pm - om)) / np.max(np.abs(om)) < MEAN_TOL, it
        assert np.max(np.abs(pv - ov) / ov) < VAR_TOL, it
        e, oe = float(g.energy()), z[f"{variant}_energy"][it]
        assert abs(e - oe) <= LL_TOL * abs(oe), (it, e, oe)


def test_pairs_filter_matches_golden():
    """The sequential pairs filter (ops.kalman_filter_pairs runs one kernel for both
    of the reference's forms; the golden values come from the oracle's sequential
    form, which its own tests show equal to the associative one)."""
    from bayesnewton_amd import kernels as K, ops
    z = _load("pairs_sum_m52_m32_n60.npz")
    kern = K.Sum([K.Matern52(variance=1.0, lengthscale=3.0), K.Matern32(variance=0.5, lengthscale=0.7)])
    ell, (fm, fP) = ops.kalman_filter_pairs(z["dt"], kern, z["y"], z["noise"])
    oll = float(z["loglik"])
    assert abs(float(ell) - oll) <= 1e-9 * abs(oll)
    fm, fP = fm.cpu().numpy(), fP.cpu().numpy()
    assert fm.shape == z["pair_mean"].shape and fP.shape == z["pair_cov"].shape
    assert np.max(np.abs(fm - z["pair_mean"])) <= 1e-8 * np.max(np.abs(z["pair_mean"]))
    assert np.max(np.abs(fP - z["pair_cov"])) <= 1e-8 * np.max(np.abs(z["pair_cov"]))


def test_predict_matches_golden():
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    z = _load("predict_studentt_vi_n300.npz")
    kern = K.Sum([K.Matern52(variance=0.8, lengthscale=1.5), K.Matern32(variance=0.4, lengthscale=4.0)])
    g = M.MarkovVariationalGP(kern, L.StudentsT(scale=0.25, df=3.0), z["x"], z["y"])
    for lr in (1.0, 0.7):
        g.inference(lr=lr)

    def close(a, key):
        a = np.asarray(a.cpu().numpy() if hasattr(a, "cpu") else a, np.float64).reshape(-1)
        o = z[key]
        np.testing.assert_allclose(a, o, rtol=MEAN_TOL, atol=MEAN_TOL * np.max(np.abs(o)), err_msg=key)

    m, v = g.predict(z["x_test"])
    close(m, "pred_mean")
    close(v, "pred_var")
    my, vy = g.predict_y(z["x_test"])
    close(my, "pred_y_mean")
    close(vy, "pred_y_var")
    nlpd = float(g.negative_log_predictive_density(z["x_test"], z["y_test"]))
    assert abs(nlpd - float(z["nlpd"])) <= MEAN_TOL * max(1.0, abs(float(z["nlpd"])))
