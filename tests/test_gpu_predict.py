"""GPU parity of prediction (MarkovGaussianProcess.predict / predict_y /
negative_log_predictive_density, basemodels.py:165-193, 766-816) against the FP64
oracle, whose prediction path tests/test_oracle_predict.py pins against the dense GP
predictive and quadrature.  Test times include points before, between, on and after
the training times and repeats.  Tolerance: 1e-4 relative (the sites and pseudo
observations the GPU path filters are FP32 per step)."""
import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _test_times(x, rng, n=40):
    lo, hi = x.min(), x.max()
    return np.concatenate([[lo - 3.0, lo - 0.01], rng.uniform(lo, hi, n), x[[0, 3, len(x) - 1]], [x[5], x[5]],
                           [hi + 0.5, hi + 7.0]])


def _close(g, o, scale):
    g = np.asarray(g.cpu().numpy() if hasattr(g, "cpu") else g, np.float64)
    np.testing.assert_allclose(g, o, rtol=TOL, atol=TOL * scale)


@pytest.mark.parametrize("case", ["sum-studentt-vi", "m32-bernoulli-newton", "m12-poisson-ep", "m52-gauss-vi"])
def test_predict_matches_oracle(case):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(3)
    n = 300
    x = np.sort(rng.choice(40 * 256, size=n, replace=False) / 256.0)
    f = np.sin(0.4 * x) + 0.5 * np.cos(1.3 * x)
    if case == "sum-studentt-vi":
        gk = K.Sum([K.Matern52(variance=0.8, lengthscale=1.5), K.Matern32(variance=0.4, lengthscale=4.0)])
        ok = OK.Sum([OK.Matern52(0.8, 1.5), OK.Matern32(0.4, 4.0)])
        gl, ol = L.StudentsT(scale=0.25, df=3.0), OL.StudentsT(0.25, 3.0)
        y = _f32(f + 0.25 * rng.standard_t(3.0, size=n))
        cls, kind, kw = M.MarkovVariationalGP, "VI", {}
    elif case == "m32-bernoulli-newton":
        gk, ok = K.Matern32(variance=1.5, lengthscale=2.0), OK.Matern32(1.5, 2.0)
        gl, ol = L.Bernoulli(), OL.Bernoulli()
        y = (rng.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(np.float64)
        cls, kind, kw = M.MarkovLaplaceGP, "Newton", {}
    elif case == "m12-poisson-ep":
        gk, ok = K.Matern12(variance=0.7, lengthscale=3.0), OK.Matern12(0.7, 3.0)
        gl, ol = L.Poisson(binsize=1.5), OL.Poisson(1.5)
        y = rng.poisson(1.5 * np.exp(f)).astype(np.float64)
        cls, kind, kw = M.MarkovExpectationPropagationGP, "EP", {"power": 0.5}
    else:
        gk, ok = K.Matern52(variance=1.0, lengthscale=1.2), OK.Matern52(1.0, 1.2)
        gl, ol = L.Gaussian(variance=0.05), OL.Gaussian(0.05)
        y = _f32(f + 0.2 * rng.standard_normal(n))
        cls, kind, kw = M.MarkovVariationalGP, "VI", {}
    y[[10, 11, 200]] = np.nan
    g = cls(gk, gl, x, y, **kw)
    o = OM.MarkovGP(ok, ol, x, y, inference=kind, **kw)
    for lr in (1.0, 0.7):
        g.inference(lr=lr)
        o.inference(lr=lr)
    xs = _test_times(x, rng)
    gm, gv = g.predict(xs)
    om, ov = o.predict(xs[:, None])
    _close(gm, om, np.max(np.abs(om)))
    _close(gv, ov, np.max(ov))
    gmy, gvy = g.predict_y(xs)
    omy, ovy = o.predict_y(xs[:, None])
    _close(gmy, omy, np.max(np.abs(omy)))
    _close(gvy, ovy, np.max(ovy))
    ys = np.where(rng.uniform(size=xs.shape[0]) < 0.1, np.nan, np.interp(xs, x[~np.isnan(y)], y[~np.isnan(y)]))
    if case == "m32-bernoulli-newton":
        ys = np.where(np.isnan(ys), np.nan, (ys > 0.5).astype(float))
    elif case == "m12-poisson-ep":
        ys = np.where(np.isnan(ys), np.nan, np.round(ys))
    gn = g.negative_log_predictive_density(xs, ys)
    on = o.negative_log_predictive_density(xs[:, None], ys)
    assert abs(gn - on) <= TOL * max(1.0, abs(on)), (gn, on)
    # default X: the training times
    gm0, gv0 = g.predict()
    om0, ov0 = o.predict()
    _close(gm0, om0, np.max(np.abs(om0)))
    _close(gv0, ov0, np.max(ov0))


def test_predict_multi_series_matches_per_series_oracle():
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(8)
    lens = [5, 120, 37]
    xs = [np.sort(rng.choice(30 * 256, size=n, replace=False) / 256.0) for n in lens]
    ys = [_f32(np.sin(x) + 0.2 * rng.standard_t(4.0, size=len(x))) for x in xs]
    lik = L.StudentsT(scale=0.2, df=4.0)
    g = M.MarkovVariationalGP.from_series(K.Matern32(variance=1.0, lengthscale=2.0), lik, xs, ys, chunk_len=16)
    g.inference(lr=1.0)
    tests = [_test_times(x, rng, n=15) if len(x) > 5 else np.array([x[0] - 1, x[2], x[-1] + 2]) for x in xs]
    gm, gv = g.predict(tests)
    gm, gv = gm.cpu().numpy(), gv.cpu().numpy()
    off = np.concatenate([[0], np.cumsum([len(t) for t in tests])])
    for b, (x, y, t) in enumerate(zip(xs, ys, tests)):
        o = OM.MarkovGP(OK.Matern32(1.0, 2.0), OL.StudentsT(0.2, 4.0), x, y, inference="VI")
        o.inference(lr=1.0)
        om, ov = o.predict(t[:, None])
        np.testing.assert_allclose(gm[off[b]:off[b + 1]], om, rtol=TOL, atol=TOL * np.max(np.abs(om)))
        np.testing.assert_allclose(gv[off[b]:off[b + 1]], ov, rtol=TOL, atol=TOL * np.max(ov))


def test_predict_sorted_order_with_a_series_without_test_times():
    """bn_predict sorts the test times by the state slot they read; outputs come back in
    input order, for series with unsorted test times, a series with none, and test
    times far outside the data (the sentinel states)."""
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(9)
    lens = [40, 7, 90, 3]
    xs = [np.sort(rng.choice(20 * 256, size=n, replace=False) / 256.0) for n in lens]
    ys = [_f32(np.cos(0.7 * x) + 0.1 * rng.standard_normal(len(x))) for x in xs]
    g = M.MarkovVariationalGP.from_series(K.Matern52(variance=1.0, lengthscale=1.5), L.Gaussian(variance=0.01), xs,
                                          ys, chunk_len=8)
    g.inference(lr=1.0)
    tests = [rng.permutation(np.concatenate([rng.uniform(x[0] - 2, x[-1] + 2, 30), [x[-1] + 1e3, x[0] - 1e3]]))
             if b != 1 else np.zeros(0) for b, x in enumerate(xs)]
    gm, gv = g.predict(tests)
    gm, gv = gm.cpu().numpy(), gv.cpu().numpy()
    off = np.concatenate([[0], np.cumsum([len(t) for t in tests])])
    assert gm.shape[0] == off[-1]
    for b, (x, y, t) in enumerate(zip(xs, ys, tests)):
        if len(t) == 0:
            continue
        o = OM.MarkovGP(OK.Matern52(1.0, 1.5), OL.Gaussian(0.01), x, y, inference="VI")
        o.inference(lr=1.0)
        om, ov = o.predict(t[:, None])
        np.testing.assert_allclose(gm[off[b]:off[b + 1]], om, rtol=TOL, atol=TOL * np.max(np.abs(om)))
        np.testing.assert_allclose(gv[off[b]:off[b + 1]], ov, rtol=TOL, atol=TOL * np.max(ov))


@pytest.mark.parametrize("g,T,nr,kind", [(3, 20, 16, "VI"), (2, 15, 9, "Newton"), (8, 12, 256, "VI")])
def test_spacetime_predict_matches_oracle(g, T, nr, kind):
    """Space-time prediction (basemodels.py:766-816): filter, full smoother with the
    neighbours' cross covariances, Kronecker-structured state conditional and the spatial
    conditional of the test locations, against the oracle's dense restatement.  With 64
    inducing points the steps carry 256 locations: with as many locations as inducing points
    B' diag(nat2) B is numerically singular (condition ~1e17), and the reference's explicit
    inverse of it (basemodels.py:683), which the oracle restates, is then rounding noise."""
    from bayesnewton_amd import kernels as K, likelihoods as L, models as GM
    from oracle import kernels as OK, likelihoods as OL, models as OM
    rng = np.random.default_rng(40 + g)
    t = np.arange(T, dtype=np.float64)
    r = rng.uniform(size=(T, nr, 2)).astype(np.float32).astype(np.float64)
    f = np.sin(0.3 * t)[:, None] + np.cos(3.0 * r[..., 0]) * np.sin(2.0 * r[..., 1])
    y = rng.poisson(2.0 * np.exp(f)).astype(np.float64)
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    z = np.stack([a.ravel(), b.ravel()], 1)
    gk = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=5.0), K.Matern32(variance=1.0, lengthscale=0.3),
                                z=z, sparse=True, conditional="Full")
    ok = OK.SpatioTemporalKernel(OK.Matern32(1.0, 5.0), OK.Matern32(1.0, 0.3), z)
    cls = {"VI": GM.MarkovVariationalGP, "Newton": GM.MarkovLaplaceGP}[kind]
    gm = cls(gk, L.Poisson(binsize=2.0), t, y, R=r)
    om = OM.MarkovGP(ok, OL.Poisson(2.0), t[:, None], y, R=r, inference=kind)
    for _ in range(2):
        gm.inference(lr=1.0)
        om.inference(lr=1.0)
    ts = np.array([-2.0, 0.0, 3.0, 4.25, 7.5, T - 1.0, T + 1.5])
    rs = rng.uniform(size=(ts.shape[0], 5, 2)).astype(np.float32).astype(np.float64)
    print(f"[st predict] g={g} T={T} R={nr} kind={kind}", flush=True)
    mu, var = gm.predict(ts, rs)
    omu, ovar = om.predict(ts[:, None], rs)
    mu, var = mu.cpu().numpy(), var.cpu().numpy()
    assert mu.shape == omu.shape and var.shape == ovar.shape
    assert np.max(np.abs(mu - omu)) <= 1e-4 * np.max(np.abs(omu))
    assert np.max(np.abs(var - ovar) / ovar) <= 1e-4
    # default test inputs: the training times and locations
    mu0, var0 = gm.predict()
    omu0, ovar0 = om.predict(t[:, None], r)
    assert np.max(np.abs(mu0.cpu().numpy() - omu0)) <= 1e-4 * np.max(np.abs(omu0))
    assert np.max(np.abs(var0.cpu().numpy() - ovar0) / ovar0) <= 1e-4


@pytest.mark.parametrize("variant,g,T,nr", [("mf-par", 3, 18, 16), ("mf-seq", 3, 18, 16), ("ih", 3, 18, 16),
                                            ("ih", 8, 10, 256)])
def test_spacetime_meanfield_and_infinite_horizon_predict_match_oracle(variant, g, T, nr):
    """Prediction of the mean-field (basemodels.py:1177-1228) and infinite-horizon
    (temporal_conditional_infinite_horizon, utils.py:142-177) space-time models against the
    oracle; the same calls on both sides (the infinite-horizon Riccati solutions advance
    with each filter / smoother call)."""
    from bayesnewton_amd import kernels as K, likelihoods as L, models as GM
    from oracle import kernels as OK, likelihoods as OL, models as OM
    rng = np.random.default_rng(60 + g)
    t = np.arange(T, dtype=np.float64)
    r = rng.uniform(size=(T, nr, 2)).astype(np.float32).astype(np.float64)
    f = np.sin(0.3 * t)[:, None] + np.cos(3.0 * r[..., 0]) * np.sin(2.0 * r[..., 1])
    y = rng.poisson(2.0 * np.exp(f)).astype(np.float64)
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    z = np.stack([a.ravel(), b.ravel()], 1)
    gk = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=5.0), K.Matern32(variance=1.0, lengthscale=0.3),
                                z=z, sparse=True, conditional="Full")
    ok = OK.SpatioTemporalKernel(OK.Matern32(1.0, 5.0), OK.Matern32(1.0, 0.3), z)
    if variant == "ih":
        gm = GM.InfiniteHorizonVariationalGP(gk, L.Poisson(binsize=2.0), t, y, R=r)
        om = OM.MarkovGP(ok, OL.Poisson(2.0), t[:, None], y, R=r, inference="VI", infinite_horizon=True)
    else:
        par = variant == "mf-par"
        gm = GM.MarkovVariationalMeanFieldGP(gk, L.Poisson(binsize=2.0), t, y, R=r, parallel=par)
        om = OM.MarkovGP(ok, OL.Poisson(2.0), t[:, None], y, R=r, inference="VI", mean_field=True, parallel=par)
    for _ in range(2):
        gm.inference(lr=1.0)
        om.inference(lr=1.0)
    ts = np.array([-2.0, 0.0, 3.0, 4.25, 7.5, T - 1.0, T + 1.5])
    rs = rng.uniform(size=(ts.shape[0], 5, 2)).astype(np.float32).astype(np.float64)
    print(f"[st predict] variant={variant} M={g * g} T={T} R={nr}", flush=True)
    for _ in range(2):
        mu, var = gm.predict(ts, rs)
        omu, ovar = om.predict(ts[:, None], rs)
        mu, var = mu.cpu().numpy(), var.cpu().numpy()
        assert mu.shape == omu.shape and var.shape == ovar.shape
        assert np.max(np.abs(mu - omu)) <= 1e-4 * np.max(np.abs(omu))
        assert np.max(np.abs(var - ovar) / ovar) <= 1e-4
