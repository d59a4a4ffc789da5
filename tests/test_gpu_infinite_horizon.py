"""GPU parity of the infinite-horizon space-time filter / smoother (ops.py:796-1068) and
of InfiniteHorizonVariationalGP / InfiniteHorizonLaplaceGP with a SpatioTemporalKernel
(basemodels.py:1257-1301) against the FP64 oracle.  FP64 throughout: states within 1e-8
relative and log likelihoods within 1e-9 on the ops; the model tests use the space-time
tolerances (1e-4 on moments, 1e-5 on the energy) over a mirrored call sequence (the
Riccati solutions advance with every filter / smoother call, on both sides).
"""
import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM, ops as OP

pytestmark = pytest.mark.gpu

STATE_TOL, LL_TOL = 1e-8, 1e-9
MEAN_TOL, VAR_TOL, E_TOL = 1e-4, 1e-4, 1e-5


def _kernels(temporal, M, seed=0):
    from bayesnewton_amd import kernels as K
    rng = np.random.default_rng(seed)
    z = rng.uniform(size=(M, 2))
    gcls, ocls = {"m12": (K.Matern12, OK.Matern12), "m32": (K.Matern32, OK.Matern32)}[temporal]
    gk = K.SpatioTemporalKernel(gcls(variance=1.3, lengthscale=4.0), K.Matern32(variance=1.0, lengthscale=0.4),
                                z=z, sparse=True, conditional="Full")
    ok = OK.SpatioTemporalKernel(ocls(1.3, 4.0), OK.Matern32(1.0, 0.4), z)
    return gk, ok


def _problem(seed, M, N):
    rng = np.random.default_rng(seed)
    dt = np.r_[0.0, np.full(N - 1, 0.7)]
    y = rng.standard_normal((N, M, 1))
    G = rng.standard_normal((N, M, M)) * (0.8 / np.sqrt(M))
    R = G @ np.swapaxes(G, -1, -2) + np.eye(M) * rng.uniform(0.05, 0.5, (N, 1, 1))
    Rt = np.linalg.inv(np.mean(np.linalg.inv(R), axis=0))
    return dt, y, R, Rt


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("temporal,M,N,hetero,iters,warm", [
    ("m32", 5, 30, True, 20, False), ("m32", 5, 30, False, 7, True), ("m12", 9, 25, True, 20, True),
    ("m32", 64, 24, True, 20, False), ("m32", 64, 24, False, 20, True), ("m32", 1, 20, True, 5, False),
])
def test_filter_and_smoother_match_oracle(temporal, M, N, hetero, iters, warm):
    from bayesnewton_amd import ops
    gk, ok = _kernels(temporal, M, seed=M)
    dt, y, R, Rt = _problem(M + N, M, N)
    D = M * (1 if temporal == "m12" else 2)
    init = None
    if warm:
        W = np.random.default_rng(5).standard_normal((D, D)) * 0.1
        init = ok.stationary_covariance() + W @ W.T
    print(f"[ih filter] temporal={temporal} M={M} N={N} hetero={hetero} iters={iters} warm={warm}", flush=True)
    kw = dict(heteroscedastic=hetero, noise_cov_tied=Rt, dare_iters=iters, dare_init=init)
    ell, (fm, (Pd, cov)) = ops.kalman_filter_infinite_horizon(dt, gk, y, R, **kw)
    oell, (ofm, (oPd, ocov)) = OP.kalman_filter_infinite_horizon(dt, ok, y, R, **kw)
    assert abs(float(ell) - oell) <= LL_TOL * abs(oell), (float(ell), oell)
    assert _rel(fm.cpu().numpy(), ofm) < STATE_TOL
    assert _rel(Pd.cpu().numpy(), oPd) < STATE_TOL
    assert _rel(cov.cpu().numpy(), ocov) < STATE_TOL
    dts = np.r_[dt[1:], 0.0]
    for full in (False, True):
        sm, sP, G, dc = ops.rauch_tung_striebel_smoother_infinite_horizon(dts, gk, ofm, (oPd, ocov),
                                                                          return_full=full, dare_iters=iters,
                                                                          dare_init=init)
        osm, osP, oG, odc = OP.rauch_tung_striebel_smoother_infinite_horizon(dts, ok, ofm, (oPd, ocov),
                                                                             return_full=full, dare_iters=iters,
                                                                             dare_init=init)
        assert sm.shape == osm.shape and sP.shape == osP.shape and G.shape == oG.shape
        assert _rel(sm.cpu().numpy(), osm) < STATE_TOL
        assert _rel(sP.cpu().numpy(), osP) < STATE_TOL
        assert _rel(G.cpu().numpy(), oG) < STATE_TOL
        assert _rel(dc.cpu().numpy(), odc) < STATE_TOL


def _grid(g):
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    return np.stack([a.ravel(), b.ravel()], 1)


def _st_data(seed, T, nr, lik):
    rng = np.random.default_rng(seed)
    t = np.arange(T, dtype=np.float64)
    r = rng.uniform(size=(T, nr, 2)).astype(np.float32).astype(np.float64)
    f = np.sin(0.3 * t)[:, None] + np.cos(3.0 * r[..., 0]) * np.sin(2.0 * r[..., 1])
    if lik == "poisson":
        y = rng.poisson(2.0 * np.exp(f)).astype(np.float64)
    else:
        y = (f + 0.3 * rng.standard_normal(f.shape)).astype(np.float32).astype(np.float64)
    return t, r, y


@pytest.mark.parametrize("kind,lik,g,T,nr", [
    ("VI", "poisson", 3, 30, 20), ("VI", "gauss", 3, 25, 16), ("Newton", "poisson", 2, 25, 16),
    ("VI", "poisson", 8, 16, 256),
])
def test_infinite_horizon_model_matches_oracle(kind, lik, g, T, nr):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as GM
    t, r, y = _st_data(31 + g, T, nr, lik)
    z = _grid(g)
    gk = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=5.0), K.Matern32(variance=1.0, lengthscale=0.3),
                                z=z, sparse=True, conditional="Full")
    ok = OK.SpatioTemporalKernel(OK.Matern32(1.0, 5.0), OK.Matern32(1.0, 0.3), z)
    gl, ol = {"poisson": (L.Poisson(binsize=2.0), OL.Poisson(2.0)),
              "gauss": (L.Gaussian(variance=0.09), OL.Gaussian(0.09))}[lik]
    cls = {"VI": GM.InfiniteHorizonVariationalGP, "Newton": GM.InfiniteHorizonLaplaceGP}[kind]
    gm = cls(gk, gl, t, y, R=r)
    om = OM.MarkovGP(ok, ol, t[:, None], y, R=r, inference=kind, infinite_horizon=True)
    assert gm.heteroscedastic == om.heteroscedastic
    print(f"[ih model] kind={kind} lik={lik} M={g * g} T={T} R={nr}", flush=True)
    for it in range(3):
        lr = 1.0 if it == 0 else 0.7
        gm.inference(lr=lr)
        om.inference(lr=lr)
        pm = gm.posterior_mean.cpu().numpy()
        pP = gm.posterior_variance.cpu().numpy()
        assert np.max(np.abs(pm - om.posterior_mean)) / np.max(np.abs(om.posterior_mean)) < MEAN_TOL
        oP = om.posterior_variance
        assert np.max(np.abs(pP - oP)) / np.max(np.abs(np.diagonal(oP, axis1=1, axis2=2))) < VAR_TOL
        gll, oll = float(gm.compute_log_lik()), om.compute_log_lik()
        assert abs(gll - oll) <= E_TOL * abs(oll), (it, gll, oll)
        ge, oe = float(gm.energy()), om.energy()
        assert abs(ge - oe) <= E_TOL * abs(oe), (it, ge, oe)
    Xf, Xs = gm.dare_solutions
    assert _rel(Xf.cpu().numpy(), om.dare_init_filter) < 1e-6
    assert _rel(Xs.cpu().numpy(), om.dare_init_smoother) < 1e-6
