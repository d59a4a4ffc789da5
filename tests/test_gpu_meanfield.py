"""GPU parity of the mean-field space-time filter / smoother (ops.py:429-706) and of the
mean-field Markov models (MarkovMeanFieldGaussianProcess, basemodels.py:1158-1175)
against the FP64 oracle.  The GPU path is FP64 throughout: filtered and smoothed blocks
within 1e-8 relative, log likelihoods and energies within 1e-9 relative on the ops; the
model tests use the space-time tolerances (1e-4 on moments, 1e-5 on the energy).
"""
import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM, ops as OP

pytestmark = pytest.mark.gpu

STATE_TOL, LL_TOL = 1e-8, 1e-9
MEAN_TOL, VAR_TOL, E_TOL = 1e-4, 1e-4, 1e-5


def _kernels(temporal, M, seed=0, var_t=1.3, ls_t=2.0):
    from bayesnewton_amd import kernels as K
    rng = np.random.default_rng(seed)
    z = rng.uniform(size=(M, 2))
    gcls, ocls = {"m12": (K.Matern12, OK.Matern12), "m32": (K.Matern32, OK.Matern32)}[temporal]
    gk = K.SpatioTemporalKernel(gcls(variance=var_t, lengthscale=ls_t), K.Matern32(variance=1.0, lengthscale=0.4),
                                z=z, sparse=True, conditional="Full")
    ok = OK.SpatioTemporalKernel(ocls(var_t, ls_t), OK.Matern32(1.0, 0.4), z)
    return gk, ok


def _problem(seed, M, N, coupled=True):
    rng = np.random.default_rng(seed)
    dt = np.concatenate([[0.0], rng.uniform(0.1, 1.5, N - 1)])
    y = rng.standard_normal((N, M, 1))
    if coupled:
        G = rng.standard_normal((N, M, M)) * (0.8 / np.sqrt(M))
        R = G @ np.swapaxes(G, -1, -2) + np.eye(M) * rng.uniform(0.05, 0.5, (N, 1, 1))
    else:
        R = np.zeros((N, M, M))
        R[:, np.arange(M), np.arange(M)] = rng.uniform(0.05, 0.8, (N, M))
    return dt, y, R


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("parallel", [False, True])
@pytest.mark.parametrize("temporal,M,N,coupled", [
    ("m32", 5, 30, True), ("m12", 7, 25, True), ("m32", 64, 40, True), ("m32", 16, 33, False),
    ("m32", 1, 20, True), ("m12", 33, 9, True),
])
def test_filter_matches_oracle(parallel, temporal, M, N, coupled):
    from bayesnewton_amd import ops
    gk, ok = _kernels(temporal, M, seed=M)
    dt, y, R = _problem(M + N, M, N, coupled)
    print(f"[mf filter] parallel={parallel} temporal={temporal} M={M} N={N}", flush=True)
    ell, (fm, fP) = ops.kalman_filter_meanfield(dt, gk, y, R, parallel=parallel)
    oell, (ofm, ofP) = OP.kalman_filter_meanfield(dt, ok, y, R, parallel=parallel)
    assert abs(float(ell) - oell) <= LL_TOL * abs(oell), (float(ell), oell)
    assert _rel(fm.cpu().numpy(), ofm) < STATE_TOL
    assert _rel(fP.cpu().numpy(), ofP) < STATE_TOL


@pytest.mark.parametrize("parallel", [False, True])
@pytest.mark.parametrize("return_full", [False, True])
def test_smoother_matches_oracle(parallel, return_full):
    from bayesnewton_amd import ops
    M, N = 12, 37
    gk, ok = _kernels("m32", M, seed=3)
    dt, y, R = _problem(9, M, N)
    _, (ofm, ofP) = OP.kalman_filter_meanfield(dt, ok, y, R, parallel=parallel)
    dts = np.concatenate([dt[1:], [0.0]])
    print(f"[mf smoother] parallel={parallel} return_full={return_full}", flush=True)
    sm, sP, G = ops.rauch_tung_striebel_smoother_meanfield(dts, gk, ofm, ofP, return_full=return_full,
                                                           parallel=parallel)
    osm, osP, oG = OP.rauch_tung_striebel_smoother_meanfield(dts, ok, ofm, ofP, return_full=return_full,
                                                             parallel=parallel)
    assert sm.shape == osm.shape and sP.shape == osP.shape and G.shape == oG.shape
    assert _rel(sm.cpu().numpy(), osm) < STATE_TOL
    assert _rel(sP.cpu().numpy(), osP) < STATE_TOL
    assert _rel(G.cpu().numpy(), oG) < STATE_TOL


def _st_data(seed, T, nr):
    rng = np.random.default_rng(seed)
    t = np.arange(T, dtype=np.float64)
    r = rng.uniform(size=(T, nr, 2)).astype(np.float32).astype(np.float64)
    f = np.sin(0.3 * t)[:, None] + np.cos(3.0 * r[..., 0]) * np.sin(2.0 * r[..., 1])
    y = rng.poisson(2.0 * np.exp(f)).astype(np.float64)
    return t, r, y


def _grid(g):
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    return np.stack([a.ravel(), b.ravel()], 1)


@pytest.mark.parametrize("kind,parallel,g,T,nr", [
    ("VI", True, 3, 30, 20), ("VI", False, 3, 30, 20), ("Newton", True, 3, 25, 16),
    ("Newton", False, 2, 25, 16), ("VI", True, 8, 16, 256), ("VI", False, 8, 16, 256),
])
def test_meanfield_model_matches_oracle(kind, parallel, g, T, nr):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as GM
    t, r, y = _st_data(11 + g, T, nr)
    z = _grid(g)
    gk = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=5.0), K.Matern32(variance=1.0, lengthscale=0.3),
                                z=z, sparse=True, conditional="Full")
    ok = OK.SpatioTemporalKernel(OK.Matern32(1.0, 5.0), OK.Matern32(1.0, 0.3), z)
    cls = {"VI": GM.MarkovVariationalMeanFieldGP, "Newton": GM.MarkovLaplaceGPMeanField}[kind]
    gm = cls(gk, L.Poisson(binsize=2.0), t, y, R=r, parallel=parallel)
    om = OM.MarkovGP(ok, OL.Poisson(2.0), t[:, None], y, R=r, inference=kind, mean_field=True, parallel=parallel)
    print(f"[mf model] kind={kind} parallel={parallel} M={g * g} T={T} R={nr}", flush=True)
    for it in range(3):
        lr = 1.0 if it == 0 else 0.7
        gm.inference(lr=lr)
        om.inference(lr=lr)
        pm = gm.posterior_mean.cpu().numpy()
        pP = gm.posterior_variance.cpu().numpy()
        assert np.max(np.abs(pm - om.posterior_mean)) / np.max(np.abs(om.posterior_mean)) < MEAN_TOL
        ov = np.diagonal(om.posterior_variance, axis1=1, axis2=2)
        assert np.max(np.abs(np.diagonal(pP, axis1=1, axis2=2) - ov) / ov) < VAR_TOL
        off = pP - np.einsum("tii->ti", pP)[..., None] * np.eye(g * g)
        assert np.all(off == 0.0)
        gll, oll = float(gm.compute_log_lik()), om.compute_log_lik()
        assert abs(gll - oll) <= E_TOL * abs(oll), (it, gll, oll)
        ge, oe = float(gm.energy()), om.energy()
        assert abs(ge - oe) <= E_TOL * abs(oe), (it, ge, oe)


def test_meanfield_runs_match_single_models():
    """from_runs with the mean-field filter == one model per run."""
    from bayesnewton_amd import kernels as K, likelihoods as L, spacetime as S
    z = _grid(3)
    mk = lambda: K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=4.0),  # noqa: E731
                                        K.Matern32(variance=1.0, lengthscale=0.3), z=z, sparse=True,
                                        conditional="Full")
    data = [_st_data(40 + k, 20, 12) for k in range(3)]
    gm = S.SpaceTimeVariationalMeanFieldGP.from_runs(mk(), L.Poisson(binsize=2.0), [d[0] for d in data],
                                                     [d[1] for d in data], [d[2] for d in data])
    singles = [S.SpaceTimeVariationalMeanFieldGP(mk(), L.Poisson(binsize=2.0), d[0], d[2], d[1]) for d in data]
    for _ in range(2):
        gm.inference(lr=1.0)
        for s in singles:
            s.inference(lr=1.0)
    e = gm.energy_per_run().cpu().numpy()
    pm = gm.posterior_mean.cpu().numpy()
    for k, s in enumerate(singles):
        assert np.max(np.abs(pm[k] - s.posterior_mean.cpu().numpy())) <= 1e-10 * np.max(np.abs(pm[k]))
        assert abs(e[k] - float(s.energy())) <= 1e-10 * abs(e[k])
