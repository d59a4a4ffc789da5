"""GPU parity of the Markov-GP models' inference() / energy() (inference.py:65-325)
against the FP64 oracle (oracle/models.py) on identical inputs.

Inputs are on a 2^-10 grid and FP32-representable, so the GPU (FP32 per-step
inputs) and the oracle (FP64) see the same data.  Tolerances: <= 1e-4 relative
on posterior means (relative to the series' scale) and variances, <= 1e-5
relative on the energy, after every inference step.
"""
import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM

pytestmark = pytest.mark.gpu

MEAN_TOL, VAR_TOL, E_TOL = 1e-4, 1e-4, 1e-5


def _grid_x(n, span, rng):
    return np.sort(rng.choice(int(span * 1024), size=n, replace=False)) / 1024.0


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _pair(kind, lik_name, kern_name, x, y, power=1.0):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    kern = {"m52": (K.Matern52(variance=1.0, lengthscale=1.0), OK.Matern52(1.0, 1.0)),
            "m32+m52": (K.Sum([K.Matern32(variance=0.5, lengthscale=10.0), K.Matern52(variance=0.5, lengthscale=10.0)]),
                        OK.Sum([OK.Matern32(0.5, 10.0), OK.Matern52(0.5, 10.0)])),
            "m52+m32": (K.Sum([K.Matern52(variance=1.0, lengthscale=2.0), K.Matern32(variance=1.0, lengthscale=2.0)]),
                        OK.Sum([OK.Matern52(1.0, 2.0), OK.Matern32(1.0, 2.0)]))}[kern_name]
    lik = {"gauss": (L.Gaussian(variance=0.01), OL.Gaussian(0.01)),
           "probit": (L.Bernoulli(), OL.Bernoulli()),
           "studentt": (L.StudentsT(scale=0.2, df=3.0), OL.StudentsT(0.2, 3.0)),
           "poisson": (L.Poisson(binsize=2.0), OL.Poisson(2.0))}[lik_name]
    cls = {"VI": M.MarkovVariationalGP, "Newton": M.MarkovLaplaceGP, "EP": M.MarkovExpectationPropagationGP}[kind]
    if kind == "EP":
        g = cls(kern[0], lik[0], x, y, power=power, parallel=True)
    else:
        g = cls(kern[0], lik[0], x, y, parallel=True)
    o = OM.MarkovGP(kern[1], lik[1], x, y, inference=kind, power=power)
    return g, o


def _data(lik_name, n, seed):
    rng = np.random.default_rng(seed)
    x = _grid_x(n, n / 20.0, rng)
    f = np.sin(0.3 * x) + 0.5 * np.sin(1.7 * x)
    if lik_name == "probit":
        y = (rng.uniform(size=n) < 0.5 * (1 + np.tanh(f))).astype(float)
    elif lik_name == "poisson":
        y = rng.poisson(np.exp(0.5 * f)).astype(float)
    elif lik_name == "studentt":
        y = f + 0.2 * rng.standard_t(3.0, size=n)
    else:
        y = f + 0.1 * rng.standard_normal(n)
    return x, _f32(y)


def _check(g, o):
    pm = g.posterior_mean.cpu().numpy().reshape(-1)
    pv = g.posterior_variance.cpu().numpy().reshape(-1)
    om = o.posterior_mean.reshape(-1)
    ov = o.posterior_variance.reshape(-1)
    em = np.max(np.abs(pm - om)) / np.max(np.abs(om))
    ev = np.max(np.abs(pv - ov) / np.abs(ov))
    assert em < MEAN_TOL, em
    assert ev < VAR_TOL, ev


@pytest.mark.parametrize("kind,lik,kern,power", [
    ("VI", "gauss", "m52", 1.0),
    ("VI", "studentt", "m52+m32", 1.0),
    ("VI", "probit", "m32+m52", 1.0),
    ("VI", "poisson", "m52", 1.0),
    ("Newton", "probit", "m32+m52", 1.0),
    ("Newton", "studentt", "m52", 1.0),
    ("EP", "probit", "m32+m52", 1.0),
    ("EP", "gauss", "m52", 0.5),
    ("EP", "studentt", "m52", 0.5),
])
def test_inference_and_energy_match_oracle(kind, lik, kern, power):
    x, y = _data(lik, 1500, 3)
    g, o = _pair(kind, lik, kern, x, y, power)
    for it in range(3):
        (gm, gj, gh), (gd1, gd2) = g.inference(lr=1.0 if it == 0 else 0.7)
        (om, oj, oh), (od1, od2) = o.inference(lr=1.0 if it == 0 else 0.7)
        _check(g, o)
        assert abs(float(gd1) - od1) <= 1e-4 * max(abs(od1), 1e-3)
        assert abs(float(gd2) - od2) <= 1e-4 * max(abs(od2), 1e-3)
        ge, oe = float(g.energy()), float(o.energy())
        assert abs(ge - oe) <= E_TOL * abs(oe), (it, ge, oe)
    # returned site terms (mean, jacobian, hessian) of the last step
    np.testing.assert_allclose(gh.cpu().numpy().reshape(-1), oh.reshape(-1), rtol=1e-4, atol=1e-6)


def test_gaussian_vi_energy_is_exact_marginal():
    """tests/test_vs_exact_marg_lik.py:41-65 on the GPU: one VI step with a Gaussian
    likelihood gives energy == -log N(y | 0, K + s2 I)."""
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(0)
    x = _grid_x(800, 40.0, rng)
    y = _f32(np.sin(x) + 0.3 * rng.standard_normal(800))
    g = M.MarkovVariationalGP(K.Matern52(variance=1.5, lengthscale=0.75), L.Gaussian(variance=0.1), x, y)
    g.inference(lr=1.0)
    dense = OM.DenseGP(OK.Matern52(1.5, 0.75), x, y, np.full(800, 0.1))
    assert abs(float(g.energy()) + dense.log_marginal()) < E_TOL * abs(dense.log_marginal())


def test_missing_observations():
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(1)
    x = _grid_x(600, 30.0, rng)
    y = _f32(np.sin(x) + 0.2 * rng.standard_normal(600))
    y[rng.uniform(size=600) < 0.1] = np.nan
    g = M.MarkovVariationalGP(K.Matern52(variance=1.0, lengthscale=1.0), L.StudentsT(scale=0.2), x, y)
    o = OM.MarkovGP(OK.Matern52(1.0, 1.0), OL.StudentsT(0.2, 3.0), x, y, inference="VI")
    for _ in range(2):
        g.inference(lr=1.0)
        o.inference(lr=1.0)
        _check(g, o)
    assert abs(float(g.energy()) - o.energy()) <= E_TOL * abs(o.energy())


def test_batched_series_match_independent_models():
    """from_series: B ragged series with per-series hypers == B independent oracle models."""
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(2)
    lens = [37, 400, 1, 250, 900, 64]
    xs, ys = [], []
    for n in lens:
        x = _grid_x(n, max(n / 10.0, 1.0), rng)
        xs.append(x)
        ys.append(_f32(np.sin(0.7 * x) + 0.2 * rng.standard_t(3.0, size=n)))
    ls = rng.uniform(0.5, 4.0, len(lens))
    LS = np.stack([ls, ls], 1)
    VAR = np.ones((len(lens), 2))
    g = M.MarkovVariationalGP.from_series(K.Sum([K.Matern52(), K.Matern32()]), L.StudentsT(scale=0.2), xs, ys,
                                          lengthscale=LS, variance=VAR, chunk_len=16)
    os_ = [OM.MarkovGP(OK.Sum([OK.Matern52(1.0, l), OK.Matern32(1.0, l)]), OL.StudentsT(0.2, 3.0), x, y)
           for x, y, l in zip(xs, ys, ls)]
    for _ in range(2):
        g.inference(lr=1.0)
        for o in os_:
            o.inference(lr=1.0)
    pm = g.posterior_mean.cpu().numpy().reshape(-1)
    pv = g.posterior_variance.cpu().numpy().reshape(-1)
    e = g.energy_per_series().cpu().numpy()
    for b, o in enumerate(os_):
        s = slice(g.offsets[b], g.offsets[b + 1])
        om = o.posterior_mean.reshape(-1)
        assert np.max(np.abs(pm[s] - om)) / max(np.max(np.abs(om)), 1e-2) < MEAN_TOL
        assert np.max(np.abs(pv[s] - o.posterior_variance.reshape(-1)) / o.posterior_variance.reshape(-1)) < VAR_TOL
        assert abs(e[b] - o.energy()) <= E_TOL * max(abs(o.energy()), 1.0)


def test_energy_keeps_posterior_fixed_under_new_hypers():
    """energy() after a hyper change uses the stored posterior and a fresh log Z
    (inference.py:197-222 with compute_kl, basemodels.py:708-724)."""
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(4)
    x = _grid_x(500, 25.0, rng)
    y = _f32(np.sin(x) + 0.2 * rng.standard_normal(500))
    g = M.MarkovVariationalGP(K.Matern52(variance=1.0, lengthscale=1.0), L.Gaussian(variance=0.05), x, y)
    o = OM.MarkovGP(OK.Matern52(1.0, 1.0), OL.Gaussian(0.05), x, y)
    g.inference(lr=1.0)
    o.inference(lr=1.0)
    g.kernel.lengthscale = 1.3
    o.kernel.lengthscale = 1.3
    assert abs(float(g.energy()) - o.energy()) <= E_TOL * abs(o.energy())
    _check(g, o)  # the posterior was not recomputed


def test_autotuned_chunk_length_gives_the_same_posterior():
    """chunk_len='auto' (bayesnewton_amd.autotune: the fused pass timed on the batch for a
    few chunk lengths) changes only the schedule: the posterior and energy equal those of a
    fixed chunk length."""
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(21)
    lens = rng.integers(40, 900, 30)
    xs = [np.sort(rng.uniform(0, n / 10.0, n)) for n in lens]
    ys = [(np.sin(x) + 0.3 * rng.standard_normal(x.size)).astype(np.float32).astype(np.float64) for x in xs]
    mk = lambda c: M.MarkovVariationalGP.from_series(  # noqa: E731
        K.Sum([K.Matern52(variance=1.0, lengthscale=2.0), K.Matern32(variance=0.5, lengthscale=0.7)]),
        L.StudentsT(scale=0.2, df=3.0), xs, ys, chunk_len=c)
    ma, m32 = mk("auto"), mk(32)
    assert ma.autotune["schedule"] in ("chunked", "sequential")
    assert ma.autotune["schedule"] == "sequential" or str(ma.autotune["chunk_len"]) in ma.autotune["ms_per_pass"]
    print(f"[autotune] {ma.autotune}", flush=True)
    for m in (ma, m32):
        m.inference(lr=1.0)
    a = ma.posterior_mean.cpu().numpy().reshape(-1)
    b = m32.posterior_mean.cpu().numpy().reshape(-1)
    assert np.max(np.abs(a - b)) <= 1e-5 * max(np.max(np.abs(b)), 1.0)
    ea, eb = float(ma.energy()), float(m32.energy())
    assert abs(ea - eb) <= 1e-9 * abs(eb)
