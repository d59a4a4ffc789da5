"""GPU parity of mini-batched inference and energy on the temporal models
(inference(batch_ind=...), energy(batch_ind=...); inference.py:65-90, 130-325 with
group_natural_params, basemodels.py:195-199) against the oracle (pinned in
tests/test_oracle_minibatch.py): a sequence of steps on different random batches,
after each the sites, the posterior, diff1 / diff2, the batch's returned terms and the
energy of a batch.  FP32 posterior transfer as in the other model tests: 1e-5."""
import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _rel(a, b):
    a = np.asarray(a.cpu().numpy() if hasattr(a, "cpu") else a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("kind", ["VI", "Newton", "EP"])
def test_minibatch_steps_match_oracle(kind):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(31)
    n = 400
    x = np.sort(rng.uniform(0, 100, n)).astype(np.float32).astype(np.float64)
    f = np.sin(0.2 * x)
    if kind == "EP":
        y = (rng.uniform(size=n) < 0.5 * (1 + np.tanh(f))).astype(np.float64)
        gl, ol = L.Bernoulli(), OL.Bernoulli()
    else:
        y = (f + 0.3 * rng.standard_t(3.0, size=n)).astype(np.float32).astype(np.float64)
        gl, ol = L.StudentsT(scale=0.3, df=3.0), OL.StudentsT(0.3, 3.0)
    y[rng.choice(n, 20, replace=False)] = np.nan
    cls = {"VI": M.MarkovVariationalGP, "Newton": M.MarkovNewtonGP, "EP": M.MarkovExpectationPropagationGP}[kind]
    kw = {"power": 0.7} if kind == "EP" else {}
    g = cls(K.Matern52(variance=1.1, lengthscale=4.0), gl, x, y, parallel=True, **kw)
    o = OM.MarkovGP(OK.Matern52(1.1, 4.0), ol, x[:, None], y[:, None], inference=kind, power=kw.get("power", 1.0))
    print(f"[minibatch] {kind}", flush=True)
    for it, (bs, lr) in enumerate(((400, 1.0), (97, 0.5), (150, 0.8), (33, 1.0))):
        batch = np.sort(rng.choice(n, bs, replace=False)) if bs < n else None
        (gm, gj, gh), (gd1, gd2) = g.inference(lr=lr, batch_ind=batch)
        (om, oj, oh), (od1, od2) = o.inference(lr=lr, batch_ind=batch)
        n1 = g.plan.scatter(g._nat1[None]).cpu().numpy().reshape(-1)
        n2 = g.plan.scatter(g._nat2[None]).cpu().numpy().reshape(-1)
        e1, e2 = _rel(n1, o.nat1), _rel(n2, o.nat2)
        em = _rel(g.posterior_mean, o.posterior_mean)
        ev = _rel(g.posterior_variance, o.posterior_variance)
        et = _rel(gm, om)
        ed = max(abs(float(gd1) - od1) / max(od1, 1e-12), abs(float(gd2) - od2) / max(od2, 1e-12))
        sel = np.arange(n) if batch is None else batch
        eb = float(g.energy(batch_ind=sel[: max(1, len(sel) // 2)]))
        ob = o.energy(batch_ind=sel[: max(1, len(sel) // 2)])
        ee = abs(eb - ob) / abs(ob)
        print(f"[minibatch] {kind} step {it + 1} (batch {bs}): nat1 {e1:.1e} nat2 {e2:.1e} mean {em:.1e} "
              f"var {ev:.1e} terms {et:.1e} diff {ed:.1e} energy {ee:.1e}", flush=True)
        assert gm.shape[0] == len(sel)
        assert max(e1, e2, em, ev, et, ed, ee) < TOL, (it, e1, e2, em, ev, et, ed, ee)


def _grid(g):
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    return np.stack([a.ravel(), b.ravel()], 1)


@pytest.mark.parametrize("kind", ["VI", "Newton"])
def test_spacetime_minibatch_steps_match_oracle(kind):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(32)
    T_, R_ = 30, 6
    z = _grid(2)
    t = np.arange(T_, dtype=np.float64)
    r = rng.uniform(size=(T_, R_, 2)).astype(np.float32).astype(np.float64)
    y = rng.poisson(2.0 * np.exp(np.sin(0.3 * t)[:, None] + r[..., 0])).astype(np.float64)
    y[4, 2] = np.nan
    gk = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=5.0), K.Matern32(variance=1.0, lengthscale=0.3),
                                z=z, sparse=True, conditional="Full")
    ok = OK.SpatioTemporalKernel(OK.Matern32(1.0, 5.0), OK.Matern32(1.0, 0.3), z, sparse=True, conditional="Full")
    cls = M.MarkovVariationalGP if kind == "VI" else M.MarkovNewtonGP
    g = cls(gk, L.Poisson(binsize=2.0), t, y, R=r)
    o = OM.MarkovGP(ok, OL.Poisson(2.0), t[:, None], y, R=r, inference=kind)
    print(f"[minibatch] space-time {kind}", flush=True)
    for it, (bs, lr) in enumerate(((30, 1.0), (11, 0.5), (7, 1.0))):
        batch = np.sort(rng.choice(T_, bs, replace=False)) if bs < T_ else None
        (gm, _, _), (gd1, _) = g.inference(lr=lr, batch_ind=batch)
        (om, _, _), (od1, _) = o.inference(lr=lr, batch_ind=batch)
        n1, _ = g.site_nat
        e1 = _rel(n1, o.nat1[..., 0])
        em = _rel(g.posterior_mean, o.posterior_mean)
        et = _rel(gm, om[..., 0])
        ed = abs(float(gd1) - od1) / od1
        sel = np.arange(T_) if batch is None else batch
        eb, ob = float(g.energy(batch_ind=sel[:4])), o.energy(batch_ind=sel[:4])
        ee = abs(eb - ob) / abs(ob)
        print(f"[minibatch] space-time {kind} step {it + 1} (batch {bs}): nat1 {e1:.1e} mean {em:.1e} "
              f"terms {et:.1e} diff {ed:.1e} energy {ee:.1e}", flush=True)
        assert gm.shape[0] == len(sel)
        assert max(e1, em, et, ed, ee) < 1e-4, (it, e1, em, et, ed, ee)
