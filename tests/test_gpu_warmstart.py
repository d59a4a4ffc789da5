"""GPU checks of the learned warm start (bayesnewton_amd/warmstart.py, bn_warmstart.hip):
network forward and gradient against the FP64 oracle (oracle/mlp.py, same BF16 weights),
features against their FP64 restatement, and the A/B of plain vs warm start to the
same convergence test on a held-out run."""
import numpy as np
import pytest
import torch

from oracle import mlp as OMLP

pytestmark = pytest.mark.gpu


def _grid(g):
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    return np.stack([a.ravel(), b.ravel()], 1)


def test_mlp_forward_and_gradient_match_oracle():
    from bayesnewton_amd import warmstart as W
    rng = np.random.default_rng(0)
    net = W.WarmStartMLP(seed=3)
    dev = net.device
    n = 333  # not a multiple of the row blocks
    X = torch.tensor(rng.standard_normal((n, W.F)), dtype=torch.float32, device=dev)
    Y = torch.tensor(rng.standard_normal((n, 2)), dtype=torch.float32, device=dev)
    norm = np.concatenate([rng.standard_normal(W.F) * 0.1, 1 + rng.uniform(size=W.F), [0.3, -0.5], [1.2, 0.8]])
    net.norm = torch.tensor(norm, dtype=torch.float32, device=dev)
    p = net.params.cpu().numpy()
    out = net.predict(X).cpu().numpy()
    ref, _ = OMLP.forward(p, norm.astype(np.float32), X.cpu().numpy())
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-4 * np.max(np.abs(ref)))
    g, loss = net.gradient(X, Y)
    L, gref = OMLP.loss_and_grad(p, norm.astype(np.float32), X.cpu().numpy(), Y.cpu().numpy())
    assert abs(float(loss) - L) <= 1e-4 * L
    np.testing.assert_allclose(g.cpu().numpy(), gref, rtol=1e-3, atol=1e-4 * np.max(np.abs(gref)))


def test_features_match_oracle():
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M, warmstart as W
    rng = np.random.default_rng(1)
    T_, R_ = 15, 20
    t = np.arange(T_, dtype=np.float64)
    r = rng.uniform(size=(T_, R_, 2)).astype(np.float32).astype(np.float64)
    y = rng.poisson(3.0, size=(T_, R_)).astype(np.float64)
    y[2, 3] = np.nan
    kern = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=10.0), K.Matern32(lengthscale=0.2),
                                  z=_grid(3))
    m = M.MarkovVariationalGP(kern, L.Poisson(binsize=2.0), t, y, R=r)
    X = W.site_features(m).cpu().numpy().reshape(T_, R_, W.F)
    ref = OMLP.features(t, r, y, 10.0, 0.2, 1.0, 2.0)
    np.testing.assert_allclose(X, ref, rtol=1e-5, atol=2e-5)


def _runs(seed0, n, T_, R_, z):
    from bayesnewton_amd import workloads
    out = []
    for k in range(n):
        d = workloads.spacetime_cox(seed=seed0 + k, T=T_, R=R_, grid=int(round(np.sqrt(len(z)))), ls_t=10.0, ls_s=0.2)
        out.append((d["t"], d["r"].astype(np.float32).astype(np.float64), d["y"]))
    return out


def test_warm_start_reaches_plain_energy_in_fewer_iterations():
    from bayesnewton_amd import kernels as K, likelihoods as L, spacetime as S, warmstart as W
    z = _grid(4)
    T_, R_ = 60, 64

    def kern():
        return K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=10.0), K.Matern32(lengthscale=0.2), z=z)

    train = _runs(500, 6, T_, R_, z)
    tm = S.SpaceTimeVariationalGP.from_runs(kern(), L.Poisson(binsize=2.0), [d[0] for d in train],
                                            [d[1] for d in train], [d[2] for d in train])
    W.run_to_convergence(tm, tol=1e-9, max_iter=100)
    X = W.site_features(tm)
    Y = W.site_marginals(tm)
    net = W.WarmStartMLP(seed=0)
    hist = net.fit(X, Y, epochs=40, batch=2048, lr=2e-3)
    assert hist[-1] < 0.5 * hist[0]
    test = _runs(900, 1, T_, R_, z)[0]
    plain = S.SpaceTimeVariationalGP(kern(), L.Poisson(binsize=2.0), test[0], test[2], test[1])
    warm = S.SpaceTimeVariationalGP(kern(), L.Poisson(binsize=2.0), test[0], test[2], test[1])
    it_p, e_p, _ = W.run_to_convergence(plain, tol=1e-7)
    net.init_sites(warm)
    it_w, e_w, _ = W.run_to_convergence(warm, tol=1e-7)
    assert abs(float(e_w[0]) - float(e_p[0])) <= 1e-5 * abs(float(e_p[0]))
    assert it_w < it_p, (it_w, it_p)


def test_shipped_network_loads_and_warm_starts_a_held_out_run():
    """bayesnewton_amd/pretrained/spacetime_cox_warmstart.npz through the package loader
    (no pickles): the parameter layout, and on a held-out Cox run (seed 1002, T = 200,
    R = 256, M = 64, the bench's lr 0.2) the same energy as the plain start in fewer
    iterations."""
    from bayesnewton_amd import kernels as K, likelihoods as L, spacetime as S, warmstart as W, workloads
    net, record = W.WarmStartMLP.pretrained()
    assert net.params.numel() == W.P and net.norm.numel() == 2 * W.F + 4
    z = _grid(8)

    def model():
        d = workloads.spacetime_cox(seed=1002, T=200)
        kern = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=10.0),
                                      K.Matern32(variance=1.0, lengthscale=0.2), z=z, sparse=True, conditional="Full")
        return S.SpaceTimeVariationalGP(kern, L.Poisson(binsize=2.0), d["t"], d["y"], d["r"])

    plain, warm = model(), model()
    it_p, e_p, _ = W.run_to_convergence(plain, tol=1e-7, max_iter=400, lr=0.2)
    net.init_sites(warm)
    it_w, e_w, _ = W.run_to_convergence(warm, tol=1e-7, max_iter=400, lr=0.2)
    print(f"[warmstart] shipped network: plain {it_p} iterations, warm {it_w}; energies {float(e_p[0])!r} "
          f"{float(e_w[0])!r}", flush=True)
    assert np.isfinite(float(e_p[0])) and np.isfinite(float(e_w[0]))
    assert abs(float(e_w[0]) - float(e_p[0])) <= 1e-5 * abs(float(e_p[0]))
    assert it_w < it_p, (it_w, it_p)
