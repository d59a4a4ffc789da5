"""Energy gradients of the space-time models (the reference differentiates energy()
through B and C, kernels.py:486-515, and the filter's log-likelihood;
demos/2d_log_gaussian_cox_process.py:46-57 trains the hypers with it).

* Gaussian likelihood after one VI step (lr = 1, exact sites): the energy is the
  collapsed bound of f_t = B_t u_t + e_t (pinned in tests/test_oracle_spacetime.py), and
  at the optimal sites its gradient in the hypers is the bound's gradient (envelope
  theorem), here by central differences of the dense closed form.
* Poisson likelihood after two VI steps: central differences of the oracle's energy with
  its sites and posterior held fixed, as objax.GradValues(energy) holds the StateVars.
Tolerance 1e-4 relative.
"""
import math

import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM

pytestmark = pytest.mark.gpu

G_TOL = 1e-4


def _grid(g):
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    return np.stack([a.ravel(), b.ravel()], 1)


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _sig(x):
    return 1.0 / (1.0 + np.exp(-x))


def _softplus_inv(x):
    return np.log(np.expm1(x))


def _collapsed(vt, lt, vs, ls, s2, t, r, y, z):
    kt, ks = OK.Matern32(vt, lt), OK.Matern32(vs, ls)
    kern = OK.SpatioTemporalKernel(kt, ks, z, sparse=True, conditional="Full")
    B, C = kern.spatial_conditional(t[:, None], r)
    Kt = kt(t[:, None], t[:, None])
    N = y.size
    Kf = np.einsum("ab,aim,bjm->aibj", Kt, B, B).reshape(N, N)
    Ky = Kf + s2 * np.eye(N)
    yv = y.reshape(-1)
    _, logdet = np.linalg.slogdet(Ky)
    logml = -0.5 * (yv @ np.linalg.solve(Ky, yv) + logdet + N * math.log(2 * math.pi))
    return -(logml - 0.5 * np.trace(C, axis1=1, axis2=2).sum() / s2)


def test_gaussian_gradient_matches_collapsed_bound():
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(21)
    T_, R_ = 12, 7
    z = _grid(3)
    t = np.sort(rng.uniform(0, 10, T_))
    r = _f32(rng.uniform(size=(T_, R_, 2)))
    y = _f32(rng.standard_normal((T_, R_)))
    vt, lt, vs, ls, s2 = 1.3, 2.0, 0.8, 0.4, 0.2
    kern = K.SpatioTemporalKernel(K.Matern32(variance=vt, lengthscale=lt), K.Matern32(variance=vs, lengthscale=ls), z=z)
    g = M.MarkovVariationalGP(kern, L.Gaussian(variance=s2), t, y, R=r)
    g.inference(lr=1.0)
    E, grad = g.energy_and_grad()
    names = g.trainable_names()
    assert names == ["temporal.transformed_lengthscale", "temporal.transformed_variance",
                     "spatial.transformed_lengthscale", "spatial.transformed_variance", "Gaussian.transformed_variance"]
    p0 = np.array([lt, vt, ls, vs, s2])

    def bound(p):
        return _collapsed(p[1], p[0], p[3], p[2], p[4], t, r, y, z)

    assert abs(E - bound(p0)) <= 1e-9 * abs(bound(p0))
    fd = []
    for i in range(5):
        h = 1e-6 * p0[i]
        e = np.zeros(5)
        e[i] = h
        fd.append((bound(p0 + e) - bound(p0 - e)) / (2 * h) * _sig(_softplus_inv(p0[i])))
    fd = np.array(fd)
    print(f"[st grad] gaussian: gpu {grad} closed form {fd}", flush=True)
    np.testing.assert_allclose(grad, fd, rtol=G_TOL, atol=G_TOL * np.max(np.abs(fd)))


@pytest.mark.parametrize("M_side", [3, 8])
def test_poisson_gradient_matches_oracle_fd(M_side):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(22 + M_side)
    T_, R_ = (10, 9) if M_side == 3 else (6, 256)
    z = _grid(M_side)
    t = np.arange(T_, dtype=np.float64)
    r = _f32(rng.uniform(size=(T_, R_, 2)))
    f = np.sin(0.3 * t)[:, None] + np.cos(3.0 * r[..., 0]) * np.sin(2.0 * r[..., 1])
    y = rng.poisson(2.0 * np.exp(f)).astype(np.float64)
    kern = K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=5.0), K.Matern32(variance=0.9, lengthscale=0.3),
                                  z=z)
    g = M.MarkovVariationalGP(kern, L.Poisson(binsize=2.0), t, y, R=r)
    okern = OK.SpatioTemporalKernel(OK.Matern32(1.0, 5.0), OK.Matern32(0.9, 0.3), z)
    o = OM.MarkovGP(okern, OL.Poisson(2.0), t[:, None], y, R=r, inference="VI")
    for lr in (1.0, 0.5):
        g.inference(lr=lr)
        o.inference(lr=lr)
    E, grad = g.energy_and_grad()
    assert abs(E - o.energy()) <= 1e-5 * abs(o.energy())
    fd = []
    for obj in (okern.temporal_kernel, okern.spatial_kernel):
        for attr in ("lengthscale", "variance"):
            v0 = getattr(obj, attr)
            h = 1e-6 * v0
            setattr(obj, attr, v0 + h)
            ep = o.energy()
            setattr(obj, attr, v0 - h)
            em = o.energy()
            setattr(obj, attr, v0)
            fd.append((ep - em) / (2 * h) * _sig(_softplus_inv(v0)))
    fd = np.array(fd)
    print(f"[st grad] poisson M={M_side * M_side}: gpu {grad} oracle fd {fd}", flush=True)
    np.testing.assert_allclose(grad, fd, rtol=G_TOL, atol=G_TOL * np.max(np.abs(fd)))
