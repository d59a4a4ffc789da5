"""Pins the oracle's space-time branch (kernels.py:385-586 SpatioTemporalKernel with the
'Full' conditional; basemodels.py:667-764 projection, KL and data conditional).

* k_t(0) B B' + C = k_t(0) Krr at every step (the spatial conditional is exact
  within a time slice: B = Krz Kzz^-1 Lzz, C = k_t(0)(Krr - Krz Kzz^-1 Kzr)).
* With a Gaussian likelihood one VI step (lr = 1) is exact, and the energy equals
  the closed-form collapsed bound of the model f_t = B_t u_t + e_t, u ~ GP(0, k_t) (x) I_M:
  -[log N(y | 0, K_f + s2 I) - tr(C) / (2 s2)], K_f[(a,i),(b,j)] = k_t(t_a,t_b) B_a[i] . B_b[j].
  A dropped C term, a transposed B or a wrong KL sign each fail it.
* With the inducing grid as the observation grid (R_t = z), C = 0 and the bound is the
  exact marginal likelihood of the separable kernel k_t (x) k_s.
"""
import math

import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM


def _grid(g):
    gz = np.linspace(0, 1, g)
    a, b = np.meshgrid(gz, gz)
    return np.stack([a.ravel(), b.ravel()], 1)


def _collapsed_energy(kt, kern, t, r, y, s2):
    B, C = kern.spatial_conditional(t[:, None], r)
    Kt = kt(t[:, None], t[:, None])
    T_, R_ = y.shape
    N = T_ * R_
    Kf = np.einsum("ab,aim,bjm->aibj", Kt, B, B).reshape(N, N)
    Ky = Kf + s2 * np.eye(N)
    yv = y.reshape(-1)
    _, logdet = np.linalg.slogdet(Ky)
    logml = -0.5 * (yv @ np.linalg.solve(Ky, yv) + logdet + N * math.log(2 * math.pi))
    return -(logml - 0.5 * np.trace(C, axis1=1, axis2=2).sum() / s2)


@pytest.mark.parametrize("seed,g,R_,ls_t,ls_s", [(1, 2, 7, 2.0, 0.4), (2, 3, 11, 0.7, 0.25)])
def test_spacetime_vi_energy_matches_collapsed_bound(seed, g, R_, ls_t, ls_s):
    rng = np.random.default_rng(seed)
    T_ = 12
    z = _grid(g)
    t = np.sort(rng.uniform(0, 10, T_))
    r = rng.uniform(size=(T_, R_, 2))
    y = rng.standard_normal((T_, R_))
    vt, vs, s2 = 1.3, 0.8, 0.2
    kt, ks = OK.Matern32(vt, ls_t), OK.Matern32(vs, ls_s)
    kern = OK.SpatioTemporalKernel(kt, ks, z, sparse=True, conditional="Full")
    o = OM.MarkovGP(kern, OL.Gaussian(s2), t[:, None], y, R=r, inference="VI")
    o.inference(lr=1.0)
    closed = _collapsed_energy(kt, kern, t, r, y, s2)
    assert abs(o.energy() - closed) <= 1e-10 * abs(closed)


def test_spatial_conditional_reproduces_slice_covariance():
    rng = np.random.default_rng(3)
    z = _grid(3)
    r = rng.uniform(size=(4, 9, 2))
    kt, ks = OK.Matern32(1.7, 3.0), OK.Matern32(0.6, 0.3)
    kern = OK.SpatioTemporalKernel(kt, ks, z)
    B, C = kern.spatial_conditional(np.arange(4.0)[:, None], r)
    for k in range(4):
        np.testing.assert_allclose(1.7 * B[k] @ B[k].T + C[k], 1.7 * ks(r[k], r[k]), atol=1e-13)


def test_gridded_observations_give_exact_separable_marginal():
    rng = np.random.default_rng(4)
    T_ = 10
    z = _grid(3)
    t = np.sort(rng.uniform(0, 5, T_))
    r = np.broadcast_to(z, (T_,) + z.shape).copy()
    y = rng.standard_normal((T_, z.shape[0]))
    kt, ks = OK.Matern32(1.1, 1.5), OK.Matern32(1.0, 0.5)
    kern = OK.SpatioTemporalKernel(kt, ks, z)
    o = OM.MarkovGP(kern, OL.Gaussian(0.3), t[:, None], y, R=r, inference="VI")
    o.inference(lr=1.0)
    K = np.kron(kt(t[:, None], t[:, None]), ks(z, z)) + 0.3 * np.eye(y.size)
    yv = y.reshape(-1)
    _, logdet = np.linalg.slogdet(K)
    logml = -0.5 * (yv @ np.linalg.solve(K, yv) + logdet + yv.size * math.log(2 * math.pi))
    assert abs(o.energy() + logml) <= 1e-9 * abs(logml)
