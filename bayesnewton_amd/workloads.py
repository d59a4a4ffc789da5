"""Seeded synthetic benchmark workloads (FP64, host).

Each generator returns plain NumPy arrays; values a workload leaves open are
fixed here once:
  single_series_10k    Matern52 (l=1, s2=1), Gaussian(0.01); N=10,000 sorted
                       U[0,100]; y = sin(0.3x) + 0.5 sin(1.7x) + 0.1 eps.    seed 1001
  binary_200k          Matern32 + Matern52 (l=10, s2=0.5 each); N=200,000 sorted
                       U[0,1000]; f = 2 sin(0.05x) + GP draw; y = 1[u < Phi(f)].
                                                                              seed 1002
  batched_4096_series  B=4096 series, n_b = floor(exp(U[ln 256, ln 16384])), x_b
                       sorted U[0, 0.1 n_b]; Matern52 + Matern32 with one lengthscale
                       per series l_b ~ logU[0.5, 20], s2 = 1; y = GP draw + 0.2 t_3.
                                                                              seed 1003
  single_series_500k   Matern52 (l=1, s2=1); N=500,000 sorted U[0,5000];
                       y = GP draw + 0.1 eps.                                 seed 1004
  spacetime_cox        space-time log-Gaussian Cox process (docstring below). seed 1005
GP draws use the reference's state-space sampler (basemodels.py:833-879):
m0 = chol(Pinf) e, m <- A m + chol(Q + 1e-8 I) e, f = H m.
"""
import math

import numpy as np


# ---- closed-form Matern blocks (kernels.py:175-300), vectorised over steps ----
def _m32_A(lam, dt):
    e = np.exp(-dt * lam)
    A = np.empty(dt.shape + (2, 2))
    A[..., 0, 0] = e * (dt * lam + 1.0)
    A[..., 0, 1] = e * dt
    A[..., 1, 0] = -e * dt * lam * lam
    A[..., 1, 1] = e * (1.0 - dt * lam)
    return A


def _m52_A(lam, dt):
    dl = dt * lam
    e = np.exp(-dl)
    A = np.empty(dt.shape + (3, 3))
    A[..., 0, 0] = e * (dt * lam * (0.5 * dl + 1.0) + 1.0)
    A[..., 0, 1] = e * dt * (dl + 1.0)
    A[..., 0, 2] = e * 0.5 * dt * dt
    A[..., 1, 0] = e * dt * (-0.5 * dl * lam * lam)
    A[..., 1, 1] = e * (dt * lam * (1.0 - dl) + 1.0)
    A[..., 1, 2] = e * dt * (1.0 - 0.5 * dl)
    A[..., 2, 0] = e * dt * lam ** 3 * (0.5 * dl - 1.0)
    A[..., 2, 1] = e * dt * lam ** 2 * (dl - 3.0)
    A[..., 2, 2] = e * (dt * lam * (0.5 * dl - 2.0) + 1.0)
    return A


def _pinf(kind, ls, var):
    ls = np.asarray(ls, np.float64)
    var = np.asarray(var, np.float64)
    if kind == "m32":
        P = np.zeros(ls.shape + (2, 2))
        P[..., 0, 0] = var
        P[..., 1, 1] = 3.0 * var / ls ** 2
        return P
    kap = 5.0 / 3.0 * var / ls ** 2
    P = np.zeros(ls.shape + (3, 3))
    P[..., 0, 0] = var
    P[..., 0, 2] = P[..., 2, 0] = -kap
    P[..., 1, 1] = kap
    P[..., 2, 2] = 25.0 * var / ls ** 4
    return P


def _trans(kind, ls, dt):
    if kind == "m32":
        return _m32_A(math.sqrt(3.0) / ls, dt)
    return _m52_A(math.sqrt(5.0) / ls, dt)


def _chol_small(Q):
    """Batched Cholesky of 2x2 / 3x3 SPD matrices (closed form, lower)."""
    k = Q.shape[-1]
    L = np.zeros_like(Q)
    L[:, 0, 0] = np.sqrt(Q[:, 0, 0])
    L[:, 1, 0] = Q[:, 1, 0] / L[:, 0, 0]
    L[:, 1, 1] = np.sqrt(Q[:, 1, 1] - L[:, 1, 0] ** 2)
    if k == 3:
        L[:, 2, 0] = Q[:, 2, 0] / L[:, 0, 0]
        L[:, 2, 1] = (Q[:, 2, 1] - L[:, 2, 0] * L[:, 1, 0]) / L[:, 1, 1]
        L[:, 2, 2] = np.sqrt(Q[:, 2, 2] - L[:, 2, 0] ** 2 - L[:, 2, 1] ** 2)
    return L


def gp_draw(kinds, ls, var, dt_list, rng):
    """State-space prior draw of f = sum of components for each series.

    kinds: component names ('m32'/'m52'); ls/var: [B, n_comp]; dt_list: list of
    per-series step vectors (dt[0] = 0).  Vectorised over the series still
    running at each time index (series are visited longest first).
    """
    B = len(dt_list)
    ls = np.asarray(ls, np.float64).reshape(B, -1)
    var = np.asarray(var, np.float64).reshape(B, -1)
    lens = np.array([len(d) for d in dt_list])
    order = np.argsort(-lens, kind="stable")
    slens = lens[order]
    T = int(slens[0])
    dtm = np.zeros((B, T))
    for i, b in enumerate(order):
        dtm[i, :lens[b]] = dt_list[b]
    f = np.zeros((B, T))
    # number of series still running at time t (series sorted by length, descending)
    active = np.searchsorted(-slens, -np.arange(1, T + 1), side="right")
    for c, kind in enumerate(kinds):
        lsc, vc = ls[order, c], var[order, c]
        Pinf = _pinf(kind, lsc, vc)
        k = Pinf.shape[-1]
        m = np.einsum("bij,bj->bi", np.linalg.cholesky(Pinf), rng.standard_normal((B, k)))
        jit = 1e-8 * np.eye(k)
        for t in range(T):
            a = int(active[t])
            A = _trans(kind, lsc[:a], dtm[:a, t])
            Q = Pinf[:a] - A @ Pinf[:a] @ np.swapaxes(A, -1, -2)
            L = _chol_small(Q + jit)
            m[:a] = np.einsum("bij,bj->bi", A, m[:a]) + np.einsum("bij,bj->bi", L, rng.standard_normal((a, k)))
            f[:a, t] += m[:a, 0]
    out = [None] * B
    for i, b in enumerate(order):
        out[b] = f[i, :lens[b]]
    return out


def _dt(x):
    return np.concatenate([[0.0], np.diff(x)])


def single_series_10k():
    rng = np.random.default_rng(1001)
    x = np.sort(rng.uniform(0, 100, 10_000))
    y = np.sin(0.3 * x) + 0.5 * np.sin(1.7 * x) + 0.1 * rng.standard_normal(x.size)
    return dict(x=x, y=y, kinds=["m52"], ls=[1.0], var=[1.0], noise=0.01)


def binary_200k():
    rng = np.random.default_rng(1002)
    x = np.sort(rng.uniform(0, 1000, 200_000))
    f = 2 * np.sin(0.05 * x) + gp_draw(["m32", "m52"], [[10.0, 10.0]], [[0.5, 0.5]], [_dt(x)], rng)[0]
    from scipy.special import ndtr
    y = (rng.uniform(size=x.size) < ndtr(f)).astype(np.float64)
    return dict(x=x, y=y, f=f, kinds=["m32", "m52"], ls=[10.0, 10.0], var=[0.5, 0.5])


def batched_4096_series(B=4096, seed=1003):
    rng = np.random.default_rng(seed)
    n = np.floor(np.exp(rng.uniform(np.log(256), np.log(16384), B))).astype(np.int64)
    xs = [np.sort(rng.uniform(0, 0.1 * nb, nb)) for nb in n]
    ls_b = np.exp(rng.uniform(np.log(0.5), np.log(20.0), B))
    ls = np.stack([ls_b, ls_b], axis=1)
    var = np.ones((B, 2))
    dts = [_dt(x) for x in xs]
    fs = gp_draw(["m52", "m32"], ls, var, dts, rng)
    ys = [f + 0.2 * rng.standard_t(3.0, size=f.size) for f in fs]
    off = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    return dict(offsets=off, x=np.concatenate(xs), dt=np.concatenate(dts), y=np.concatenate(ys),
                kinds=["m52", "m32"], ls=ls, var=var, df=3.0, scale=0.2)


def single_series_500k():
    rng = np.random.default_rng(1004)
    x = np.sort(rng.uniform(0, 5000, 500_000))
    f = gp_draw(["m52"], [[1.0]], [[1.0]], [_dt(x)], rng)[0]
    y = f + 0.1 * rng.standard_normal(x.size)
    return dict(x=x, y=y, kinds=["m52"], ls=[1.0], var=[1.0], noise=0.01)


# ---- space-time Cox process ----------------------------------------------------
def _m32_cov(r, ls, var=1.0):
    """Matern-3/2 covariance of distances r (kernels.py:175-200)."""
    s = math.sqrt(3.0) * r / ls
    return var * (1.0 + s) * np.exp(-s)


def _sqdist(a, b):
    d = a[..., :, None, :] - b[..., None, :, :]
    return np.sum(d * d, axis=-1)


def spacetime_cox(seed=1005, T=2000, R=256, grid=8, ls_t=10.0, ls_s=0.2, var_t=1.0, binsize=2.0, device=None):
    """Space-time log-Gaussian Cox process on [0,1]^2 x {0..T-1}.

    Inducing inputs z: grid x grid points of linspace(0,1,grid)^2.  Each step draws R
    locations U[0,1]^2.  The latent field is drawn from the model's own prior
    (kernels.py:486-506 with the 'Full' conditional): M = grid^2 independent
    Matern-3/2 temporal states u (whitened inducing process), f_t = B_t u_t + e_t,
    e_t ~ N(0, C_t), B_t = Krz Kzz^-1 Lzz, C_t = k_t(0) (Krr - Krz Kzz^-1 Kzr).
    Counts y ~ Poisson(binsize exp(f)) (mean log(binsize) folded into the binsize).
    All random draws come from numpy's PCG64 stream of `seed`; with `device` (a torch
    device) the per-step linear algebra runs there in FP64 (same draws, same data up
    to rounding).
    """
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, grid)
    zA, zB = np.meshgrid(g, g)
    z = np.stack([zA.reshape(-1), zB.reshape(-1)], axis=1)
    M = z.shape[0]
    t = np.arange(T, dtype=np.float64)
    r = rng.uniform(0.0, 1.0, size=(T, R, 2))
    Kzz = _m32_cov(np.sqrt(np.maximum(_sqdist(z, z), 1e-36)), ls_s)
    Lzz = np.linalg.cholesky(Kzz)
    # inducing temporal states: x0 = chol(Pinf) e, x <- A x + chol(Q + 1e-8 I) e (basemodels.py:833-879)
    lam = math.sqrt(3.0) / ls_t
    Pinf = np.array([[var_t, 0.0], [0.0, 3.0 * var_t / ls_t ** 2]])
    A = _m32_A(lam, np.array(1.0))
    Q = Pinf - A @ Pinf @ A.T
    LQ = np.linalg.cholesky(Q + 1e-8 * np.eye(2))
    u = np.empty((T, M))
    x = rng.standard_normal((M, 2)) @ np.linalg.cholesky(Pinf).T
    u[0] = x[:, 0]
    for k in range(1, T):
        x = x @ A.T + rng.standard_normal((M, 2)) @ LQ.T
        u[k] = x[:, 0]
    eps = rng.standard_normal((T, R))
    if device is None:
        Krz = _m32_cov(np.sqrt(np.maximum(_sqdist(r, z[None]), 1e-36)), ls_s)          # [T, R, M]
        Bt = np.linalg.solve(Lzz, np.swapaxes(Krz, 1, 2))                             # Lzz^-1 Kzr  [T, M, R]
        f = np.einsum("tmr,tm->tr", Bt, u)
        jit = 1e-8 * np.eye(R)
        for k in range(T):
            Krr = _m32_cov(np.sqrt(np.maximum(_sqdist(r[k], r[k]), 1e-36)), ls_s)
            C = var_t * (Krr - Bt[k].T @ Bt[k])
            try:
                f[k] += np.linalg.cholesky(C + jit) @ eps[k]
            except np.linalg.LinAlgError:
                w, V = np.linalg.eigh(C)
                f[k] += V @ (np.sqrt(np.maximum(w, 0.0)) * (V.T @ eps[k]))
    else:
        import torch
        dev = torch.device(device)
        tr = torch.tensor(r, dtype=torch.float64, device=dev)
        tz = torch.tensor(z, dtype=torch.float64, device=dev)
        s3 = math.sqrt(3.0) / ls_s

        def m32(d2):
            d = torch.sqrt(torch.clamp(d2, min=1e-36)) * s3
            return (1.0 + d) * torch.exp(-d)

        Krz = m32(((tr[:, :, None, :] - tz[None, None]) ** 2).sum(-1))                 # [T, R, M]
        tLi = torch.tensor(np.linalg.solve(Lzz, np.eye(M)), dtype=torch.float64, device=dev)
        Bt = torch.matmul(tLi, Krz.transpose(1, 2))                                     # Lzz^-1 Kzr [T, M, R]
        fk = torch.einsum("tmr,tm->tr", Bt, torch.tensor(u, dtype=torch.float64, device=dev))
        te = torch.tensor(eps, dtype=torch.float64, device=dev)
        jit = 1e-8 * torch.eye(R, dtype=torch.float64, device=dev)
        for k0 in range(0, T, 250):
            k1 = min(T, k0 + 250)
            rr = tr[k0:k1]
            Krr = m32(((rr[:, :, None, :] - rr[:, None, :, :]) ** 2).sum(-1))
            C = var_t * (Krr - Bt[k0:k1].transpose(1, 2) @ Bt[k0:k1])
            Lc, info = torch.linalg.cholesky_ex(C + jit)
            bad = torch.nonzero(info).reshape(-1).tolist()
            fk[k0:k1] += (Lc @ te[k0:k1, :, None])[..., 0]
            for b in bad:  # not PD even with jitter: square root from the spectrum
                w, V = torch.linalg.eigh(C[b])
                fk[k0 + b] = (Bt[k0 + b].T @ torch.tensor(u[k0 + b], dtype=torch.float64, device=dev)
                              + V @ (torch.sqrt(torch.clamp(w, min=0.0)) * (V.T @ te[k0 + b])))
        f = fk.cpu().numpy()
    y = rng.poisson(binsize * np.exp(f)).astype(np.float64)
    return dict(t=t, r=r, y=y, f=f, z=z, ls_t=ls_t, ls_s=ls_s, var_t=var_t, binsize=binsize)
