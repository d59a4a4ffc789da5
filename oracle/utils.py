"""FP64 numerics helpers of the oracle (test infrastructure only).

Restates bayesnewton/utils.py of the reference.
"""
import math

import numpy as np

LOG2PI = math.log(2 * math.pi)
INV2PI = (2 * math.pi) ** -1


def softplus(x):
    """bayesnewton/utils.py:54-56"""
    return np.log(1.0 + np.exp(x))


def softplus_inv(x):
    """bayesnewton/utils.py:67-75"""
    return np.log(np.exp(x) - 1.0)


def cho_solve_batched(P, Q):
    """P^-1 Q for a (batch of) PSD P via Cholesky - bayesnewton/utils.py:14-19."""
    L = np.linalg.cholesky(P)
    Z = np.linalg.solve(L, Q)
    return np.linalg.solve(np.swapaxes(L, -1, -2), Z)


def solve(P, Q):
    """bayesnewton/utils.py:14-19"""
    return cho_solve_batched(P, Q)


def inv(P):
    """bayesnewton/utils.py:22-27 (and inv_vmap 30-35)"""
    eye = np.broadcast_to(np.eye(P.shape[-1]), P.shape)
    return cho_solve_batched(P, eye)


def diag(P):
    """bayesnewton/utils.py:38-42"""
    return np.diagonal(P, axis1=-2, axis2=-1)


def ensure_diagonal_positive_precision(K):
    """bayesnewton/utils.py:89-96: keep the diagonal, replace negatives with 1e-2."""
    d = diag(K)
    d = np.where(d < 0, 1e-2, d)
    out = np.zeros_like(K)
    idx = np.arange(K.shape[-1])
    out[..., idx, idx] = d
    return out


def mvn_logpdf(x, mean, cov, mask=None):
    """Batched restatement of bayesnewton/utils.py:376-396.

    x, mean: [..., p, 1]; cov: [..., p, p]; mask: [..., p, 1] bool (True = missing).
    Masked entries are made independent with variance 1/(2 pi) so that they add 0.
    """
    x = np.array(x, dtype=np.float64)
    mean = np.array(mean, dtype=np.float64)
    cov = np.array(cov, dtype=np.float64)
    if mask is not None:
        maskv = np.asarray(mask, dtype=bool).reshape(x.shape)
        x = np.where(maskv, 0.0, x)
        mean = np.where(maskv, 0.0, mean)
        pair = maskv | np.swapaxes(maskv, -1, -2)
        cov = np.where(pair, 0.0, cov)
        p = cov.shape[-1]
        idx = np.arange(p)
        dm = maskv[..., 0]
        cdiag = cov[..., idx, idx]
        cov[..., idx, idx] = np.where(dm, INV2PI, cdiag)
    n = mean.shape[-2]
    L = np.linalg.cholesky(cov)
    log_det = 2 * np.sum(np.log(np.abs(diag(L))), axis=-1)
    diff = x - mean
    z = np.linalg.solve(L, diff)
    distance = np.sum(z[..., 0] ** 2, axis=-1)
    return -0.5 * (distance + n * LOG2PI + log_det)


def mvn_logpdf_and_derivs(x, mean, cov):
    """bayesnewton/utils.py:448-466 (unmasked form used by Gaussian.moment_match_)."""
    n = mean.shape[-2]
    L = np.linalg.cholesky(cov)
    precision = inv(cov)
    log_det = 2 * np.sum(np.log(np.abs(diag(L))), axis=-1)
    diff = x - mean
    scaled_diff = precision @ diff
    distance = (np.swapaxes(diff, -1, -2) @ scaled_diff)[..., 0, 0]
    return -0.5 * (distance + n * LOG2PI + log_det), scaled_diff, -precision


def pep_constant(var, power, mask=None):
    """bayesnewton/utils.py:431-445 (batched over leading dims)."""
    dim = var.shape[-1]
    L = np.linalg.cholesky(var)
    log_diag = np.log(np.abs(diag(L)))
    if mask is not None:
        m = np.asarray(mask, dtype=bool).reshape(log_diag.shape)
        log_diag = np.where(m, 0.0, log_diag)
        dim = dim - np.sum(m, axis=-1)
    logdet = 2 * np.sum(log_diag, axis=-1)
    return 0.5 * dim * ((1 - power) * LOG2PI - np.log(power)) + 0.5 * (1 - power) * logdet


def gaussian_expected_log_lik(Y, q_mu, q_covar, noise, mask=None):
    """Batched restatement of bayesnewton/utils.py:510-531.

    E_q[log N(Y | f, noise)] with q = N(q_mu, q_covar).  Shapes [..., p, 1] / [..., p, p].
    """
    Y = np.array(Y, dtype=np.float64)
    q_mu = np.array(q_mu, dtype=np.float64)
    q_covar = np.array(q_covar, dtype=np.float64)
    noise = np.array(noise, dtype=np.float64)
    if mask is not None:
        maskv = np.asarray(mask, dtype=bool).reshape(Y.shape)
        q_mu = np.where(maskv, Y, q_mu)
        pair = maskv | np.swapaxes(maskv, -1, -2)
        p = noise.shape[-1]
        idx = np.arange(p)
        dm = maskv[..., 0]
        noise = np.where(pair, 0.0, noise)
        nd = noise[..., idx, idx]
        noise[..., idx, idx] = np.where(dm, INV2PI, nd)
        q_covar = np.where(pair, 0.0, q_covar)
        qd = q_covar[..., idx, idx]
        q_covar[..., idx, idx] = np.where(dm, 1e-20, qd)
    ml = mvn_logpdf(Y, q_mu, noise)
    trace_term = -0.5 * np.trace(solve(noise, q_covar), axis1=-2, axis2=-1)
    return ml + trace_term


def compute_cavity(post_mean, post_cov, site_nat1, site_nat2, power, jitter=1e-8):
    """bayesnewton/utils.py:534-541 (batched)."""
    p = post_cov.shape[-1]
    post_nat2 = inv(post_cov + jitter * np.eye(p))
    cav_cov = inv(post_nat2 - power * site_nat2)
    cav_mean = cav_cov @ (post_nat2 @ post_mean - power * site_nat1)
    return cav_mean, cav_cov


def input_admin(t, y, r=None):
    """bayesnewton/utils.py:234-265: sort by time, dt_0 = 0, dt_n = t_n - t_{n-1}."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if t.ndim < 2:
        t = t[:, None]
    if y.ndim < 2:
        y = y[:, None]
    if r is None:
        if t.shape[1] > 1:
            r = t[:, 1:]
            t = t[:, :1]
        else:
            r = np.nan * t
    r = np.asarray(r, dtype=np.float64)
    if r.ndim < 2:
        r = r[:, None]
    ind = np.argsort(t[:, 0], axis=0, kind="stable")
    t_train = t[ind]
    y_train = y[ind]
    r_train = r[ind]
    dt = np.concatenate([np.array([0.0]), np.diff(t_train[:, 0])])
    return t_train, y_train, r_train, dt
