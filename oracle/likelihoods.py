"""FP64 likelihoods and Gauss-Hermite cubature of the oracle (test infrastructure only).

Restates bayesnewton/likelihoods.py (Gaussian 684-803, Bernoulli/probit 806-871,
Poisson 891-1008, StudentsT 1011-1044, base-class site maps 326-412) and
bayesnewton/cubature.py (gauss_hermite 76-84, variational_expectation_cubature
198-246, log_density_cubature 249-276, moment_match_cubature 310-371).

All site maps here are for scalar sites (p = 1 per site, the only case on the
hot path); they are vectorised over sites: y, m, v are 1-D arrays.  The
reference's autodiff site derivatives (jacrev, likelihoods.py:326-335) are
replaced by analytic derivatives, pinned against central finite differences.
"""
import math

import numpy as np
from numpy.polynomial.hermite import hermgauss
from scipy.special import erf, gammaln

LOG2PI = math.log(2 * math.pi)


def gauss_hermite(num_quad_pts=20):
    """cubature.py:76-84 (dim=1): nodes sqrt(2) x_i, weights w_i / sqrt(pi)."""
    x, w = hermgauss(num_quad_pts)
    return np.sqrt(2.0) * x, w * np.pi ** -0.5


def _cubature_points(m, v, num=20):
    x, w = gauss_hermite(num)
    sig = np.sqrt(v)[:, None] * x[None, :] + m[:, None]
    return sig, w


def variational_expectation_cubature(loglik_fn, y, m, v, num=20):
    """cubature.py:198-246 for 1-D sites: E, dE/dm, d2E/dm2 = 2 dE/dv."""
    sig, w = _cubature_points(m, v, num)
    wl = w[None, :] * loglik_fn(y[:, None], sig)
    E = np.sum(wl, axis=1)
    invv = 1.0 / v
    dE_dm = np.sum(invv[:, None] * (sig - m[:, None]) * wl, axis=1)
    dE_dv = np.sum((0.5 * invv[:, None] ** 2 * (sig - m[:, None]) ** 2 - 0.5 * invv[:, None]) * wl, axis=1)
    return E, dE_dm, 2.0 * dE_dv


def moment_match_cubature(loglik_fn, y, m, v, power=1.0, num=20):
    """cubature.py:310-371 for 1-D sites: log Z, dlogZ/dm, d2logZ/dm2 (Z clipped at 1e-8)."""
    sig, w = _cubature_points(m, v, num)
    wl = w[None, :] * np.exp(power * loglik_fn(y[:, None], sig))
    Z = np.sum(wl, axis=1)
    Zc = np.maximum(Z, 1e-8)
    lZ = np.log(Zc)
    Zinv = 1.0 / Zc
    invC = 1.0 / v
    d = sig - m[:, None]
    dZ = np.sum(invC[:, None] * d * wl, axis=1)
    dlZ = Zinv * dZ
    d2Z = np.sum((invC[:, None] ** 2 * d ** 2 - invC[:, None]) * wl, axis=1)
    d2lZ = -dlZ * dlZ + Zinv * d2Z
    return lZ, dlZ, d2lZ


def log_density_cubature(loglik_fn, y, m, v, num=20):
    """cubature.py:249-276"""
    sig, w = _cubature_points(m, v, num)
    Z = np.sum(w[None, :] * np.exp(loglik_fn(y[:, None], sig)), axis=1)
    return np.log(np.maximum(Z, 1e-8))


def predict_cubature(cond_moments, m, v, num=20):
    """cubature.py:438-465: E[y] = sum_i w_i E[y|f_i], Var[y] = sum_i w_i (Var[y|f_i] + E[y|f_i]^2) - E[y]^2."""
    sig, w = _cubature_points(m, v, num)
    mu, s2 = cond_moments(sig)
    Ey = np.sum(w[None, :] * mu, axis=1)
    Ey2 = np.sum(w[None, :] * (s2 + mu ** 2), axis=1)
    return Ey, Ey2 - Ey ** 2


class Likelihood:
    """Base-class site maps of likelihoods.py:299-412 (scalar sites, NaN = missing)."""

    def log_lik(self, y, f):
        raise NotImplementedError

    def log_lik_derivs(self, y, f):
        raise NotImplementedError

    # likelihoods.py:337-355
    def log_likelihood_gradients(self, y, f):
        y = np.asarray(y, np.float64)
        f = np.asarray(f, np.float64)
        mask = np.isnan(y)
        yy = np.where(mask, f, y)
        ll, J, H = self.log_lik_derivs(yy, f)
        return np.where(mask, 0.0, ll), np.where(mask, np.nan, J), np.where(mask, np.nan, H)

    def variational_expectation_(self, y, m, v, num=20):
        return variational_expectation_cubature(self.log_lik, y, m, v, num)

    # likelihoods.py:363-383
    def variational_expectation(self, y, m, v, num=20):
        y = np.asarray(y, np.float64)
        m = np.asarray(m, np.float64)
        v = np.asarray(v, np.float64)
        mask = np.isnan(y)
        yy = np.where(mask, m, y)
        E, dE, d2E = self.variational_expectation_(yy, m, v, num)
        return np.where(mask, 0.0, E), np.where(mask, np.nan, dE), np.where(mask, np.nan, d2E)

    def moment_match_(self, y, m, v, power=1.0, num=20):
        return moment_match_cubature(self.log_lik, y, m, v, power, num)

    # likelihoods.py:401-412
    def moment_match(self, y, m, v, power=1.0, num=20):
        y = np.asarray(y, np.float64)
        m = np.asarray(m, np.float64)
        v = np.asarray(v, np.float64)
        mask = np.isnan(y)
        yy = np.where(mask, m, y)
        return self.moment_match_(yy, m, v, power, num)

    def log_density(self, y, m, v, num=20):
        return log_density_cubature(self.log_lik, y, m, v, num)

    # likelihoods.py:493-505 (Likelihood.predict -> predict_cubature)
    def predict(self, m, v, num=20):
        return predict_cubature(self.conditional_moments, np.asarray(m, np.float64), np.asarray(v, np.float64), num)


class Gaussian(Likelihood):
    """likelihoods.py:684-803"""

    def __init__(self, variance=0.1):
        self.variance = float(variance)

    def log_lik(self, y, f):
        return -0.5 * np.log(2 * np.pi * self.variance) - 0.5 * (y - f) ** 2 / self.variance

    def log_lik_derivs(self, y, f):
        return self.log_lik(y, f), (y - f) / self.variance, -np.ones_like(f) / self.variance

    def variational_expectation_(self, y, m, v, num=20):
        """likelihoods.py:727-753 (closed form)"""
        E = -0.5 * np.log(2 * np.pi) - 0.5 * np.log(self.variance) - 0.5 * ((y - m) ** 2 + v) / self.variance
        return E, (y - m) / self.variance, -np.ones_like(m) / self.variance

    def moment_match_(self, y, m, v, power=1.0, num=20):
        """likelihoods.py:755-782: log N(y | m, s2/power + v) + pep_constant(s2, power)."""
        s2 = self.variance
        var = s2 / power + v
        lZ = -0.5 * (LOG2PI + np.log(var) + (y - m) ** 2 / var)
        const = 0.5 * ((1 - power) * LOG2PI - np.log(power)) + 0.5 * (1 - power) * np.log(s2)
        return lZ + const, (y - m) / var, -1.0 / var

    def log_density(self, y, m, v, num=20):
        """likelihoods.py:784-800"""
        var = self.variance + v
        return -0.5 * (LOG2PI + np.log(var) + (y - m) ** 2 / var)

    def predict(self, m, v, num=20):
        """likelihoods.py:802-803"""
        return np.asarray(m, np.float64), np.asarray(v, np.float64) + self.variance


PROBIT_JITTER = 1e-3  # likelihoods.py:828


def _probit_link(f):
    """likelihoods.py:828-829: Phi(f) (1 - 2 eps) + eps."""
    return 0.5 * (1.0 + erf(f / np.sqrt(2.0))) * (1 - 2 * PROBIT_JITTER) + PROBIT_JITTER


class Bernoulli(Likelihood):
    """likelihoods.py:806-871 (probit link; y in {0, 1}, y == 1 selects the link)."""

    def __init__(self, link="probit"):
        if link != "probit":
            raise NotImplementedError("only the probit link is on the hot path")
        self.link = link

    def log_lik(self, y, f):
        p = _probit_link(f)
        return np.log(np.where(np.equal(y, 1), p, 1 - p))

    def log_lik_derivs(self, y, f):
        """Analytic d/df and d2/df2 of log(eps + (1-2eps)Phi(+-f)) (replaces jacrev)."""
        eps = PROBIT_JITTER
        phi = np.exp(-0.5 * f ** 2) / np.sqrt(2 * np.pi)
        pos = np.equal(y, 1)
        s = np.where(pos, 1.0, -1.0)
        p = eps + (1 - 2 * eps) * 0.5 * (1.0 + erf(s * f / np.sqrt(2.0)))
        J = s * (1 - 2 * eps) * phi / p
        H = -(1 - 2 * eps) * s * f * phi / p - J ** 2
        return np.log(p), J, H

    def conditional_moments(self, f):
        """likelihoods.py:854-860"""
        p = _probit_link(f)
        return p, p - p ** 2


Probit = Bernoulli


class Poisson(Likelihood):
    """likelihoods.py:891-1008 (exp link)."""

    def __init__(self, binsize=1.0, link="exp"):
        if link != "exp":
            raise NotImplementedError("only the exp link is on the hot path")
        self.binsize = float(binsize)

    def log_lik(self, y, f):
        mu = np.exp(f) * self.binsize
        return y * np.log(mu) - mu - gammaln(y + 1)

    def log_lik_derivs(self, y, f):
        mu = np.exp(f) * self.binsize
        return self.log_lik(y, f), y - mu, -mu

    def variational_expectation_(self, y, m, v, num=20):
        """likelihoods.py:979-1008 (closed form)"""
        emc = self.binsize * np.exp(m + v / 2)
        E = y * np.log(self.binsize) + y * m - emc - gammaln(y + 1.0)
        return E, y - emc, -emc

    def conditional_moments(self, f):
        """likelihoods.py:952-960"""
        mu = np.exp(f) * self.binsize
        return mu, mu


class StudentsT(Likelihood):
    """likelihoods.py:1011-1044"""

    def __init__(self, scale=1.0, df=3.0):
        self.scale = float(scale)
        self.df = float(df)

    def log_lik(self, y, f):
        df, s = self.df, self.scale
        const = gammaln((df + 1.0) * 0.5) - gammaln(df * 0.5) - 0.5 * (np.log(s ** 2) + np.log(df) + np.log(np.pi))
        return const - 0.5 * (df + 1.0) * np.log(1.0 + (1.0 / df) * ((y - f) / s) ** 2)

    def log_lik_derivs(self, y, f):
        df, s2 = self.df, self.scale ** 2
        r = y - f
        den = df * s2 + r ** 2
        J = (df + 1.0) * r / den
        H = (df + 1.0) * (r ** 2 - df * s2) / den ** 2
        return self.log_lik(y, f), J, H

    def conditional_moments(self, f):
        """likelihoods.py:1043-1044"""
        return f, (self.scale ** 2) * (self.df / (self.df - 2.0)) * np.ones_like(f)
