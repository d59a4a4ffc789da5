"""State-space kernels of the Markov-GP hot path.

Mirrors the reference's kernel classes for the configurations this package
accelerates (bayesnewton/kernels.py): Matern12 (123-169), Matern32 (175-230),
Matern52 (233-300), their Sum (1624-1663) and the SpatioTemporalKernel
(385-586).  As in the reference (kernels.py:74-95), hyperparameters are stored
softplus-inverse transformed and exposed on the natural scale by properties.

The per-step transition A(dt) and process noise Q(dt) are never materialised
by the hot path: the HIP kernels build them in registers from the component
kinds and hyperparameters held here (`desc()` / `hypers()`).
"""
import math

import numpy as np
import torch

from . import _native as nat


def softplus(x):
    """bayesnewton/utils.py:54-56"""
    return math.log1p(math.exp(x)) if x < 30 else x + math.log1p(math.exp(-x))


def softplus_inv(x):
    """bayesnewton/utils.py:67-75"""
    return math.log(math.expm1(x)) if x < 30 else x + math.log(-math.expm1(-x))


class StationaryKernel:
    """bayesnewton/kernels.py:70-120 (hypers held transformed)."""
    state_dim = None
    kind = None  # BN_MATERN12 / 32 / 52

    def __init__(self, variance=1.0, lengthscale=1.0, fix_variance=False, fix_lengthscale=False):
        self.transformed_variance = softplus_inv(float(variance))
        self.transformed_lengthscale = softplus_inv(float(lengthscale))
        self.fix_variance = fix_variance
        self.fix_lengthscale = fix_lengthscale

    @property
    def variance(self):
        return softplus(self.transformed_variance)

    @variance.setter
    def variance(self, v):
        self.transformed_variance = softplus_inv(float(v))

    @property
    def lengthscale(self):
        return softplus(self.transformed_lengthscale)

    @lengthscale.setter
    def lengthscale(self, v):
        self.transformed_lengthscale = softplus_inv(float(v))

    # ---- descriptor for the HIP library ------------------------------------------
    @property
    def components(self):
        return [self]

    def desc(self):
        comps = self.components
        if len(comps) > 2:
            raise NotImplementedError("the HIP path supports sums of at most two Matern components")
        d = nat.KernelDesc()
        d.n_comp = len(comps)
        for i, c in enumerate(comps):
            d.kind[i] = c.kind
        return d

    def hypers(self, device):
        """(lengthscale[n_comp], variance[n_comp]) as FP64 device tensors."""
        comps = self.components
        ls = torch.tensor([c.lengthscale for c in comps], dtype=torch.float64, device=device)
        var = torch.tensor([c.variance for c in comps], dtype=torch.float64, device=device)
        return ls, var

    def measurement_model(self):
        H = np.zeros((1, self.state_dim))
        H[0, 0] = 1.0
        return H

    def K(self, X, X2):
        X = np.asarray(X, dtype=np.float64).reshape(len(X), -1)
        X2 = np.asarray(X2, dtype=np.float64).reshape(len(X2), -1)
        r2 = np.sum(((X[:, None, :] - X2[None, :, :]) / self.lengthscale) ** 2, axis=-1)
        return self.K_r(np.sqrt(np.maximum(r2, 1e-36)))

    __call__ = K


class Matern12(StationaryKernel):
    """bayesnewton/kernels.py:123-169"""
    state_dim = 1
    kind = nat.BN_MATERN12

    def K_r(self, r):
        return self.variance * np.exp(-r)


Exponential = Matern12


class Matern32(StationaryKernel):
    """bayesnewton/kernels.py:175-230"""
    state_dim = 2
    kind = nat.BN_MATERN32

    def K_r(self, r):
        s3 = math.sqrt(3.0)
        return self.variance * (1.0 + s3 * r) * np.exp(-s3 * r)


class Matern52(StationaryKernel):
    """bayesnewton/kernels.py:233-300"""
    state_dim = 3
    kind = nat.BN_MATERN52

    def K_r(self, r):
        s5 = math.sqrt(5.0)
        return self.variance * (1.0 + s5 * r + 5.0 / 3.0 * r ** 2) * np.exp(-s5 * r)


class Sum(StationaryKernel):
    """bayesnewton/kernels.py:1624-1663: block-diagonal state, H = [H_1, H_2]."""

    def __init__(self, kernels):
        self.kernels = list(kernels)
        for k in self.kernels:
            if not isinstance(k, (Matern12, Matern32, Matern52)):
                raise NotImplementedError("Sum supports Matern12/32/52 components")
        self.state_dim = sum(k.state_dim for k in self.kernels)

    @property
    def components(self):
        return self.kernels

    def measurement_model(self):
        return np.concatenate([k.measurement_model() for k in self.kernels], axis=1)

    def K(self, X, X2):
        return sum(k.K(X, X2) for k in self.kernels)

    __call__ = K
