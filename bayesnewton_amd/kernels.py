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

    shared_with = None  # kernel whose lengthscale this one uses (kernels.py:712-714)

    @property
    def lengthscale(self):
        if self.shared_with is not None:
            return self.shared_with.lengthscale
        return softplus(self.transformed_lengthscale)

    @lengthscale.setter
    def lengthscale(self, v):
        if self.shared_with is not None:
            self.shared_with.lengthscale = v
        else:
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


class SpatioTemporalKernel:
    """bayesnewton/kernels.py:385-586: k((t, r), (t', r')) = k_t(t, t') k_s(r, r') in
    state-space form over M spatial inducing points z, sparse in space.

    The conditional of the data given the inducing states (kernels.py:486-506) is
    f(t, R) | u(t) ~ N(B u(t), C) with B = Krz Kzz^-1 Lzz and
    C = k_t(0) (Krr - Krz Kzz^-1 Kzr).  Only diag(C) reaches the factorised
    likelihoods (likelihoods.py:363-383), so 'Full' and 'FIC' ('FITC') give the
    same model here; 'DTC' drops C.  The temporal kernel is Matern12 or Matern32
    (state dim 1 or 2), the spatial kernel Matern12/32/52 on 2-D locations.
    """

    def __init__(self, temporal_kernel, spatial_kernel, z=None, conditional=None, sparse=True, opt_z=False,
                 spatial_dims=None):
        if not isinstance(temporal_kernel, (Matern12, Matern32)):
            raise NotImplementedError("the space-time HIP path supports Matern12 / Matern32 temporal kernels")
        if not isinstance(spatial_kernel, (Matern12, Matern32, Matern52)):
            raise NotImplementedError("spatial kernel must be Matern12, Matern32 or Matern52")
        if opt_z:
            raise NotImplementedError("optimising the inducing inputs z is not implemented on the HIP path")
        if conditional is None:
            conditional = "Full" if sparse else "DTC"
        if conditional not in ("Full", "full", "FIC", "FITC", "fic", "fitc", "DTC", "dtc"):
            raise NotImplementedError("conditional method not recognised")
        self.temporal_kernel = temporal_kernel
        self.spatial_kernel = spatial_kernel
        if z is None:
            # kernels.py:418-428
            if spatial_dims == 1:
                raise NotImplementedError("1-D space: the HIP path takes 2-D locations")
            g = np.linspace(-3.0, 3.0, 5)
            a, b = np.meshgrid(g, g)
            z = np.stack([a.reshape(-1), b.reshape(-1)], axis=1)
        z = np.asarray(z, dtype=np.float64)
        if z.ndim != 2 or z.shape[1] != 2:
            raise NotImplementedError("spatial inducing inputs z must be [M, 2]")
        self.z = z
        self.M = z.shape[0]
        self.sparse = sparse
        # not sparse: B = Lzz and C = 0 (kernels.py:502-505), which the sparse
        # formulas reproduce when the locations are z itself; DTC drops C in both
        self.conditional = "DTC" if (not sparse or conditional in ("DTC", "dtc")) else "Full"

    @property
    def variance(self):
        return self.temporal_kernel.variance

    @property
    def temporal_lengthscale(self):
        return self.temporal_kernel.lengthscale

    @property
    def spatial_lengthscale(self):
        return self.spatial_kernel.lengthscale

    @property
    def state_dim(self):
        return self.temporal_kernel.state_dim * self.M

    def K(self, X, X2):
        """kernels.py:469-474"""
        X = np.asarray(X, np.float64)
        X2 = np.asarray(X2, np.float64)
        return self.temporal_kernel(X[:, :1], X2[:, :1]) * self.spatial_kernel(X[:, 1:], X2[:, 1:])

    __call__ = K

    def inducing_precision(self):
        """kernels.py:508-515: (Kzz^-1, chol(Kzz)), FP64 on the host (M x M)."""
        Kzz = self.spatial_kernel(self.z, self.z)
        Lzz = np.linalg.cholesky(Kzz)
        Qzz = np.linalg.solve(Kzz, np.eye(self.M))
        return Qzz, Lzz


class SpatialMatern12(SpatioTemporalKernel):
    """kernels.py:664-688: Matern12 in time and space with one shared lengthscale."""

    def __init__(self, variance=1.0, lengthscale=1.0, z=None, sparse=True, opt_z=False, conditional=None):
        super().__init__(Matern12(variance=variance, lengthscale=lengthscale),
                         Matern12(variance=1.0, lengthscale=lengthscale, fix_variance=True), z=z,
                         conditional=conditional, sparse=sparse, opt_z=opt_z)
        self.spatial_kernel.shared_with = self.temporal_kernel


class SpatialMatern32(SpatioTemporalKernel):
    """kernels.py:691-715: Matern32 in time and space with one shared lengthscale."""

    def __init__(self, variance=1.0, lengthscale=1.0, z=None, sparse=True, opt_z=False, conditional=None):
        super().__init__(Matern32(variance=variance, lengthscale=lengthscale),
                         Matern32(variance=1.0, lengthscale=lengthscale, fix_variance=True), z=z,
                         conditional=conditional, sparse=sparse, opt_z=opt_z)
        self.spatial_kernel.shared_with = self.temporal_kernel
