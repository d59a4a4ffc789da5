"""FP64 state-space kernels of the oracle (test infrastructure only).

Restates bayesnewton/kernels.py of the reference: Matern12/32/52 (70-300), the
Sum kernel (1499-1663) and the SpatioTemporalKernel (385-586).  Hyperparameters
are held as plain natural-scale floats; the softplus transform of the reference
(kernels.py:74-95) lives in the model layer.
"""
import numpy as np
import scipy.linalg


def _square_distance(X, X2):
    """bayesnewton/utils.py:580-598 (GPflow form: |x|^2 + |x2|^2 - 2 x.x2)."""
    Xs = np.sum(X ** 2, axis=-1)
    X2s = np.sum(X2 ** 2, axis=-1)
    dist = -2 * np.tensordot(X, X2, [[-1], [-1]])
    dist += Xs[..., None] + X2s[None, ...] if X2s.ndim == 1 else Xs[..., None, None] + X2s
    return dist


class StationaryKernel:
    """bayesnewton/kernels.py:70-120"""
    state_dim = None

    def __init__(self, variance=1.0, lengthscale=1.0):
        self.variance = float(variance)
        self.lengthscale = float(lengthscale)

    def K(self, X, X2):
        X = np.asarray(X, dtype=np.float64)
        X2 = np.asarray(X2, dtype=np.float64)
        if X.ndim < 2:
            X = X[:, None]
        if X2.ndim < 2:
            X2 = X2[:, None]
        r2 = _square_distance(X / self.lengthscale, X2 / self.lengthscale)
        r = np.sqrt(np.maximum(r2, 1e-36))  # kernels.py:101-104
        return self.K_r(r)

    __call__ = K

    def state_transition_batch(self, dt):
        dt = np.asarray(dt, dtype=np.float64).reshape(-1)
        return np.stack([self.state_transition(d) for d in dt])


class Matern12(StationaryKernel):
    """bayesnewton/kernels.py:123-169"""
    state_dim = 1

    def K_r(self, r):
        return self.variance * np.exp(-r)

    def feedback_matrix(self):
        return np.array([[-1.0 / self.lengthscale]])

    def noise_effect(self):
        return np.array([[1.0]]), np.array([[2.0 * self.variance / self.lengthscale]])

    def stationary_covariance(self):
        return np.array([[self.variance]])

    def measurement_model(self):
        return np.array([[1.0]])

    def state_transition(self, dt):
        return np.array([[np.exp(-dt / self.lengthscale)]])


class Matern32(StationaryKernel):
    """bayesnewton/kernels.py:175-230"""
    state_dim = 2

    def K_r(self, r):
        s3 = np.sqrt(3.0)
        return self.variance * (1.0 + s3 * r) * np.exp(-s3 * r)

    def feedback_matrix(self):
        lam = 3.0 ** 0.5 / self.lengthscale
        return np.array([[0.0, 1.0], [-lam ** 2, -2 * lam]])

    def noise_effect(self):
        """L and Qc of kernels.py:195-205"""
        L = np.array([[0.0], [1.0]])
        Qc = np.array([[12.0 * 3.0 ** 0.5 / self.lengthscale ** 3.0 * self.variance]])
        return L, Qc

    def stationary_covariance(self):
        """kernels.py:207-210"""
        return np.array([[self.variance, 0.0],
                         [0.0, 3.0 * self.variance / self.lengthscale ** 2.0]])

    def measurement_model(self):
        return np.array([[1.0, 0.0]])

    def state_transition(self, dt):
        """kernels.py:216-224: closed-form expm(F dt)."""
        lam = np.sqrt(3.0) / self.lengthscale
        return np.exp(-dt * lam) * (dt * np.array([[lam, 1.0], [-lam ** 2.0, -lam]]) + np.eye(2))


class Matern52(StationaryKernel):
    """bayesnewton/kernels.py:233-300"""
    state_dim = 3

    def K_r(self, r):
        s5 = np.sqrt(5.0)
        return self.variance * (1.0 + s5 * r + 5.0 / 3.0 * r ** 2) * np.exp(-s5 * r)

    def feedback_matrix(self):
        lam = 5.0 ** 0.5 / self.lengthscale
        return np.array([[0.0, 1.0, 0.0],
                         [0.0, 0.0, 1.0],
                         [-lam ** 3.0, -3.0 * lam ** 2.0, -3.0 * lam]])

    def noise_effect(self):
        """L and Qc of kernels.py:253-267"""
        L = np.array([[0.0], [0.0], [1.0]])
        Qc = np.array([[self.variance * 400.0 * 5.0 ** 0.5 / 3.0 / self.lengthscale ** 5.0]])
        return L, Qc

    def stationary_covariance(self):
        """kernels.py:288-293"""
        kappa = 5.0 / 3.0 * self.variance / self.lengthscale ** 2.0
        return np.array([[self.variance, 0.0, -kappa],
                         [0.0, kappa, 0.0],
                         [-kappa, 0.0, 25.0 * self.variance / self.lengthscale ** 4.0]])

    def measurement_model(self):
        return np.array([[1.0, 0.0, 0.0]])

    def state_transition(self, dt):
        """kernels.py:273-286: closed-form expm(F dt)."""
        lam = np.sqrt(5.0) / self.lengthscale
        dtlam = dt * lam
        return np.exp(-dtlam) * (
            dt * np.array([[lam * (0.5 * dtlam + 1.0), dtlam + 1.0, 0.5 * dt],
                           [-0.5 * dtlam * lam ** 2, lam * (1.0 - dtlam), 1.0 - 0.5 * dtlam],
                           [lam ** 3 * (0.5 * dtlam - 1.0), lam ** 2 * (dtlam - 3), lam * (0.5 * dtlam - 2.0)]])
            + np.eye(3))


class Sum:
    """bayesnewton/kernels.py:1624-1663 (block-diagonal state, concatenated H)."""

    def __init__(self, kernels):
        self.kernels = list(kernels)
        self.state_dim = sum(k.state_dim for k in self.kernels)

    def K(self, X, X2):
        return sum(k.K(X, X2) for k in self.kernels)

    __call__ = K

    def feedback_matrix(self):
        return scipy.linalg.block_diag(*[k.feedback_matrix() for k in self.kernels])

    def noise_effect(self):
        Ls, Qs = zip(*[k.noise_effect() for k in self.kernels])
        return scipy.linalg.block_diag(*Ls), scipy.linalg.block_diag(*Qs)

    def stationary_covariance(self):
        return scipy.linalg.block_diag(*[k.stationary_covariance() for k in self.kernels])

    def measurement_model(self):
        return np.concatenate([k.measurement_model() for k in self.kernels], axis=1)

    def state_transition(self, dt):
        return scipy.linalg.block_diag(*[k.state_transition(dt) for k in self.kernels])

    def state_transition_batch(self, dt):
        dt = np.asarray(dt, dtype=np.float64).reshape(-1)
        return np.stack([self.state_transition(d) for d in dt])


class SpatioTemporalKernel:
    """bayesnewton/kernels.py:385-586 (sparse=True, conditional='Full' - the default when sparse)."""

    def __init__(self, temporal_kernel, spatial_kernel, z, sparse=True, conditional="Full"):
        self.temporal_kernel = temporal_kernel
        self.spatial_kernel = spatial_kernel
        z = np.asarray(z, dtype=np.float64)
        if z.ndim < 2:
            z = z[:, None]
        self.z = z
        self.M = z.shape[0]
        self.sparse = sparse
        if conditional not in ("Full", "full", "DTC", "dtc"):
            raise NotImplementedError(conditional)
        self.conditional = "Full" if conditional in ("Full", "full") else "DTC"
        if not sparse:
            self.conditional = "DTC"
        self.state_dim = self.M * temporal_kernel.state_dim

    def K(self, X, X2):
        """kernels.py:469-474"""
        return self.temporal_kernel(X[:, :1], X2[:, :1]) * self.spatial_kernel(X[:, 1:], X2[:, 1:])

    __call__ = K

    def inducing_precision(self):
        """kernels.py:508-515: Qzz = Kzz^-1 and Lzz = chol(Kzz)."""
        Kzz = self.spatial_kernel(self.z, self.z)
        Lzz = np.linalg.cholesky(Kzz)
        Qzz = scipy.linalg.cho_solve((Lzz, True), np.eye(self.M))
        return Qzz, Lzz

    def spatial_conditional(self, X, R, predict=False):
        """kernels.py:486-506: B = Krz Kzz^-1 Lzz, C = k_t(0) (Krr - Krz Kzz^-1 Kzr)."""
        Qzz, Lzz = self.inducing_precision()
        if self.sparse or predict:
            R = np.asarray(R, dtype=np.float64)
            R = R.reshape((R.shape[0], -1) + self.z.shape[1:])
            Krz = np.stack([self.spatial_kernel(Rt, self.z) for Rt in R])
            K = Krz @ Qzz
            B = K @ Lzz
            if self.conditional == "Full":
                kt0 = np.array([self.temporal_kernel(np.atleast_2d(x[:1]), np.atleast_2d(x[:1]))[0, 0]
                                for x in np.asarray(X).reshape(R.shape[0], -1)])
                Krr = np.stack([self.spatial_kernel(Rt, Rt) for Rt in R])
                C = kt0[:, None, None] * (Krr - K @ np.swapaxes(Krz, -1, -2))
            else:
                C = np.zeros((R.shape[0], 1, 1))
        else:
            B = Lzz
            C = np.zeros([B.shape[0], B.shape[0]])
        return B, C

    def stationary_covariance(self):
        """kernels.py:517-524"""
        return np.kron(np.eye(self.M), self.temporal_kernel.stationary_covariance())

    def measurement_model(self):
        """kernels.py:534-541"""
        return np.kron(np.eye(self.M), self.temporal_kernel.measurement_model())

    def state_transition(self, dt):
        """kernels.py:543-551"""
        return np.kron(np.eye(self.M), self.temporal_kernel.state_transition(dt))

    def state_transition_batch(self, dt):
        dt = np.asarray(dt, dtype=np.float64).reshape(-1)
        return np.stack([self.state_transition(d) for d in dt])

    def feedback_matrix(self):
        return np.kron(np.eye(self.M), self.temporal_kernel.feedback_matrix())
