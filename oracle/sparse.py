"""FP64 restatement of the reference's sparse Markov GP (test infrastructure only).

bayesnewton/basemodels.py:928-1152 (SparseMarkovGaussianProcess) with the inference
mixins of bayesnewton/inference.py:21-222 (newton_update, InferenceMixin.inference,
Newton, VariationalInference) and the helpers of bayesnewton/utils.py:185-215
(compute_conditional_statistics), 218-224 (sum_natural_params_by_group), 544-559
(build_joint, set_z_stats); models.py:135-157, 302-310, 330 bind the mixins.

The inducing states u_0..u_{M+1} sit at the sorted Z with sentinels at -/+1e10; the
sites are 2d-dimensional Gaussians on the pairs (u_i, u_{i+1}) of each of the M + 1
transitions, filtered by the pairs filter (ops.py:383-426, oracle.ops.kalman_filter_pairs)
and smoothed by the RTS smoother over the inducing states.  Every data point n between
Z_i and Z_{i+1} sees f_n = H x(t_n), x(t_n) | (u_i, u_{i+1}) ~ N(P_n [u_i; u_{i+1}], T_n).
"""
import numpy as np

from . import ops
from .models import compute_conditional_statistics
from .utils import gaussian_expected_log_lik, input_admin, inv


def build_joint(i, mean, cov, gain):
    """utils.py:544-553: joint smoothed moments of (u_i, u_{i+1})."""
    mj = np.concatenate([mean[i], mean[i + 1]])
    cross = gain[i] @ cov[i + 1]
    cj = np.block([[cov[i], cross], [cross.T, cov[i + 1]]])
    return mj, cj


def set_z_stats(t, z):
    """utils.py:556-559"""
    ind = np.searchsorted(z.reshape(-1), t[:, :1].reshape(-1)) - 1
    num_neighbours = np.array([np.sum(ind == m) for m in range(z.shape[0] - 1)])
    return ind, num_neighbours


class SparseMarkovGP:
    """basemodels.py:928-1152 with inference 'VI' (SparseMarkovVariationalGP) or
    'Newton' (SparseMarkovNewtonGP = SparseMarkovLaplaceGP)."""

    def __init__(self, kernel, likelihood, X, Y, Z, inference="VI", gh_points=20):
        X, Y, _, _ = input_admin(X, Y)
        self.X, self.Y = X, Y
        self.kernel, self.likelihood = kernel, likelihood
        self.inference_kind = inference
        self.gh_points = gh_points
        self.num_data = X.shape[0]
        d = kernel.stationary_covariance().shape[0]
        self.state_dim = d
        Z = np.sort(np.asarray(Z, np.float64).reshape(-1, 1), axis=0)
        inf = np.array([[1e10]])
        self.Z = np.concatenate([-inf, Z, inf], axis=0)
        self.dz = np.diff(self.Z[:, 0])
        self.num_transitions = self.dz.shape[0]
        zeros = np.zeros([self.num_transitions, 2 * d, 1])
        eyes = np.tile(np.eye(2 * d), [self.num_transitions, 1, 1])
        # basemodels.py:953-955: match the Markov GP's first step when Z = X
        nat2 = 1e-8 * eyes
        nat2[:-1, d, d] = 1e-2
        self.nat1 = zeros.copy()
        self.nat2 = nat2
        self.site_cov = inv(nat2)
        self.site_mean = self.site_cov @ self.nat1
        self.posterior_mean = zeros.copy()
        self.posterior_variance = eyes.copy()
        self.ind, self.num_neighbours = set_z_stats(self.X, self.Z)
        self.minf = np.zeros([d, 1])
        self.last_filter_loglik = None

    def update_nat_params(self, nat1, nat2):
        """basemodels.py:97-100"""
        self.nat1, self.nat2 = nat1, nat2
        self.site_cov = inv(nat2)
        self.site_mean = self.site_cov @ nat1

    # ---- posterior -------------------------------------------------------------------
    def _smoothed_states(self):
        ell, (fm, fP) = ops.kalman_filter_pairs(self.dz, self.kernel, self.site_mean, self.site_cov)
        self.last_filter_loglik = ell
        sm, sP, gain = ops.rauch_tung_striebel_smoother(self.dz[1:], self.kernel, fm, fP, return_full=True)
        return ell, sm, sP, gain

    def update_posterior(self):
        """basemodels.py:980-1005"""
        _, sm, sP, gain = self._smoothed_states()
        Pinf = self.kernel.stationary_covariance()[None]
        mean_aug = np.concatenate([self.minf[None], sm, self.minf[None]])
        cov_aug = np.concatenate([Pinf, sP, Pinf])
        gain = np.concatenate([np.zeros_like(gain[:1]), gain])
        pm, pc = zip(*[build_joint(i, mean_aug, cov_aug, gain) for i in range(self.num_transitions)])
        self.posterior_mean, self.posterior_variance = np.stack(pm), np.stack(pc)

    def compute_log_lik(self):
        """basemodels.py:1007-1016"""
        ell, _ = ops.kalman_filter_pairs(self.dz, self.kernel, self.site_mean, self.site_cov)
        return ell

    def expected_density_pseudo(self):
        """basemodels.py:214-222 (mask_pseudo_y is None for the sparse model)"""
        return np.sum(gaussian_expected_log_lik(self.site_mean, self.posterior_mean, self.posterior_variance,
                                                self.site_cov))

    def compute_kl(self):
        """basemodels.py:1018-1030"""
        return self.expected_density_pseudo() - self.compute_log_lik()

    # ---- data conditional ------------------------------------------------------------
    def conditional_posterior_to_data(self, post_mean=None, post_cov=None):
        """basemodels.py:1074-1105 (temporal inputs): mean_f = W q_mean, cov_f = W q_cov W' + nu."""
        if post_mean is None:
            post_mean = self.posterior_mean
        if post_cov is None:
            post_cov = self.posterior_variance
        ind = self.ind
        pm, pc = post_mean[ind], post_cov[ind]
        P, Tn = compute_conditional_statistics(self.X[:, :1], self.Z, self.kernel, ind)
        H = self.kernel.measurement_model()
        self.conditional_mean = H @ P
        nu = H @ Tn @ H.T
        mean_f = self.conditional_mean @ pm
        cov_f = self.conditional_mean @ pc @ np.swapaxes(self.conditional_mean, -1, -2) + nu
        return mean_f, cov_f

    def conditional_data_to_posterior(self, mean_f, cov_f):
        """basemodels.py:1107-1113"""
        W = self.conditional_mean
        return np.swapaxes(W, -1, -2) @ mean_f, np.swapaxes(W, -1, -2) @ cov_f @ W

    def group_natural_params(self, nat1_n, nat2_n):
        """basemodels.py:1115-1140: sum the data points' pair sites per transition; a
        transition without data keeps its old site; 1e-8 I keeps nat2 away from zero."""
        new1 = np.zeros_like(self.nat1)
        new2 = np.zeros_like(self.nat2)
        counter = np.zeros(self.nat1.shape[0])
        for i, n1, n2 in zip(self.ind, nat1_n, nat2_n):
            new1[i] += n1
            new2[i] += n2
            counter[i] += 1.0
        nn = np.maximum(self.num_neighbours, 1).reshape(-1, 1, 1)
        c = counter.reshape(-1, 1, 1)
        nat1 = new1 + (1.0 - c / nn) * self.nat1
        nat2 = new2 + (1.0 - c / nn) * self.nat2
        nat2 = nat2 + 1e-8 * np.eye(nat2.shape[1])
        return nat1, nat2

    # ---- inference (inference.py:65-90) -----------------------------------------------
    def update_variational_params(self, ensure_psd=True):
        mean_f, cov_f = self.conditional_posterior_to_data()
        y = self.Y.reshape(-1)
        m = mean_f.reshape(-1)
        if self.inference_kind == "VI":
            # inference.py:170-195
            _, d1, d2 = self.likelihood.variational_expectation(y, m, cov_f.reshape(-1), self.gh_points)
        elif self.inference_kind == "Newton":
            # inference.py:105-128
            _, d1, d2 = self.likelihood.log_likelihood_gradients(y, m)
        else:
            raise NotImplementedError(self.inference_kind)
        d1 = d1.reshape(-1, 1, 1)
        d2 = d2.reshape(-1, 1, 1)
        if ensure_psd:
            d2 = np.where(np.isnan(d2), d2, -np.where(-d2 < 0, 1e-2, -d2))  # utils.py:89-96, scalar sites
        jac, hess = self.conditional_data_to_posterior(d1, d2)
        return self.posterior_mean[self.ind], jac, hess  # sparse Markov case (inference.py:190-195)

    def inference(self, lr=1.0, ensure_psd=True):
        """inference.py:65-90 with group_natural_params."""
        self.update_posterior()
        mean, jac, hess = self.update_variational_params(ensure_psd)
        # newton_update (inference.py:21-39), missing data included
        hess = np.where(np.isnan(hess), -1e-6, hess)
        jac = np.where(np.isnan(jac), hess @ mean, jac)
        nat1_n = jac - hess @ mean
        nat2_n = -hess
        nat1, nat2 = self.group_natural_params(nat1_n, nat2_n)
        diff1 = np.mean(np.abs(nat1 - self.nat1))
        diff2 = np.mean(np.abs(nat2 - self.nat2))
        self.update_nat_params((1 - lr) * self.nat1 + lr * nat1, (1 - lr) * self.nat2 + lr * nat2)
        self.update_posterior()
        return (mean, jac, hess), (diff1, diff2)

    def energy(self):
        """VI: inference.py:197-222; Newton: inference.py:130-154."""
        mean_f, cov_f = self.conditional_posterior_to_data()
        y = self.Y.reshape(-1)
        if self.inference_kind == "VI":
            ell, _, _ = self.likelihood.variational_expectation(y, mean_f.reshape(-1), cov_f.reshape(-1),
                                                                self.gh_points)
        else:
            ell, _, _ = self.likelihood.log_likelihood_gradients(y, mean_f.reshape(-1))
        return -(np.nansum(ell) - self.compute_kl())

    def predict(self, X):
        """basemodels.py:1032-1072 (temporal inputs): the smoothed inducing states
        conditioned at the test times (utils.py:125-141)."""
        from .models import temporal_conditional
        _, sm, sP, gain = self._smoothed_states()
        X = np.asarray(X, np.float64).reshape(-1, 1)
        m, P = temporal_conditional(self.Z, X, sm, sP, gain, self.kernel)  # Z holds the sentinels
        H = self.kernel.measurement_model()
        return np.squeeze(H @ m), np.squeeze(H @ P @ H.T)
