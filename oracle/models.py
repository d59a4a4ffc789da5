"""FP64 Markov-GP models of the oracle (test infrastructure only).

Restates the reference's model/inference stack for the Markov GP:
bayesnewton/basemodels.py:52-262 (sites, BaseModel) and 625-764
(MarkovGaussianProcess) with prediction 766-816 and 165-193 (utils.py:99-215),
its mean-field form 1158-1175 (space-time kernels), the infinite-horizon model 1257-1301,
bayesnewton/inference.py:21-325 (newton_update,
InferenceMixin.inference, Newton/Laplace, VariationalInference,
ExpectationPropagation) and bayesnewton/models.py:118-132, 216-231, 297-299.

`DenseGP` restates the kernel-matrix GP (basemodels.py:265-330 with
ops.py:52-80) and is used only to pin the Markov oracle.
"""
import numpy as np

from . import ops
from .utils import (inv, solve, diag, input_admin, mvn_logpdf, mvn_logpdf_and_derivs, gaussian_expected_log_lik,
                    compute_cavity, pep_constant, ensure_diagonal_positive_precision, LOG2PI)
from .kernels import SpatioTemporalKernel


def T(P):
    return np.swapaxes(P, -1, -2)


def reparametrise(param1, param2):
    """basemodels.py:85-90: (mean, cov) <-> (nat1, nat2) with nat2 = cov^-1 (no -1/2)."""
    return solve(param2, param1), inv(param2)


def newton_update(mean, jacobian, hessian):
    """inference.py:21-39"""
    hessian = np.where(np.isnan(hessian), -1e-6, hessian)
    jacobian = np.where(np.isnan(jacobian), hessian @ mean, jacobian)
    return jacobian - hessian @ mean, -hessian


def _diag_mat(v):
    """[N, p] -> [N, p, p] diagonal."""
    out = np.zeros(v.shape + (v.shape[-1],))
    idx = np.arange(v.shape[-1])
    out[..., idx, idx] = v
    return out


def compute_conditional_statistics(x_test, x, kernel, ind):
    """utils.py:185-215: p(x_n | u-, u+) = N(P_n [u-; u+], T_n) for every test time."""
    n = x_test.shape[0]
    Pinf = kernel.stationary_covariance()
    d = Pinf.shape[0]
    P = np.empty((n, d, 2 * d))
    Tn = np.empty((n, d, d))
    for i in range(n):
        A_fwd = kernel.state_transition(x_test[i, 0] - x[ind[i], 0])
        A_back = kernel.state_transition(x[ind[i] + 1, 0] - x_test[i, 0])
        Q_fwd = Pinf - A_fwd @ Pinf @ A_fwd.T
        Q_back = Pinf - A_back @ Pinf @ A_back.T
        A_back_Q_fwd = A_back @ Q_fwd
        Q_mp = Q_back + A_back @ A_back_Q_fwd.T
        L = np.linalg.cholesky(Q_mp + 1e-8 * np.eye(d))
        V = np.linalg.solve(L.T, np.linalg.solve(L, A_back))
        Tn[i] = Q_fwd - A_back_Q_fwd.T @ V @ Q_fwd
        W = Q_fwd @ V.T
        P[i] = np.concatenate([A_fwd - W @ A_back @ A_fwd, W], axis=-1)
    return P, Tn


def temporal_conditional(X, X_test, mean, cov, gain, kernel):
    """utils.py:125-141 with predict_from_state_ (utils.py:110-122)."""
    Pinf = kernel.stationary_covariance()[None, ...]
    minf = np.zeros([1, Pinf.shape[1], 1])
    mean_aug = np.concatenate([minf, mean, minf])
    cov_aug = np.concatenate([Pinf, cov, Pinf])
    gain = np.concatenate([np.zeros_like(gain[:1]), gain])
    ind = np.searchsorted(X.reshape(-1), X_test.reshape(-1)) - 1
    P, Tn = compute_conditional_statistics(X_test, X, kernel, ind)
    d = Pinf.shape[1]
    out_m = np.empty((X_test.shape[0], d, 1))
    out_P = np.empty((X_test.shape[0], d, d))
    for i, k in enumerate(ind):
        mean_joint = np.concatenate([mean_aug[k], mean_aug[k + 1]])
        cross = gain[k] @ cov_aug[k + 1]
        cov_joint = np.block([[cov_aug[k], cross], [cross.T, cov_aug[k + 1]]])
        out_m[i] = P[i] @ mean_joint
        out_P[i] = P[i] @ cov_joint @ P[i].T + Tn[i]
    return out_m, out_P


class MarkovGP:
    """basemodels.py:625-764 with the inference mixins of inference.py.

    inference: 'VI' (MarkovVariationalGP), 'Newton' (MarkovLaplaceGP /
    MarkovNewtonGP) or 'EP' (MarkovExpectationPropagationGP).
    """

    def __init__(self, kernel, likelihood, X, Y, R=None, parallel=False, inference="VI", power=1.0,
                 gh_points=20, mean_field=False, infinite_horizon=False, dare_iters=20):
        X, Y, R, dt = input_admin(X, Y, R)
        # InfiniteHorizonGaussianProcess (basemodels.py:1257-1301): equispaced steps, the
        # infinite-horizon filter / smoother, each Riccati solve started from the previous one
        self.infinite_horizon = bool(infinite_horizon)
        self.dare_iters = int(dare_iters)
        # MarkovMeanFieldGaussianProcess (basemodels.py:1158-1175) swaps in the
        # mean-field filter and smoother (ops.py:581-608, 681-706); nothing else changes.
        self.mean_field = bool(mean_field)
        if self.mean_field and not isinstance(kernel, SpatioTemporalKernel):
            raise NotImplementedError("the mean-field model is restated for space-time kernels only")
        self.X, self.Y, self.R, self.dt = X, Y, R, dt
        self.kernel = kernel
        self.likelihood = likelihood
        self.parallel = parallel
        self.inference_kind = inference
        self.power = float(power)
        self.gh_points = gh_points
        H = kernel.measurement_model()
        self.func_dim = H.shape[0]
        self.num_data = X.shape[0]
        self.obs_dim = Y.shape[1]
        p = self.obs_dim
        # basemodels.py:130-133: sites N(0, 1e2 I)
        self.site_mean = np.zeros([self.num_data, p, 1])
        self.site_cov = 1e2 * np.tile(np.eye(p), [self.num_data, 1, 1])
        self.nat1, self.nat2 = reparametrise(self.site_mean, self.site_cov)
        self.posterior_mean = np.zeros([self.num_data, self.func_dim, 1])
        self.posterior_variance = np.tile(np.eye(self.func_dim), [self.num_data, 1, 1])
        self.mask_y = np.isnan(self.Y).reshape(Y.shape[0], Y.shape[1])
        self.spatio_temporal = bool(np.any(~np.isnan(self.R)))
        if self.func_dim == self.obs_dim:
            self.mask_pseudo_y = self.mask_y
        else:
            self.mask_pseudo_y = np.tile(self.mask_y, [1, p])
        if self.func_dim != self.obs_dim and self.spatio_temporal:
            self.mask_pseudo_y = None
        self.last_filter_loglik = None
        if self.infinite_horizon:
            if np.max(np.abs(np.diff(self.dt[1:]))) >= 1e-6:
                raise ValueError("the infinite-horizon model needs equispaced time steps")
            self.heteroscedastic = bool(np.any(np.isnan(self.Y))) or type(likelihood).__name__ != "Gaussian"
            Pinf = kernel.stationary_covariance()
            self.dare_init_filter = Pinf
            self.dare_init_smoother = Pinf

    # ---- sites ----------------------------------------------------------------
    def update_nat_params(self, nat1, nat2):
        """basemodels.py:97-100"""
        self.nat1, self.nat2 = nat1, nat2
        self.site_mean, self.site_cov = reparametrise(nat1, nat2)

    def _mask3(self):
        if self.mask_pseudo_y is None:
            return None
        return self.mask_pseudo_y.reshape(self.num_data, -1, 1)

    # ---- posterior ------------------------------------------------------------
    def spatial_conditional(self):
        return self.kernel.spatial_conditional(self.X, self.R)

    def compute_full_pseudo_lik(self):
        """basemodels.py:676-687"""
        if self.spatio_temporal:
            B, C = self.spatial_conditional()
            nat1f = T(B) @ self.nat1
            nat2f = T(B) @ self.nat2 @ B
            Mi = nat2f.shape[-1]
            pvar = inv(nat2f + 1e-12 * np.eye(Mi))
            return pvar @ nat1f, pvar
        return self.site_mean, self.site_cov

    def compute_full_pseudo_nat(self):
        """basemodels.py:667-674"""
        if self.spatio_temporal:
            B, C = self.spatial_conditional()
            return T(B) @ self.nat1, T(B) @ self.nat2 @ B
        return self.nat1, self.nat2

    def _filter(self, *args, **kwargs):
        if self.infinite_horizon:
            # basemodels.py:1276-1290
            if self.spatio_temporal:
                B, C = self.spatial_conditional()
                tied = np.mean(T(B) @ self.nat2 @ B, axis=0)
            else:
                tied = np.mean(self.nat2, axis=0)
            out = ops.kalman_filter_infinite_horizon(*args, **kwargs, heteroscedastic=self.heteroscedastic,
                                                     noise_cov_tied=inv(tied), dare_iters=self.dare_iters,
                                                     dare_init=self.dare_init_filter)
            self.dare_init_filter = out[1][1][0]
            return out
        if self.mean_field:
            return ops.kalman_filter_meanfield(*args, **kwargs)
        return ops.kalman_filter(*args, **kwargs)

    def _smoother(self, *args, **kwargs):
        if self.infinite_horizon:
            # basemodels.py:1292-1298
            means, covs, gains, dare_cov = ops.rauch_tung_striebel_smoother_infinite_horizon(
                *args, **kwargs, dare_iters=self.dare_iters, dare_init=self.dare_init_smoother)
            self.dare_init_smoother = dare_cov
            return means, covs, gains
        if self.mean_field:
            return ops.rauch_tung_striebel_smoother_meanfield(*args, **kwargs)
        return ops.rauch_tung_striebel_smoother(*args, **kwargs)

    def update_posterior(self):
        """basemodels.py:689-706"""
        pseudo_y, pseudo_var = self.compute_full_pseudo_lik()
        ell, (fm, fP) = self._filter(self.dt, self.kernel, pseudo_y, pseudo_var,
                                     mask=self._mask3(), parallel=self.parallel)
        self.last_filter_loglik = ell
        dt = np.concatenate([self.dt[1:], np.array([0.0])])
        sm, sP, _ = self._smoother(dt, self.kernel, fm, fP, parallel=self.parallel)
        self.posterior_mean, self.posterior_variance = sm, sP

    # ---- prediction (basemodels.py:766-816, 165-193) --------------------------
    def predict(self, X=None, R=None):
        """basemodels.py:766-816: filter, full smoother with gains, then the conditional of
        every test time on its two neighbouring states.  The mean-field model filters with
        the sequential form whatever `parallel` is (basemodels.py:1177-1228); the
        infinite-horizon model uses the smoother's steady covariance at every test time
        (temporal_conditional_infinite_horizon, utils.py:142-177) and advances both Riccati
        solutions, as its filter and smoother calls do."""
        X = self.X if X is None else np.asarray(X, np.float64).reshape(len(X), -1)
        if self.spatio_temporal and R is None:
            R = self.R
        X = X[:, :1]
        pseudo_y, pseudo_var = self.compute_full_pseudo_lik()
        _, (fm, fP) = self._filter(self.dt, self.kernel, pseudo_y, pseudo_var, mask=self._mask3(),
                                   parallel=False if self.mean_field else self.parallel)
        dt = np.concatenate([self.dt[1:], np.array([0.0])])
        sm, sP, gain = self._smoother(dt, self.kernel, fm, fP, return_full=True, parallel=self.parallel)
        inf = 1e10 * np.ones_like(self.X[0, :1])
        X_aug = np.block([[-inf], [self.X[:, :1]], [inf]])
        if self.infinite_horizon:
            mean_aug = np.concatenate([np.zeros((1,) + sm.shape[1:]), sm, np.zeros((1,) + sm.shape[1:])])
            ind = np.searchsorted(X_aug.reshape(-1), X.reshape(-1)) - 1
            P, _ = compute_conditional_statistics(X, X_aug, self.kernel, ind)
            state_mean = np.stack([P[i] @ np.concatenate([mean_aug[k], mean_aug[k + 1]]) for i, k in enumerate(ind)])
            state_cov = np.broadcast_to(sP[0], (X.shape[0],) + sP.shape[1:])
        else:
            state_mean, state_cov = temporal_conditional(X_aug, X, sm, sP, gain, self.kernel)
        H = self.kernel.measurement_model()
        if self.spatio_temporal:
            # basemodels.py:805-812: f(t, R) = B H x(t) + N(0, C); the spatial covariance
            # is reduced to its diagonal
            B, C = self.kernel.spatial_conditional(X, R, predict=True)
            W = B @ H
            test_mean = W @ state_mean
            test_var = np.diagonal(W @ state_cov @ T(W) + C, axis1=1, axis2=2)
            return np.squeeze(test_mean), np.squeeze(test_var)
        test_mean, test_var = H @ state_mean, H @ state_cov @ T(H)
        return np.squeeze(test_mean), np.squeeze(test_var)

    def predict_y(self, X, R=None):
        """basemodels.py:165-175"""
        mean_f, var_f = self.predict(X, R)
        return self.likelihood.predict(np.atleast_1d(mean_f), np.atleast_1d(var_f), self.gh_points)

    def negative_log_predictive_density(self, X, Y, R=None):
        """basemodels.py:177-193"""
        mean_f, var_f = self.predict(X, R)
        ld = self.likelihood.log_density(np.asarray(Y, np.float64).reshape(-1), np.atleast_1d(mean_f),
                                         np.atleast_1d(var_f), self.gh_points)
        return -np.nanmean(ld)

    def compute_log_lik(self, pseudo_y=None, pseudo_var=None):
        """basemodels.py:726-741"""
        if pseudo_y is None:
            pseudo_y, pseudo_var = self.compute_full_pseudo_lik()
        ell, _ = self._filter(self.dt, self.kernel, pseudo_y, pseudo_var,
                              mask=self._mask3(), parallel=self.parallel)
        return ell

    def compute_kl(self):
        """basemodels.py:708-724"""
        pseudo_y, pseudo_var = self.compute_full_pseudo_lik()
        log_lik_pseudo = self.compute_log_lik(pseudo_y, pseudo_var)
        ed = gaussian_expected_log_lik(pseudo_y, self.posterior_mean, self.posterior_variance,
                                       pseudo_var, self._mask3())
        return np.sum(ed) - log_lik_pseudo

    def conditional_posterior_to_data(self, post_mean=None, post_cov=None):
        """basemodels.py:743-764"""
        if post_mean is None:
            post_mean = self.posterior_mean
        if post_cov is None:
            post_cov = self.posterior_variance
        if self.spatio_temporal:
            B, C = self.spatial_conditional()
            return B @ post_mean, B @ post_cov @ T(B) + C
        return post_mean, post_cov

    # ---- site maps ------------------------------------------------------------
    def _variational_expectation(self, mean_f, cov_f):
        """likelihoods.py:363-383 vmapped over time (inference.py:179-184)."""
        Nn, p = mean_f.shape[0], mean_f.shape[1]
        y = self.Y.reshape(-1)
        m = mean_f.reshape(-1)
        v = diag(cov_f).reshape(-1)
        E, dE, d2E = self.likelihood.variational_expectation(y, m, v, self.gh_points)
        return E.reshape(Nn, p), dE.reshape(Nn, p, 1), _diag_mat(d2E.reshape(Nn, p))

    def _log_likelihood_gradients(self, mean_f):
        """likelihoods.py:337-355 vmapped over time (inference.py:113-117)."""
        Nn, p = mean_f.shape[0], mean_f.shape[1]
        ll, J, H = self.likelihood.log_likelihood_gradients(self.Y.reshape(-1), mean_f.reshape(-1))
        return ll.reshape(Nn, p), J.reshape(Nn, p, 1), _diag_mat(H.reshape(Nn, p))

    def cavity_distribution(self, power):
        """basemodels.py:230-245"""
        nat1f, nat2f = self.compute_full_pseudo_nat()
        return compute_cavity(self.posterior_mean, self.posterior_variance, nat1f, nat2f, power)

    def _moment_match(self, cav_mean_f, cav_cov_f):
        """likelihoods.py:401-412 vmapped over time.  Several locations per step (space-time
        kernels): the multivariate closed form of the Gaussian likelihood
        (likelihoods.py:755-782, mvn_logpdf_and_derivs with sigma^2 I / power + cavity
        covariance, plus pep_constant(sigma^2 I, power)), missing outputs replaced by the
        cavity mean (likelihoods.py:401-412); other likelihoods integrate over all the
        step's locations at once with a product Gauss-Hermite rule (cubature.py:310-370),
        which is not restated."""
        if cav_mean_f.shape[1] != 1:
            if type(self.likelihood).__name__ != "Gaussian":
                raise NotImplementedError("multivariate EP moment matching is restated for the Gaussian "
                                          "likelihood only")
            Nn, p = cav_mean_f.shape[0], cav_mean_f.shape[1]
            y = self.Y.reshape(Nn, p, 1)
            y = np.where(np.isnan(y), cav_mean_f, y)
            s2 = self.likelihood.variance
            lik_cov = s2 * np.eye(p)
            lZ, dlZ, d2lZ = mvn_logpdf_and_derivs(y, cav_mean_f, lik_cov / self.power + cav_cov_f)
            lZ = lZ + pep_constant(np.broadcast_to(lik_cov, (Nn, p, p)), self.power)
            return lZ, dlZ, d2lZ
        y = self.Y.reshape(-1)
        m = cav_mean_f.reshape(-1)
        v = cav_cov_f.reshape(-1)
        lZ, dlZ, d2lZ = self.likelihood.moment_match(y, m, v, self.power, self.gh_points)
        Nn = cav_mean_f.shape[0]
        return lZ.reshape(Nn), dlZ.reshape(Nn, 1, 1), d2lZ.reshape(Nn, 1, 1)

    def update_variational_params(self, ensure_psd=True):
        kind = self.inference_kind
        if kind == "VI":
            # inference.py:170-195
            mean_f, cov_f = self.conditional_posterior_to_data()
            _, dE, d2E = self._variational_expectation(mean_f, cov_f)
            if ensure_psd:
                d2E = -ensure_diagonal_positive_precision(-d2E)
            return mean_f, dE, d2E
        if kind == "Newton":
            # inference.py:105-128
            mean_f, _ = self.conditional_posterior_to_data()
            _, J, H = self._log_likelihood_gradients(mean_f)
            if ensure_psd:
                H = -ensure_diagonal_positive_precision(-H)
            return mean_f, J, H
        if kind == "EP":
            # inference.py:238-284
            cav_mean, cav_cov = self.cavity_distribution(self.power)
            cav_mean_f, cav_cov_f = self.conditional_posterior_to_data(cav_mean, cav_cov)
            _, dlel, d2lel = self._moment_match(cav_mean_f, cav_cov_f)
            cav_prec = inv(cav_cov_f)
            scale_factor = cav_prec @ inv(d2lel + cav_prec) / self.power
            dlel = scale_factor @ dlel
            d2lel = scale_factor @ d2lel
            if self.mask_pseudo_y is not None:
                mask = self.mask_pseudo_y[..., None]
                dlel = np.where(mask, np.nan, dlel)
                d2m = np.where(mask | T(mask), 0.0, d2lel)
                idx = np.arange(d2m.shape[-1])
                dd = d2m[..., idx, idx]
                d2m[..., idx, idx] = np.where(mask[..., 0], np.nan, dd)
                d2lel = d2m
            if ensure_psd:
                d2lel = -ensure_diagonal_positive_precision(-d2lel)
            return cav_mean_f, dlel, d2lel
        raise NotImplementedError(kind)

    def inference(self, lr=1.0, ensure_psd=True, batch_ind=None):
        """inference.py:65-90.  batch_ind (indices into the time-sorted data): the site
        map runs on those points only and group_natural_params (basemodels.py:195-199)
        writes them into the full site arrays.  The site map is elementwise in the data
        points, so it is evaluated on all of them and the batch is selected."""
        if batch_ind is not None and np.asarray(batch_ind).shape[0] == self.num_data:
            batch_ind = None
        self.update_posterior()
        mean, jac, hess = self.update_variational_params(ensure_psd=ensure_psd)
        if batch_ind is not None:
            b = np.asarray(batch_ind)
            mean, jac, hess = mean[b], jac[b], hess[b]
        nat1_n, nat2_n = newton_update(mean, jac, hess)
        if batch_ind is not None:
            n1, n2 = self.nat1.copy(), self.nat2.copy()
            n1[b], n2[b] = nat1_n, nat2_n
            nat1_n, nat2_n = n1, n2
        diff1 = np.mean(np.abs(nat1_n - self.nat1))
        diff2 = np.mean(np.abs(nat2_n - self.nat2))
        self.update_nat_params((1 - lr) * self.nat1 + lr * nat1_n, (1 - lr) * self.nat2 + lr * nat2_n)
        self.update_posterior()
        return (mean, jac, hess), (diff1, diff2)

    def energy(self, batch_ind=None):
        """inference.py:130-154, 197-222, 286-325; with batch_ind the data term sums the
        batch and is scaled by N / B, the KL / pseudo terms keep every site."""
        kind = self.inference_kind
        sel = np.arange(self.num_data) if batch_ind is None else np.asarray(batch_ind)
        scale = self.num_data / sel.shape[0]
        if kind == "VI":
            # inference.py:197-222
            mean_f, cov_f = self.conditional_posterior_to_data()
            ell, _, _ = self._variational_expectation(mean_f, cov_f)
            return -(scale * np.nansum(ell[sel]) - self.compute_kl())
        if kind == "Newton":
            # inference.py:130-154
            mean_f, _ = self.conditional_posterior_to_data()
            ll, _, _ = self._log_likelihood_gradients(mean_f)
            return -(scale * np.nansum(ll[sel]) - self.compute_kl())
        if kind == "EP":
            # inference.py:286-325 with basemodels.py:247-262
            if self.mask_y.ndim == 2 and self.mask_y.shape[1] > 1:
                # inference.py:313-317: mask_y [N, n_loc] against one lel per step
                raise NotImplementedError("masking in spatio-temporal models not implemented for EP")
            power = self.power
            cav_mean, cav_cov = self.cavity_distribution(power)
            pseudo_y, pseudo_var = self.compute_full_pseudo_lik()
            mask = self._mask3()
            lel_pseudo = mvn_logpdf(pseudo_y, cav_mean, pseudo_var / power + cav_cov, mask)
            lel_pseudo = lel_pseudo + pep_constant(pseudo_var, power,
                                                   None if mask is None else mask[..., 0])
            lZ = self.compute_log_lik(pseudo_y, pseudo_var)
            cav_mean_f, cav_cov_f = self.conditional_posterior_to_data(cav_mean, cav_cov)
            lel, _, _ = self._moment_match(cav_mean_f, cav_cov_f)
            lel = np.where(np.squeeze(self.mask_y), 0.0, np.squeeze(lel))
            return -(lZ + 1.0 / power * (scale * np.nansum(np.atleast_1d(lel)[sel]) - np.nansum(lel_pseudo)))
        raise NotImplementedError(kind)


class DenseGP:
    """Kernel-matrix GP used to pin the Markov oracle (basemodels.py:265-330, ops.py:52-80).

    Holds the same site parameterisation as MarkovGP; posterior and pseudo
    log-normaliser come from dense Cholesky algebra, O(N^3).
    """

    def __init__(self, kernel, X, site_mean, site_var):
        self.kernel = kernel
        self.X = np.asarray(X, np.float64)
        if self.X.ndim < 2:
            self.X = self.X[:, None]
        self.y = np.asarray(site_mean, np.float64).reshape(-1)
        self.r = np.asarray(site_var, np.float64).reshape(-1)

    def posterior(self):
        """ops.py:52-80 gaussian_conditional."""
        K = self.kernel(self.X, self.X)
        Ky = K + np.diag(self.r)
        L = np.linalg.cholesky(Ky)
        A = np.linalg.solve(L, K)
        alpha = np.linalg.solve(L.T, np.linalg.solve(L, self.y))
        mean = K @ alpha
        cov = K - A.T @ A
        return mean, cov

    def log_marginal(self):
        """basemodels.py:307-330: log N(y | 0, K + R)."""
        K = self.kernel(self.X, self.X)
        Ky = K + np.diag(self.r)
        L = np.linalg.cholesky(Ky)
        z = np.linalg.solve(L, self.y)
        return -0.5 * z @ z - np.sum(np.log(np.diag(L))) - 0.5 * len(self.y) * LOG2PI
