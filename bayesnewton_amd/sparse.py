"""Sparse Markov GP on the HIP path (bayesnewton/basemodels.py:928-1152 with the VI and
Newton mixins, models.py:135-157, 302-310, 330).

The M inducing states sit at the sorted inducing times Z, with sentinels at -/+1e10;
each of the M + 1 transitions (u_i, u_{i+1}) carries a 2d-dimensional Gaussian site.
update_posterior runs the filter over pairs of states (bn_pairs_filter,
ops.py:383-426) and the RTS smoother over the inducing states (bn_rts_smoother with
the full state and gains, ops.py:288-354) on the pair filter's first-state moments, as
the reference does; the site update and the energy's data and KL terms run per
transition in bn_sparse_sites.  Temporal inputs only (no spatial R).
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from . import cubature as cub
from .models import input_admin
from .plan import SeriesPlan, default_chunk_len

_EDGE = 1e10


class SparseMarkovGaussianProcess:
    """basemodels.py:928-1152 (temporal inputs)."""
    method = None

    def __init__(self, kernel, likelihood, X, Y, Z, R=None, parallel=None, device=None):
        if R is not None:
            raise NotImplementedError("the sparse Markov model is implemented for temporal inputs")
        if self.method is None:
            raise TypeError("use SparseMarkovVariationalGP or SparseMarkovLaplaceGP")
        if not torch.cuda.is_available():
            raise nat.NativeError("bayesnewton_amd models need a ROCm GPU")
        t, y, _ = input_admin(X, Y)
        self.X, self.Y = t[:, None], y[:, None]
        self.num_data = t.shape[0]
        self.kernel, self.likelihood = kernel, likelihood
        self.parallel = True if parallel is None else bool(parallel)
        d = kernel.state_dim
        self.state_dim = d
        z = np.sort(np.asarray(Z, dtype=np.float64).reshape(-1))
        z_aug = np.concatenate([[-_EDGE], z, [_EDGE]])
        self.Z = z_aug[:, None]
        self.dz = np.diff(z_aug)
        self.n_z = z.shape[0]
        self.num_transitions = self.n_z + 1
        # set_z_stats (utils.py:556-559): transition of every data point
        self.ind = np.searchsorted(z_aug, t) - 1
        counts = np.bincount(self.ind, minlength=self.num_transitions)
        self.num_neighbours = counts
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        f64 = dict(dtype=torch.float64, device=dev)
        self._desc = kernel.desc()
        self._t = torch.tensor(t, **f64)
        self._y = torch.tensor(y, **f64)
        self._z = torch.tensor(z_aug, **f64)
        self._dz = torch.tensor(self.dz, **f64)
        self._tr_off = torch.tensor(np.concatenate([[0], np.cumsum(counts)]), dtype=torch.int64, device=dev)
        self._pair_off = torch.tensor([0, self.num_transitions], dtype=torch.int64, device=dev)
        n_tr, d2 = self.num_transitions, 2 * d
        # sites (basemodels.py:950-962): nat1 = 0, nat2 = 1e-8 I with 1e-2 on the second state's
        # first entry of every transition but the last (the Markov GP's first step when Z = X)
        nat2 = np.tile(1e-8 * np.eye(d2), (n_tr, 1, 1))
        nat2[:-1, d, d] = 1e-2
        cov = np.linalg.inv(nat2)
        self._nat1 = torch.zeros((n_tr, d2), **f64)
        self._nat2 = torch.tensor(nat2, **f64)
        self._site_mean = torch.zeros((n_tr, d2), **f64)
        self._site_cov = torch.tensor(cov, **f64)
        self._pair_mean = torch.zeros((n_tr, d2), **f64)
        self._pair_cov = torch.eye(d2, **f64).expand(n_tr, d2, d2).contiguous()
        # the inducing-state chain the smoother runs over
        self.plan = SeriesPlan(np.array([0, self.n_z], dtype=np.int64),
                               chunk_len=default_chunk_len([self.n_z], self.parallel), device=dev)
        P = self.plan
        self._dts = P.gather(torch.tensor(self.dz[1:], dtype=torch.float32, device=dev))
        self._pf_mean = torch.empty((n_tr, d2), **f64)
        self._pf_cov = torch.empty((n_tr, d2, d2), **f64)
        self._ll = torch.zeros(1, **f64)
        self._sm_t = torch.empty((d, P.n_slots), **f64)
        self._sP_t = torch.empty((d * d, P.n_slots), **f64)
        self._G_t = torch.empty((d * d, P.n_slots), **f64)
        self._sm = torch.zeros((self.n_z, d), **f64)
        self._sP = torch.zeros((self.n_z, d * d), **f64)
        self._G = torch.zeros((self.n_z, d * d), **f64)
        self._p1 = torch.empty(n_tr, **f64)
        self._p2 = torch.empty(n_tr, **f64)
        self._s = torch.empty(2, **f64)
        self._hyper_key = None
        self._hyper_version = 0
        self._sites_version = 0
        self._post_key = None
        self._ll_key = None

    # ---- state keys ------------------------------------------------------------------------
    def _hypers(self):
        key = tuple((c.lengthscale, c.variance) for c in self.kernel.components)
        if key != self._hyper_key:
            self._ls, self._var = self.kernel.hypers(self.device)
            self._hyper_key = key
            self._hyper_version += 1
        return self._ls, self._var

    def _state_key(self):
        self._hypers()
        return (self._hyper_version, self._sites_version)

    # ---- posterior (basemodels.py:980-1005) ----------------------------------------------------
    def _pairs_filter(self):
        ls, var = self._hypers()
        nat.check(nat.lib().bn_pairs_filter(ctypes.byref(self._desc), nat.ptr(ls), nat.ptr(var), 0, 1,
                                            nat.ptr(self._pair_off), nat.ptr(self._dz), nat.ptr(self._site_mean),
                                            nat.ptr(self._site_cov), nat.ptr(self._pf_mean), nat.ptr(self._pf_cov),
                                            nat.ptr(self._ll), nat.stream_ptr()))
        self._ll_key = self._state_key()

    def update_posterior(self):
        d = self.state_dim
        self._pairs_filter()
        ls, var = self._hypers()
        P = self.plan
        # the pair filter's first-state moments after every step but the first, as filtered moments
        fm = P.gather(self._pf_mean[1:, :d].contiguous())
        fP = P.gather(self._pf_cov[1:, :d, :d].reshape(self.n_z, d * d).contiguous())
        nat.check(nat.lib().bn_rts_smoother(P.handle, ctypes.byref(self._desc), nat.ptr(ls), nat.ptr(var), 0,
                                            nat.ptr(self._dts), nat.ptr(fm), nat.ptr(fP), nat.ptr(self._sm_t),
                                            nat.ptr(self._sP_t), nat.ptr(self._G_t), nat.BN_RETURN_FULL,
                                            nat.ptr(P.workspace(self._desc)), nat.stream_ptr()))
        self._sm.copy_(P.scatter(self._sm_t))
        self._sP.copy_(P.scatter(self._sP_t))
        self._G.copy_(P.scatter(self._G_t))
        self._post_key = self._state_key()
        self._run_sites(nat.BN_SPARSE_ENERGY)  # pair joint moments of the new posterior

    def _ensure_posterior(self):
        if self._post_key != self._state_key():
            self.update_posterior()

    def _ensure_loglik(self):
        if self._ll_key != self._state_key():
            self._pairs_filter()

    def _run_sites(self, mode, lr=1.0, cubature=None, ensure_psd=True):
        ls, var = self._hypers()
        lik = self.likelihood.desc()
        nat.check(nat.lib().bn_sparse_sites(
            ctypes.byref(self._desc), nat.ptr(ls), nat.ptr(var), ctypes.byref(lik), self.method, mode,
            cub.num_points(cubature), float(lr), 1 if ensure_psd else 0, self.n_z, nat.ptr(self._z),
            nat.ptr(self._sm), nat.ptr(self._sP), nat.ptr(self._G), nat.ptr(self._tr_off), nat.ptr(self._t),
            nat.ptr(self._y), nat.ptr(self._nat1), nat.ptr(self._nat2), nat.ptr(self._site_mean),
            nat.ptr(self._site_cov), nat.ptr(self._pair_mean), nat.ptr(self._pair_cov), nat.ptr(self._p1),
            nat.ptr(self._p2), nat.stream_ptr()))

    def _sums(self):
        nat.check(nat.lib().bn_st_sum(1, self.num_transitions, nat.ptr(self._p1), nat.ptr(self._s[:1]),
                                      nat.stream_ptr()))
        nat.check(nat.lib().bn_st_sum(1, self.num_transitions, nat.ptr(self._p2), nat.ptr(self._s[1:]),
                                      nat.stream_ptr()))
        return self._s[0].clone(), self._s[1].clone()

    @property
    def posterior_mean(self):
        """Joint smoothed mean of every transition's pair of states [M + 1, 2d, 1]."""
        self._ensure_posterior()
        return self._pair_mean[..., None].clone()

    @property
    def posterior_variance(self):
        """Joint smoothed covariance of every transition's pair of states [M + 1, 2d, 2d]."""
        self._ensure_posterior()
        return self._pair_cov.clone()

    @property
    def site_nat(self):
        return self._nat1[..., None].clone(), self._nat2.clone()

    def compute_log_lik(self):
        """basemodels.py:1007-1016: log normaliser of the pairs model."""
        self._ensure_loglik()
        return self._ll[0].clone()

    # ---- inference (inference.py:65-90) ----------------------------------------------------------
    def inference(self, lr=1.0, batch_ind=None, cubature=None, ensure_psd=True, **kwargs):
        """One site update of every transition and the posterior update.  Returns
        (None, (diff1, diff2)): the per-point terms of the pair-space update are not kept."""
        if batch_ind is not None and len(batch_ind) != self.num_data:
            raise NotImplementedError("mini-batched site updates are not implemented on the HIP path")
        self._ensure_posterior()
        self._run_sites(nat.BN_SPARSE_UPDATE, lr, cubature, ensure_psd)
        self._sites_version += 1
        s1, s2 = self._sums()
        d2 = 2 * self.state_dim
        diff1 = s1 / (self.num_transitions * d2)
        diff2 = s2 / (self.num_transitions * d2 * d2)
        self.update_posterior()
        return None, (diff1, diff2)

    def energy(self, batch_ind=None, cubature=None, **kwargs):
        """VI (inference.py:197-222) / Newton (130-154) energy with the sparse KL
        (basemodels.py:1018-1030): -(sum_n data term - (E_q log N(sites) - log Z))."""
        if batch_ind is not None and len(batch_ind) != self.num_data:
            raise NotImplementedError("mini-batched energies are not implemented on the HIP path")
        self._ensure_loglik()
        self._run_sites(nat.BN_SPARSE_ENERGY, cubature=cubature)
        ell, edp = self._sums()
        return -(ell - (edp - self._ll[0]))

    # ---- prediction (basemodels.py:1032-1072) ---------------------------------------------------
    def predict(self, X=None, R=None):
        """Predictive mean and variance of f at the test times X (default: the data times):
        the smoothed inducing states conditioned at every test time (utils.py:99-141,
        sentinels (0, Pinf) at -/+1e10)."""
        if R is not None:
            raise NotImplementedError("temporal inputs only")
        self._ensure_posterior()
        X = self.X if X is None else X
        x = torch.tensor(np.asarray(X, dtype=np.float64).reshape(-1), dtype=torch.float64, device=self.device)
        n = x.shape[0]
        ls, var = self._hypers()
        f64 = dict(dtype=torch.float64, device=self.device)
        xtr = self._z[1:-1].contiguous()
        tr_off = torch.tensor([0, self.n_z], dtype=torch.int64, device=self.device)
        te_off = torch.tensor([0, n], dtype=torch.int64, device=self.device)
        mean = torch.empty(n, **f64)
        var_out = torch.empty(n, **f64)
        nat.check(nat.lib().bn_predict(self.plan.handle, ctypes.byref(self._desc), nat.ptr(ls), nat.ptr(var), 0,
                                       nat.ptr(tr_off), nat.ptr(xtr), nat.ptr(self._sm_t), nat.ptr(self._sP_t),
                                       nat.ptr(self._G_t), nat.ptr(te_off), nat.ptr(x), n, nat.ptr(mean),
                                       nat.ptr(var_out), nat.stream_ptr()))
        return mean, var_out


class SparseMarkovVariationalGP(SparseMarkovGaussianProcess):
    """models.py:135-157"""
    method = nat.BN_VI


class SparseMarkovNewtonGP(SparseMarkovGaussianProcess):
    """models.py:302-310"""
    method = nat.BN_NEWTON


SparseMarkovLaplaceGP = SparseMarkovNewtonGP
SparseMarkovGP = SparseMarkovGaussianProcess
