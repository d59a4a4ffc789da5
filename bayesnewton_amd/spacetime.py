"""Space-time Markov GPs: MarkovGaussianProcess with a SpatioTemporalKernel
(bayesnewton/basemodels.py:625-764 on its `spatio_temporal` branch, kernels.py:385-586).

The reference builds the model as MarkovVariationalGP(kernel, likelihood, X=t, R=r, Y=Y)
with t [T], r [T, R, 2] and Y [T, R] (NaN = missing); the same call works here.
`from_runs` puts K independent runs (same T and R, same hypers) into one model so
that the per-run filter/smoother workgroups fill the GPU.

Data flow of one inference step (inference.py:65-90), all on the device:
  k_st_sites(UPDATE)   posterior -> data (B m, diag(B P B') + diag C), new sites,
                       whitened projection L L' = B' diag(nat2) B + 1e-12 I, z = L^-1 B' nat1
  k_st_filter          Kalman filter of the projected pseudo observations (state 2M)
  k_st_smoother        RTS smoother: posterior H m, H P H' at every step
energy() (inference.py:130-154, 197-222) uses the stored posterior, the current sites
and the current hypers: -(sum E_q[log p(y|f)] - (sum E_q[log N(pseudo_y|f, pseudo_var)] - log Z)).
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from . import cubature as cub
from .kernels import SpatioTemporalKernel


def _admin(t, r, y):
    """utils.py:234-265 for one run: sort by time, dt[0] = 0."""
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    r = np.asarray(r, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    T = t.shape[0]
    if r.ndim != 3 or r.shape[0] != T or r.shape[2] != 2:
        raise ValueError("R must be [T, n_loc, 2]")
    if y.shape != r.shape[:2]:
        raise ValueError("Y must be [T, n_loc]")
    ind = np.argsort(t, kind="stable")
    t, r, y = t[ind], r[ind], y[ind]
    dt = np.concatenate([[0.0], np.diff(t)])
    return t, r, y, dt


class SpaceTimeMarkovGP:
    """Markov GP over a SpatioTemporalKernel, K independent runs per model."""
    method = None

    def __init__(self, kernel, likelihood, X, Y, R, parallel=None, device=None):
        self.parallel = parallel
        self._build(kernel, likelihood, [X], [R], [Y], device)
        self._single = True

    @classmethod
    def from_runs(cls, kernel, likelihood, Xs, Rs, Ys, parallel=None, device=None):
        """K runs (lists of t [T], R [T, n_loc, 2], Y [T, n_loc]) under one set of hypers."""
        self = cls.__new__(cls)
        self.parallel = parallel
        self._build(kernel, likelihood, Xs, Rs, Ys, device)
        self._single = False
        return self

    # ---- construction ----------------------------------------------------------------
    def _build(self, kernel, likelihood, Xs, Rs, Ys, device):
        if not isinstance(kernel, SpatioTemporalKernel):
            raise TypeError("kernel must be a SpatioTemporalKernel")
        if not torch.cuda.is_available():
            raise nat.NativeError("bayesnewton_amd models need a ROCm GPU")
        self.kernel = kernel
        self.likelihood = likelihood
        runs = [_admin(t, r, y) for t, r, y in zip(Xs, Rs, Ys)]
        T, nr = runs[0][2].shape
        for run in runs:
            if run[2].shape != (T, nr):
                raise ValueError("all runs need the same number of steps and locations")
        K, M = len(runs), kernel.M
        KT = kernel.temporal_kernel.state_dim
        self.K, self.T, self.n_loc, self.M, self.D = K, T, nr, M, KT * M
        self.X = np.stack([run[0] for run in runs])
        self.R = np.stack([run[1] for run in runs])
        self.Y = np.stack([run[2] for run in runs])
        self.num_data = T
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self._dims = nat.StDims(K, T, nr, M, kernel.temporal_kernel.kind)
        f64 = dict(dtype=torch.float64, device=dev)
        self._dt = torch.tensor(np.stack([run[3] for run in runs]), **f64)
        self._r = torch.tensor(self.R, dtype=torch.float32, device=dev).contiguous()
        self._y = torch.tensor(self.Y, dtype=torch.float32, device=dev).contiguous()
        self._z = torch.tensor(kernel.z, **f64).contiguous()
        # sites N(0, 1e2 I) (basemodels.py:130-133)
        self._nat1 = torch.zeros((K, T, nr), **f64)
        self._nat2 = torch.full((K, T, nr), 1e-2, **f64)
        self._B = torch.empty((K, T, nr, M), **f64)
        self._cdiag = torch.empty((K, T, nr), **f64)
        self._Lw = torch.empty((K, T, M, M), **f64)
        self._zw = torch.empty((K, T, M), **f64)
        self._ldl = torch.empty((K, T), **f64)
        self._ell = torch.empty((K, T), **f64)
        self._loglik = torch.zeros(K, **f64)
        self._alloc_filter()
        # posterior initialised as in basemodels.py:134-135 (zeros, identity)
        self._pm = torch.zeros((K, T, M), **f64)
        self._pP = torch.eye(M, **f64).expand(K, T, M, M).contiguous()
        self._p0 = torch.empty((K, T), **f64)
        self._p1 = torch.empty((K, T), **f64)
        self._s0 = torch.empty(K, **f64)
        self._s1 = torch.empty(K, **f64)
        self._hyp_key = None
        self._cond_key = None
        self._proj_key = None
        self._post_key = None
        self._ll_key = None
        self._sites_version = 0

    def _alloc_filter(self):
        """Filter outputs and scratch: dense state [K,T,D] and packed covariances."""
        K, T, D = self.K, self.T, self.D
        f64 = dict(dtype=torch.float64, device=self.device)
        self._fm = torch.empty((K, T, D), **f64)
        self._fP = torch.empty((K, T, D * (D + 1) // 2), **f64)
        nb = nat.lib().bn_st_workspace_bytes(ctypes.byref(self._dims))
        nat.check(0 if nb > 0 else -1)
        self._ws = torch.empty(int(nb), dtype=torch.uint8, device=self.device)

    # ---- hyperparameters ---------------------------------------------------------------
    def _hyp(self):
        kt, ks = self.kernel.temporal_kernel, self.kernel.spatial_kernel
        key = (kt.lengthscale, kt.variance, ks.lengthscale, ks.variance)
        if key != self._hyp_key:
            dev = self.device
            self._hls = torch.tensor([key[0]], dtype=torch.float64, device=dev)
            self._hvt = torch.tensor([key[1]], dtype=torch.float64, device=dev)
            self._hvs = torch.tensor([key[3]], dtype=torch.float64, device=dev)
            self._hyp_key = key
        h = nat.StHypers(nat.ptr(self._hls), nat.ptr(self._hvt), nat.ptr(self._hvs), 0)
        return h, key

    def _ensure_conditional(self):
        h, key = self._hyp()
        ckey = (key[1], key[2], key[3], self.kernel.conditional)
        if ckey == self._cond_key:
            return h
        _, Lzz = self.kernel.inducing_precision()
        self._Lzz = torch.tensor(np.ascontiguousarray(Lzz), dtype=torch.float64, device=self.device)
        nat.check(nat.lib().bn_st_conditional(
            ctypes.byref(self._dims), ctypes.byref(h), self.kernel.spatial_kernel.kind, float(key[2]),
            nat.ptr(self._z), nat.ptr(self._Lzz), nat.ptr(self._r), nat.ptr(self._B), nat.ptr(self._cdiag),
            nat.stream_ptr()))
        if self.kernel.conditional == "DTC":
            self._cdiag.zero_()
        self._cond_key = ckey
        return h

    # ---- sites ---------------------------------------------------------------------------
    def _sites(self, mode, lr=1.0, cubature=None, ensure_psd=True, terms=None, batch=None):
        lik = self.likelihood.desc()
        om, oj, oh = terms if terms is not None else (None, None, None)
        nat.check(nat.lib().bn_st_sites(
            ctypes.byref(self._dims), ctypes.byref(lik), self.method, mode, cub.num_points(cubature), float(lr),
            1 if ensure_psd else 0, nat.ptr(self._y), nat.ptr(self._B), nat.ptr(self._cdiag), nat.ptr(self._pm),
            nat.ptr(self._pP), nat.ptr(batch), nat.ptr(self._nat1), nat.ptr(self._nat2), nat.ptr(self._Lw),
            nat.ptr(self._zw), nat.ptr(self._ldl), nat.ptr(self._p0), nat.ptr(self._p1), nat.ptr(om), nat.ptr(oj),
            nat.ptr(oh), nat.stream_ptr()))

    def _batch_mask(self, batch_ind):
        """[K, T] mask of a mini-batch of time steps (indices into the time-sorted steps,
        the same for every run), or None for all of them (inference.py:67-68)."""
        if batch_ind is None:
            return None
        idx = np.asarray(batch_ind).reshape(-1)
        if idx.shape[0] == self.num_data:
            return None
        if idx.size == 0:
            raise ValueError("empty mini-batch")
        if idx.min() < 0 or idx.max() >= self.num_data:
            raise IndexError("batch_ind out of range")
        m = torch.zeros((self.K, self.T), dtype=torch.float32, device=self.device)
        m[:, torch.as_tensor(idx, dtype=torch.int64, device=self.device)] = 1.0
        return m

    def _sum(self, per_step, out):
        nat.check(nat.lib().bn_st_sum(self.K, self.T, nat.ptr(per_step), nat.ptr(out), nat.stream_ptr()))
        return out

    def _ensure_projection(self):
        self._ensure_conditional()
        key = (self._cond_key, self._sites_version)
        if key != self._proj_key:
            self._sites(nat.BN_ST_PROJECT)
            self._proj_key = key

    def _run_filter(self):
        h, key = self._hyp()
        nat.check(nat.lib().bn_st_filter(
            ctypes.byref(self._dims), ctypes.byref(h), nat.ptr(self._dt), nat.ptr(self._Lw), nat.ptr(self._zw),
            nat.ptr(self._ldl), nat.ptr(self._fm), nat.ptr(self._fP), nat.ptr(self._ell), nat.ptr(self._loglik),
            nat.ptr(self._ws), nat.stream_ptr()))
        self._ll_key = (self._proj_key, key[0], key[1])
        return h

    # ---- posterior (basemodels.py:689-706) ----------------------------------------------
    def update_posterior(self):
        self._ensure_projection()
        h = self._run_filter()
        nat.check(nat.lib().bn_st_smoother(
            ctypes.byref(self._dims), ctypes.byref(h), nat.ptr(self._dt), nat.ptr(self._fm), nat.ptr(self._fP),
            nat.ptr(self._pm), nat.ptr(self._pP), nat.ptr(self._ws), nat.stream_ptr()))
        self._post_key = self._ll_key

    def _current_key(self):
        _, key = self._hyp()
        ckey = (key[1], key[2], key[3], self.kernel.conditional)
        return ((ckey, self._sites_version), key[0], key[1])

    def _ensure_posterior(self):
        if self._post_key != self._current_key():
            self.update_posterior()

    def _ensure_loglik(self):
        self._ensure_projection()
        if self._ll_key != self._current_key():
            self._run_filter()

    def compute_log_lik(self):
        """basemodels.py:726-741 (per run when the model holds several)."""
        self._ensure_loglik()
        return self._loglik.clone() if not self._single else self._loglik[0].clone()

    def _out(self, a):
        return a[0] if self._single else a

    @property
    def posterior_mean(self):
        """H m of the smoothed state, [T, M, 1] ([K, T, M, 1] for several runs)."""
        return self._out(self._pm[..., None])

    @property
    def posterior_variance(self):
        """H P H' of the smoothed state, [T, M, M] ([K, T, M, M])."""
        return self._out(self._pP)

    @property
    def site_nat(self):
        """Current sites (nat1, diagonal of nat2), [T, n_loc] each ([K, T, n_loc])."""
        return self._out(self._nat1), self._out(self._nat2)

    def set_sites(self, nat1, nat2):
        """Replace the sites (update_nat_params, basemodels.py:97-100); nat2 is the diagonal."""
        self._nat1.copy_(torch.as_tensor(nat1, dtype=torch.float64).reshape(self._nat1.shape))
        self._nat2.copy_(torch.as_tensor(nat2, dtype=torch.float64).reshape(self._nat2.shape))
        self._sites_version += 1

    # ---- prediction (basemodels.py:766-816, 165-193) ------------------------------------------
    def predict(self, X=None, R=None):
        """Predictive mean and variance of f at test times X [n_test] and locations
        R [n_test, n_loc, 2] (default: the training inputs), from the filter and a
        full-state smoother with the current sites and hypers; [n_test, n_loc] FP64 device
        tensors ([K, n_test, n_loc] for several runs).  The spatial covariance is reduced to
        its diagonal, as the reference does."""
        self._ensure_projection()
        return _predict_spacetime(self, X, R)

    # ---- inference (inference.py:65-90) --------------------------------------------------
    def inference(self, lr=1.0, batch_ind=None, cubature=None, ensure_psd=True, return_terms=True, **kwargs):
        """One site update + posterior update.  Returns ((mean, jacobian, hessian), (diff1,
        diff2)) with the site terms per location, [T, n_loc] each (the hessian's diagonal),
        and the mean absolute change of the natural parameters.  batch_ind: a mini-batch
        of time steps; only their sites change, and the terms are those steps' [B, n_loc]."""
        batch = self._batch_mask(batch_ind)
        self._ensure_conditional()
        self._ensure_posterior()
        terms = None
        if return_terms:
            terms = tuple(torch.empty((self.K, self.T, self.n_loc), dtype=torch.float32, device=self.device)
                          for _ in range(3))
        self._sites(nat.BN_ST_UPDATE, lr, cubature, ensure_psd, terms, batch)
        self._sites_version += 1
        self._proj_key = (self._cond_key, self._sites_version)
        # inference.py:79-80 averages over nat1 [T, R, 1] and nat2 [T, R, R] (whose
        # off-diagonal entries are zero)
        n = float(self.T * self.n_loc)
        d1 = self._sum(self._p0, self._s0) / n
        d2 = self._sum(self._p1, self._s1) / (n * self.n_loc)
        self.update_posterior()
        out_terms = tuple(a.to(torch.float64) for a in terms) if terms else (None, None, None)
        if terms and batch is not None:
            sel = torch.as_tensor(np.asarray(batch_ind).reshape(-1), dtype=torch.int64, device=self.device)
            out_terms = tuple(a[:, sel] for a in out_terms)
        out_terms = tuple(self._out(a) for a in out_terms) if terms else out_terms
        return out_terms, (self._out(d1.clone()), self._out(d2.clone()))

    # ---- energy ------------------------------------------------------------------------------
    def energy_per_run(self, cubature=None, batch_ind=None):
        """Energy of every run; with batch_ind the data term sums those time steps, scaled
        by T / B (inference.py:130-154, 197-222)."""
        batch = self._batch_mask(batch_ind)
        scale = 1.0 if batch is None else self.num_data / float(np.asarray(batch_ind).size)
        self._ensure_loglik()
        self._sites(nat.BN_ST_ENERGY, cubature=cubature, batch=batch)
        ell = self._sum(self._p0, self._s0).clone()
        ed = self._sum(self._p1, self._s1).clone()
        return -(scale * ell - (ed - self._loglik))

    def energy(self, batch_ind=None, cubature=None, **kwargs):
        e = self.energy_per_run(cubature, batch_ind)
        return e[0] if self._single else e.sum()

    # ---- hyperparameter gradients ------------------------------------------------------
    def _trainables(self):
        """(owner, attribute) of every trainable raw (softplus-inverse) hyperparameter:
        the temporal kernel's, the spatial kernel's (kernels.py:396-444 holds both as
        sub-kernels), then the likelihood's."""
        out = []
        for k in (self.kernel.temporal_kernel, self.kernel.spatial_kernel):
            if not k.fix_lengthscale:
                out.append((k, "transformed_lengthscale"))
            if not k.fix_variance:
                out.append((k, "transformed_variance"))
        lik = self.likelihood
        for attr, fix in (("transformed_variance", "fix_variance"), ("transformed_scale", "fix_scale")):
            if hasattr(lik, attr) and not getattr(lik, fix, False):
                out.append((lik, attr))
        return out

    def trainable_names(self):
        names = {id(self.kernel.temporal_kernel): "temporal", id(self.kernel.spatial_kernel): "spatial"}
        return [f"{names.get(id(o), type(o).__name__)}.{a}" for o, a in self._trainables()]

    def get_raw_hypers(self):
        return np.array([getattr(o, a) for o, a in self._trainables()], dtype=np.float64)

    def set_raw_hypers(self, raw):
        for (o, a), v in zip(self._trainables(), np.asarray(raw, dtype=np.float64).reshape(-1)):
            setattr(o, a, float(v))

    def energy_and_grad_per_run(self, rel_step=1e-5, cubature=None):
        """Energy of every run and its gradient w.r.t. the raw hypers (trainable_names()
        order) with the sites and the posterior held fixed, as objax.GradValues(energy)
        differentiates the reference's energy (README.md:56-76; the gradient reaches the
        spatial hypers through B and C, kernels.py:486-515, and the temporal ones through
        C and the filter's log-likelihood).  Central differences of the FP64 energy on the
        device, step rel_step * |natural value| on the natural scale, chained through
        softplus; every energy evaluation recomputes B, C, the site projection and the
        filter at the shifted hypers and leaves the stored posterior alone.
        Returns (E [K], dE/draw [K, n_trainable])."""
        E0 = self.energy_per_run(cubature).clone()
        raw = self.get_raw_hypers()
        grads = []
        for i, (o, a) in enumerate(self._trainables()):
            nat_attr = a.replace("transformed_", "")
            v0 = getattr(o, nat_attr)
            h = rel_step * max(abs(v0), 1e-12)
            setattr(o, nat_attr, v0 + h)
            ep = self.energy_per_run(cubature).clone()
            setattr(o, nat_attr, v0 - h)
            em = self.energy_per_run(cubature).clone()
            self.set_raw_hypers(raw)  # exact restore of the stored raw value
            sig = 1.0 / (1.0 + np.exp(-raw[i]))  # d softplus / d raw
            grads.append((ep - em) / (2.0 * h) * sig)
        # leave the projection and log-likelihood at the original hypers
        self.energy_per_run(cubature)
        return E0, torch.stack(grads, 1) if grads else torch.zeros((self.K, 0), dtype=torch.float64,
                                                                      device=self.device)

    def energy_and_grad(self, rel_step=1e-5, cubature=None):
        """(energy, dE/draw) summed over runs, numpy gradient in trainable_names() order."""
        E, g = self.energy_and_grad_per_run(rel_step, cubature)
        if self._single:
            return float(E[0]), g[0].cpu().numpy()
        return float(E.sum()), g.sum(0).cpu().numpy()


def _test_inputs(model, X, R):
    K = model.K
    if X is None:
        return model.X, model.R
    Xt = np.broadcast_to(np.asarray(X, dtype=np.float64).reshape(-1), (K, np.size(X)))
    Rt = np.asarray(R, dtype=np.float64)
    if Rt.ndim != 3 or Rt.shape[0] != Xt.shape[1] or Rt.shape[2] != 2:
        raise ValueError("R must be [n_test, n_loc, 2] for X [n_test]")
    return Xt, np.broadcast_to(Rt, (K,) + Rt.shape)


def _predict_from_state(model, h, X, R, sm, sP=None, sx=None, steady=None):
    """f at test times / locations from the smoothed full state (layout c*M + m):
    temporal conditional on the neighbours (k_st_predict) and the spatial conditional of
    the test locations (basemodels.py:787-812)."""
    K, M = model.K, model.M
    dev = model.device
    f64 = dict(dtype=torch.float64, device=dev)
    Xt, Rt = _test_inputs(model, X, R)
    ns, rs = Xt.shape[1], Rt.shape[2]
    tdims = nat.StDims(K, ns, rs, M, model.kernel.temporal_kernel.kind)
    rt = torch.tensor(np.ascontiguousarray(Rt), dtype=torch.float32, device=dev)
    Bt = torch.empty((K, ns, rs, M), **f64)
    ct = torch.empty((K, ns, rs), **f64)
    model._ensure_conditional()
    nat.check(nat.lib().bn_st_conditional(
        ctypes.byref(tdims), ctypes.byref(h), model.kernel.spatial_kernel.kind,
        float(model.kernel.spatial_kernel.lengthscale), nat.ptr(model._z), nat.ptr(model._Lzz), nat.ptr(rt),
        nat.ptr(Bt), nat.ptr(ct), nat.stream_ptr()))
    if model.kernel.conditional == "DTC":
        ct.zero_()
    xtr = torch.tensor(np.ascontiguousarray(model.X), **f64)
    xte = torch.tensor(np.ascontiguousarray(Xt), **f64)
    mean = torch.empty((K, ns, rs), **f64)
    var = torch.empty((K, ns, rs), **f64)
    nat.check(nat.lib().bn_st_predict(
        ctypes.byref(model._dims), ctypes.byref(h), nat.ptr(xtr), ns, rs, nat.ptr(xte), nat.ptr(Bt), nat.ptr(ct),
        nat.ptr(sm), nat.ptr(sP), nat.ptr(sx), nat.ptr(steady), nat.ptr(mean), nat.ptr(var), nat.stream_ptr()))
    return model._out(mean), model._out(var)


def _predict_spacetime(model, X=None, R=None):
    """MarkovGaussianProcess.predict on the space-time branch (basemodels.py:766-816)."""
    h = model._run_filter()
    K, T, M, D = model.K, model.T, model.M, model.D
    f64 = dict(dtype=torch.float64, device=model.device)
    sm = torch.empty((K, T, D), **f64)
    sP = torch.empty((K, T, D, D), **f64)
    sx = torch.empty((K, T, D, D), **f64)
    pm = torch.empty((K, T, M), **f64)
    pP = torch.empty((K, T, M, M), **f64)
    nat.check(nat.lib().bn_st_smoother_full(
        ctypes.byref(model._dims), ctypes.byref(h), nat.ptr(model._dt), nat.ptr(model._fm), nat.ptr(model._fP),
        nat.ptr(pm), nat.ptr(pP), nat.ptr(sm), nat.ptr(sP), nat.ptr(sx), nat.ptr(model._ws), nat.stream_ptr()))
    return _predict_from_state(model, h, X, R, sm, sP, sx)


class SpaceTimeVariationalGP(SpaceTimeMarkovGP):
    """MarkovVariationalGP with a SpatioTemporalKernel (models.py:118-132)."""
    method = nat.BN_VI


class SpaceTimeLaplaceGP(SpaceTimeMarkovGP):
    """MarkovLaplaceGP / MarkovNewtonGP with a SpatioTemporalKernel (models.py:297-299, 328)."""
    method = nat.BN_NEWTON


class SpaceTimeExpectationPropagationGP(SpaceTimeMarkovGP):
    """MarkovExpectationPropagationGP with a SpatioTemporalKernel (models.py:216-231): power
    EP whose cavity removes the projected sites B' diag(nat2) B, B' nat1 from the posterior
    of the inducing states (basemodels.py:230-262 with compute_full_pseudo_nat, 667-674)
    and is carried to the data before moment matching (inference.py:238-284).

    Several locations per step: the reference moment-matches a step's locations jointly;
    the Gaussian likelihood's closed form separates per location, which is what runs here,
    and other likelihoods are refused.  energy() follows the reference, whose EP energy
    refuses several locations per step (inference.py:313-317); with one location per step
    it is -(log Z + (sum log Z_tilted - sum lel_pseudo) / power) (inference.py:286-325)."""
    method = nat.BN_EP

    def __init__(self, kernel, likelihood, X, Y, R, power=1.0, parallel=None, device=None):
        self.power = float(power)
        if not 0.0 < self.power <= 1.0:
            raise ValueError("EP power must be in (0, 1]")
        super().__init__(kernel, likelihood, X, Y, R, parallel=parallel, device=device)
        self._check_lik()

    @classmethod
    def from_runs(cls, kernel, likelihood, Xs, Rs, Ys, power=1.0, parallel=None, device=None):
        self = cls.__new__(cls)
        self.power = float(power)
        self.parallel = parallel
        self._build(kernel, likelihood, Xs, Rs, Ys, device)
        self._single = False
        self._check_lik()
        return self

    def _check_lik(self):
        if self.n_loc > 1 and type(self.likelihood).__name__ != "Gaussian":
            raise NotImplementedError(
                "EP with several locations per step moment-matches them jointly (a product Gauss-Hermite rule over "
                "all of them in the reference); only the Gaussian likelihood has that in closed form")
        f64 = dict(dtype=torch.float64, device=self.device)
        self._cm = torch.empty((self.K, self.T, self.M), **f64)
        self._cP = torch.empty((self.K, self.T, self.M, self.M), **f64)

    def _sites(self, mode, lr=1.0, cubature=None, ensure_psd=True, terms=None, batch=None):
        lik = self.likelihood.desc()
        om, oj, oh = terms if terms is not None else (None, None, None)
        if mode == nat.BN_ST_PROJECT:
            # the projected pseudo observation does not depend on the inference method
            nat.check(nat.lib().bn_st_sites(
                ctypes.byref(self._dims), ctypes.byref(lik), nat.BN_VI, mode, cub.num_points(cubature), 1.0, 1,
                nat.ptr(self._y), nat.ptr(self._B), nat.ptr(self._cdiag), nat.ptr(self._pm), nat.ptr(self._pP), None,
                nat.ptr(self._nat1), nat.ptr(self._nat2), nat.ptr(self._Lw), nat.ptr(self._zw), nat.ptr(self._ldl),
                nat.ptr(self._p0), nat.ptr(self._p1), None, None, None, nat.stream_ptr()))
            return
        if batch is not None:
            raise NotImplementedError("mini-batched EP on the space-time models is not implemented")
        nat.check(nat.lib().bn_st_ep_sites(
            ctypes.byref(self._dims), ctypes.byref(lik), mode, cub.num_points(cubature), float(lr),
            1 if ensure_psd else 0, self.power, nat.ptr(self._y), nat.ptr(self._B), nat.ptr(self._cdiag),
            nat.ptr(self._pm), nat.ptr(self._pP), nat.ptr(self._nat1), nat.ptr(self._nat2), nat.ptr(self._Lw),
            nat.ptr(self._zw), nat.ptr(self._ldl), nat.ptr(self._cm), nat.ptr(self._cP), nat.ptr(self._p0),
            nat.ptr(self._p1), nat.ptr(om), nat.ptr(oj), nat.ptr(oh), nat.stream_ptr()))

    @property
    def cavity(self):
        """Cavity of the inducing states from the last site update or energy, [T, M] and
        [T, M, M] ([K, ...] for several runs)."""
        return self._out(self._cm), self._out(self._cP)

    def energy_per_run(self, cubature=None, batch_ind=None):
        if self._batch_mask(batch_ind) is not None:
            raise NotImplementedError("mini-batched EP on the space-time models is not implemented")
        if self.n_loc > 1:
            raise NotImplementedError("masking in spatio-temporal models not implemented for EP")
        self._ensure_loglik()
        self._sites(nat.BN_ST_ENERGY, cubature=cubature)
        lel = self._sum(self._p0, self._s0).clone()
        lel_pseudo = self._sum(self._p1, self._s1).clone()
        logz = self._loglik.clone()
        if self.M == self.n_loc:
            # the filter masks the log-likelihood of a missing pseudo observation
            # (mask_pseudo_y, basemodels.py:138-144, 698)
            miss = torch.isnan(self._y[..., 0])
            logz = logz - torch.where(miss, self._ell, torch.zeros_like(self._ell)).sum(1)
        return -(logz + (lel - lel_pseudo) / self.power)


class SpaceTimeMeanFieldGP(SpaceTimeMarkovGP):
    """MarkovMeanFieldGaussianProcess (basemodels.py:1158-1175): the same model with the
    mean-field filter and smoother (ops.py:581-608, 681-706), which keep only the M
    diagonal KT x KT blocks of the state covariance.  The projected pseudo observation is
    formed explicitly (noise = (B' diag(nat2) B + 1e-12 I)^-1, as compute_full_pseudo_lik
    does), then filtered.  parallel=None picks the associative-scan form, as the reference
    does on a GPU backend (basemodels.py:640-641); parallel=False the sequential form.  The
    two are different approximations when the pseudo noise couples the blocks."""

    def _alloc_filter(self):
        K, T, M = self.K, self.T, self.M
        KT = self.kernel.temporal_kernel.state_dim
        self.parallel = True if self.parallel is None else bool(self.parallel)
        f64 = dict(dtype=torch.float64, device=self.device)
        self._noise = torch.empty((K, T, M, M), **f64)
        self._ypse = torch.empty((K, T, M), **f64)
        self._fm = torch.empty((K, T, M, KT), **f64)
        self._fP = torch.empty((K, T, M, KT, KT), **f64)
        dts = self._dt.clone()
        dts[:, :-1] = self._dt[:, 1:]
        dts[:, -1] = 0.0
        self._dt_s = dts.contiguous()
        nb = nat.lib().bn_mf_workspace_bytes(ctypes.byref(self._dims))
        nat.check(0 if nb > 0 else -1)
        self._ws = torch.empty(int(nb), dtype=torch.uint8, device=self.device)
        self._pseudo_key = None

    def _run_filter(self):
        h, key = self._hyp()
        if self._pseudo_key != self._proj_key:
            nat.check(nat.lib().bn_mf_pseudo(
                ctypes.byref(self._dims), nat.ptr(self._Lw), nat.ptr(self._zw), nat.ptr(self._noise),
                nat.ptr(self._ypse), nat.stream_ptr()))
            self._pseudo_key = self._proj_key
        nat.check(nat.lib().bn_mf_filter(
            ctypes.byref(self._dims), ctypes.byref(h), 1 if self.parallel else 0, nat.ptr(self._dt),
            nat.ptr(self._ypse), nat.ptr(self._noise), nat.ptr(self._fm), nat.ptr(self._fP), nat.ptr(self._ell),
            nat.ptr(self._loglik), nat.ptr(self._ws), nat.stream_ptr()))
        self._ll_key = (self._proj_key, key[0], key[1])
        return h

    def predict(self, X=None, R=None):
        """MarkovMeanFieldGaussianProcess.predict (basemodels.py:1177-1228): the sequential
        mean-field filter whatever `parallel` is (as the reference), the smoother's full
        blocks, then the dense temporal conditional of the block-diagonal state (layout
        expanded to c*M + m) and the spatial conditional.  Returns [n_test, n_loc] mean
        and variance."""
        self._ensure_projection()
        h, key = self._hyp()
        if self._pseudo_key != self._proj_key:
            nat.check(nat.lib().bn_mf_pseudo(
                ctypes.byref(self._dims), nat.ptr(self._Lw), nat.ptr(self._zw), nat.ptr(self._noise),
                nat.ptr(self._ypse), nat.stream_ptr()))
            self._pseudo_key = self._proj_key
        K, T, M, D = self.K, self.T, self.M, self.D
        KT = self.kernel.temporal_kernel.state_dim
        f64 = dict(dtype=torch.float64, device=self.device)
        fm = torch.empty((K, T, M, KT), **f64)
        fP = torch.empty((K, T, M, KT, KT), **f64)
        ell = torch.empty((K, T), **f64)
        ll = torch.empty(K, **f64)
        nat.check(nat.lib().bn_mf_filter(
            ctypes.byref(self._dims), ctypes.byref(h), 0, nat.ptr(self._dt), nat.ptr(self._ypse), nat.ptr(self._noise),
            nat.ptr(fm), nat.ptr(fP), nat.ptr(ell), nat.ptr(ll), None, nat.stream_ptr()))
        bm = torch.empty((K, T, M, KT), **f64)
        bP = torch.empty((K, T, M, KT, KT), **f64)
        G = torch.empty((K, T, M, KT, KT), **f64)
        nat.check(nat.lib().bn_mf_smoother(
            ctypes.byref(self._dims), ctypes.byref(h), nat.ptr(self._dt_s), nat.ptr(fm), nat.ptr(fP), None, None, 0, 0,
            nat.ptr(bm), nat.ptr(bP), nat.ptr(G), nat.stream_ptr()))
        # cross covariance of neighbours per block: G_t P_s,t+1; at the end G = I and the
        # right neighbour is the sentinel (Pinf)
        Cx = torch.empty_like(bP)
        Cx[:, :-1] = G[:, :-1] @ bP[:, 1:]
        Cx[:, -1] = torch.tensor(_pinf_temporal(self.kernel.temporal_kernel), **f64)
        idx = torch.arange(M, device=self.device)
        sm = bm.permute(0, 1, 3, 2).reshape(K, T, D).contiguous()
        sP = torch.zeros((K, T, D, D), **f64)
        sx = torch.zeros((K, T, D, D), **f64)
        for c in range(KT):
            for c2 in range(KT):
                sP[:, :, c * M + idx, c2 * M + idx] = bP[:, :, :, c, c2]
                sx[:, :, c * M + idx, c2 * M + idx] = Cx[:, :, :, c, c2]
        return _predict_from_state(self, h, X, R, sm, sP, sx)

    def update_posterior(self):
        self._ensure_projection()
        h = self._run_filter()
        M = self.M
        # h P_s h' of block i lands on the diagonal of the [K,T,M,M] posterior covariance,
        # whose off-diagonal entries stay zero (mean-field posterior)
        nat.check(nat.lib().bn_mf_smoother(
            ctypes.byref(self._dims), ctypes.byref(h), nat.ptr(self._dt_s), nat.ptr(self._fm), nat.ptr(self._fP),
            nat.ptr(self._pm), nat.ptr(self._pP), M * M, M + 1, None, None, None, nat.stream_ptr()))
        self._post_key = self._ll_key

    @property
    def pseudo_likelihood_full(self):
        """(pseudo_y [T, M, 1], pseudo_var [T, M, M]) of compute_full_pseudo_lik."""
        self._ensure_projection()
        self._run_filter()
        return self._out(self._ypse[..., None]), self._out(self._noise)

    @property
    def filter_state(self):
        """Filtered block means [T, M, KT, 1] and covariances [T, M, KT, KT]."""
        return self._out(self._fm[..., None]), self._out(self._fP)


class SpaceTimeVariationalMeanFieldGP(SpaceTimeMeanFieldGP):
    """MarkovVariationalMeanFieldGP (models.py:154-156) with a SpatioTemporalKernel."""
    method = nat.BN_VI


class SpaceTimeLaplaceMeanFieldGP(SpaceTimeMeanFieldGP):
    """MarkovNewtonMeanFieldGP / MarkovLaplaceGPMeanField (models.py:307-309, 332)."""
    method = nat.BN_NEWTON


def _pinf_temporal(kt):
    """Stationary state covariance of the temporal Matern12 / Matern32 kernel
    (kernels.py:146-150, 212-216)."""
    v, ls = float(kt.variance), float(kt.lengthscale)
    if kt.state_dim == 1:
        return np.array([[v]])
    return np.array([[v, 0.0], [0.0, 3.0 * v / ls ** 2]])


def state_perm(M, KT):
    """Index map between the reference's state order (kron(I_M, .): entry m*KT + c) and
    this library's (c*M + m): x_here = x_ref[perm]."""
    return np.array([m * KT + c for c in range(KT) for m in range(M)], dtype=np.int64)


class SpaceTimeInfiniteHorizonGP(SpaceTimeMarkovGP):
    """InfiniteHorizonGaussianProcess (basemodels.py:1257-1301) with a SpatioTemporalKernel:
    equispaced steps, a steady-state covariance from `dare_iters` Riccati iterations per
    filter / smoother call, each started from the previous call's solution, as the
    reference keeps them (dare_init_filter / dare_init_smoother).  Like the reference,
    inference() runs the filter and smoother before and after the site update and every
    log-likelihood evaluation (energy) runs the filter again, so the Riccati solutions
    advance with each call.  Heteroscedastic (per-step gains) unless the likelihood is
    Gaussian with no missing data (basemodels.py:1270).  `parallel` is accepted: both of the
    reference's forms compute the same means."""

    def __init__(self, kernel, likelihood, X, Y, R, dare_iters=20, parallel=None, device=None):
        self.dare_iters = int(dare_iters)
        super().__init__(kernel, likelihood, X, Y, R, parallel=parallel, device=device)

    @classmethod
    def from_runs(cls, kernel, likelihood, Xs, Rs, Ys, dare_iters=20, parallel=None, device=None):
        self = cls.__new__(cls)
        self.dare_iters = int(dare_iters)
        self.parallel = parallel
        self._build(kernel, likelihood, Xs, Rs, Ys, device)
        self._single = False
        return self

    def _alloc_filter(self):
        K, T, M, D = self.K, self.T, self.M, self.D
        if T < 2:
            raise ValueError("the infinite-horizon model needs at least two time steps")
        dt = self._dt.cpu().numpy()
        if np.max(np.abs(np.diff(dt[:, 1:], axis=1)), initial=0.0) >= 1e-6:
            raise ValueError("the infinite-horizon model needs equispaced time steps")
        self.heteroscedastic = bool(np.any(np.isnan(self.Y))) or type(self.likelihood).__name__ != "Gaussian"
        f64 = dict(dtype=torch.float64, device=self.device)
        self._step = self._dt[:, 1].contiguous()
        self._noise = torch.empty((K, T, M, M), **f64)
        self._ypse = torch.empty((K, T, M), **f64)
        self._tied = torch.empty((K, M, M), **f64)
        pinf = np.kron(_pinf_temporal(self.kernel.temporal_kernel), np.eye(M))
        self._Xf = torch.tensor(pinf, **f64).expand(K, D, D).contiguous()
        self._Xs = self._Xf.clone()
        self._cov = torch.empty((K, D, D), **f64)
        self._gain = torch.empty((K, D, D), **f64)
        self._fm = torch.empty((K, T, D), **f64)
        nb = nat.lib().bn_ih_workspace_bytes(ctypes.byref(self._dims), 1 if self.heteroscedastic else 0)
        nat.check(0 if nb > 0 else -1)
        self._ws = torch.empty(int(nb), dtype=torch.uint8, device=self.device)
        self._pseudo_key = None

    def _run_filter(self):
        h, key = self._hyp()
        if self._pseudo_key != self._proj_key:
            nat.check(nat.lib().bn_mf_pseudo(
                ctypes.byref(self._dims), nat.ptr(self._Lw), nat.ptr(self._zw), nat.ptr(self._noise),
                nat.ptr(self._ypse), nat.stream_ptr()))
            nat.check(nat.lib().bn_ih_tied(ctypes.byref(self._dims), nat.ptr(self._Lw), nat.ptr(self._tied),
                                           nat.ptr(self._ws), nat.stream_ptr()))
            self._pseudo_key = self._proj_key
        nat.check(nat.lib().bn_ih_filter(
            ctypes.byref(self._dims), ctypes.byref(h), nat.ptr(self._step), self.dare_iters,
            1 if self.heteroscedastic else 0, nat.ptr(self._ypse), nat.ptr(self._noise), nat.ptr(self._tied),
            nat.ptr(self._Xf), nat.ptr(self._cov), nat.ptr(self._fm), nat.ptr(self._ell), nat.ptr(self._loglik),
            nat.ptr(self._ws), nat.stream_ptr()))
        self._ll_key = (self._proj_key, key[0], key[1])
        return h

    def update_posterior(self):
        self._ensure_projection()
        h = self._run_filter()
        nat.check(nat.lib().bn_ih_smoother(
            ctypes.byref(self._dims), ctypes.byref(h), nat.ptr(self._step), self.dare_iters, nat.ptr(self._fm),
            nat.ptr(self._Xf), nat.ptr(self._cov), nat.ptr(self._Xs), nat.ptr(self._gain), nat.ptr(self._pm),
            nat.ptr(self._pP), None, nat.ptr(self._ws), nat.stream_ptr()))
        self._post_key = self._ll_key

    def predict(self, X=None, R=None):
        """MarkovGaussianProcess.predict with temporal_conditional_infinite_horizon
        (basemodels.py:766-816, 1303-1308; utils.py:142-177): the filter and the full
        smoother (both advance their Riccati solutions, as in the reference), the mean by
        the conditional on the neighbouring smoothed means, the smoother's steady covariance
        as the state covariance.  Returns [n_test, n_loc] mean and variance."""
        self._ensure_projection()
        h = self._run_filter()
        f64 = dict(dtype=torch.float64, device=self.device)
        sm = torch.empty((self.K, self.T, self.D), **f64)
        nat.check(nat.lib().bn_ih_smoother(
            ctypes.byref(self._dims), ctypes.byref(h), nat.ptr(self._step), self.dare_iters, nat.ptr(self._fm),
            nat.ptr(self._Xf), nat.ptr(self._cov), nat.ptr(self._Xs), nat.ptr(self._gain), None, None, nat.ptr(sm),
            nat.ptr(self._ws), nat.stream_ptr()))
        return _predict_from_state(self, h, X, R, sm, steady=self._Xs)

    # the reference re-runs the filter (and advances its Riccati solution) on every call
    def _ensure_posterior(self):
        self.update_posterior()

    def _ensure_loglik(self):
        self._ensure_projection()
        self._run_filter()

    def energy_and_grad_per_run(self, rel_step=1e-5, cubature=None):
        raise NotImplementedError("hyperparameter gradients of the infinite-horizon models are not implemented: "
                                  "every energy evaluation advances the Riccati solutions (basemodels.py:1276-1298)")

    @property
    def dare_solutions(self):
        """(filter Pdare, smoother covariance) in the reference's state order, [D, D] each
        ([K, D, D] for several runs)."""
        p = torch.as_tensor(state_perm(self.M, self.kernel.temporal_kernel.state_dim), device=self.device)
        ip = torch.argsort(p)
        f = lambda X: X[:, ip][:, :, ip]  # noqa: E731
        return self._out(f(self._Xf)), self._out(f(self._Xs))


class SpaceTimeVariationalInfiniteHorizonGP(SpaceTimeInfiniteHorizonGP):
    """InfiniteHorizonVariationalGP (models.py:162-164) with a SpatioTemporalKernel."""
    method = nat.BN_VI


class SpaceTimeLaplaceInfiniteHorizonGP(SpaceTimeInfiniteHorizonGP):
    """InfiniteHorizonNewtonGP / InfiniteHorizonLaplaceGP (models.py:315-317, 336) with a
    SpatioTemporalKernel."""
    method = nat.BN_NEWTON
