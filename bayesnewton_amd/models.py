"""Markov-GP models on the HIP path (mirrors bayesnewton/models.py:118-132,
216-231, 297-299, 328 and their bases basemodels.py:625-764 /
inference.py:65-325).

A model holds every per-step array on the GPU in the plan's tiled layout:
time steps, observations, the sites (natural parameters in FP64, pseudo
observations in FP32) and the posterior.  `update_posterior` is one fused
filter + smoother pass; `inference` is the reference's site step between two
posterior updates; `energy` reuses the filter log-likelihood of the current
sites (compute_log_lik) instead of running the filter again.

Besides the reference's single-series constructor, `from_series` builds one
model over many independent series of uneven lengths (one launch per pass for
the whole batch), optionally with per-series kernel hyperparameters.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from . import cubature as cub
from . import ops
from . import spacetime
from .kernels import SpatioTemporalKernel
from .plan import SeriesPlan, default_chunk_len


def input_admin(t, y):
    """bayesnewton/utils.py:234-265 for the temporal case: sort by t, dt_0 = 0."""
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    if t.shape[0] != y.shape[0]:
        raise ValueError("X and Y must have the same number of rows")
    ind = np.argsort(t, kind="stable")
    t, y = t[ind], y[ind]
    dt = np.concatenate([[0.0], np.diff(t)])
    return t, y, dt


class PseudoLikelihood:
    """The sites (GaussianDistribution, basemodels.py:52-100): nat2 = 1 / variance."""

    def __init__(self, model):
        self._m = model

    @property
    def nat1(self):
        return self._m.plan.scatter(self._m._nat1[None]).reshape(-1, 1, 1)

    @property
    def nat2(self):
        return self._m.plan.scatter(self._m._nat2[None]).reshape(-1, 1, 1)

    @property
    def mean(self):
        return self.nat1 / self.nat2

    @property
    def covariance(self):
        return 1.0 / self.nat2


class MarkovGaussianProcess:
    """Sum-of-Matern state-space GP with scalar sites (basemodels.py:625-764)."""
    method = None

    def __new__(cls, kernel=None, likelihood=None, X=None, Y=None, R=None, *args, **kwargs):
        # basemodels.py:646-650: a spatio-temporal kernel with locations R selects the
        # space-time branch of the model (bayesnewton_amd.spacetime)
        if isinstance(kernel, SpatioTemporalKernel):
            if R is None:
                raise ValueError("a SpatioTemporalKernel needs the locations R [T, n_loc, 2]")
            if cls.method == nat.BN_EP:
                power = kwargs.get("power", args[0] if args else 1.0)
                return spacetime.SpaceTimeExpectationPropagationGP(kernel, likelihood, X, Y, R, power=power,
                                                                   device=kwargs.get("device"))
            st_cls = {nat.BN_VI: spacetime.SpaceTimeVariationalGP,
                      nat.BN_NEWTON: spacetime.SpaceTimeLaplaceGP}.get(cls.method)
            if st_cls is None:
                raise NotImplementedError("space-time models support VI, Newton / Laplace and EP")
            return st_cls(kernel, likelihood, X, Y, R, device=kwargs.get("device"))
        return super().__new__(cls)

    def __init__(self, kernel, likelihood, X, Y, R=None, parallel=None, power=1.0, chunk_len=None, device=None):
        if R is not None or (np.ndim(X) > 1 and np.shape(X)[1] > 1):
            raise NotImplementedError("spatio-temporal inputs need a SpatioTemporalKernel")
        t, y, dt = input_admin(X, Y)
        self._build(kernel, likelihood, [t], [y], [dt], None, None, parallel, power, chunk_len, device)

    @classmethod
    def from_series(cls, kernel, likelihood, xs, ys, lengthscale=None, variance=None, parallel=True, power=1.0,
                    chunk_len=None, device=None):
        """One model over B independent series (lists of per-series inputs / outputs).

        lengthscale / variance: optional per-series hypers [B, n_comp] (natural
        scale); default: the kernel's own hypers for every series."""
        self = cls.__new__(cls)
        ts, yy, dts = [], [], []
        for x, y in zip(xs, ys):
            t, yv, d = input_admin(x, y)
            ts.append(t)
            yy.append(yv)
            dts.append(d)
        self._build(kernel, likelihood, ts, yy, dts, lengthscale, variance, parallel, power, chunk_len, device)
        return self

    # ---- construction --------------------------------------------------------------
    def _build(self, kernel, likelihood, ts, ys, dts, ls, var, parallel, power, chunk_len, device):
        if not torch.cuda.is_available():
            raise nat.NativeError("bayesnewton_amd models need a ROCm GPU")
        self.kernel = kernel
        self.likelihood = likelihood
        self.power = float(power)
        self.parallel = True if parallel is None else bool(parallel)
        lens = np.array([len(t) for t in ts], dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        self.offsets = off
        self.X = np.concatenate(ts)[:, None]
        self.Y = np.concatenate(ys)[:, None]
        self.dt = np.concatenate(dts)
        self.num_data = int(off[-1])
        self.num_series = len(ts)
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        if chunk_len is None:
            chunk_len = default_chunk_len(lens, self.parallel)
        self._desc = kernel.desc()
        n_comp = self._desc.n_comp
        if ls is not None or var is not None:
            B = self.num_series
            base_ls, base_var = [c.lengthscale for c in kernel.components], [c.variance for c in kernel.components]
            ls = np.broadcast_to(np.asarray(base_ls if ls is None else ls, np.float64).reshape(-1, n_comp), (B, n_comp))
            var = np.broadcast_to(np.asarray(base_var if var is None else var, np.float64).reshape(-1, n_comp),
                                  (B, n_comp))
            self._hyper_stride = n_comp
            self._ls = torch.tensor(np.ascontiguousarray(ls), dtype=torch.float64, device=dev)
            self._var = torch.tensor(np.ascontiguousarray(var), dtype=torch.float64, device=dev)
            self._per_series = True
        else:
            self._hyper_stride = 0
            self._per_series = False
            self._hyper_key = None
        self.autotune = None
        if isinstance(chunk_len, str):
            if chunk_len != "auto" or not self.parallel:
                raise ValueError("chunk_len must be an int, None or 'auto' (with the associative form)")
            # time the fused pass on this batch's own steps at the initial sites
            # (bayesnewton_amd.autotune); the values do not change the pass's cost
            from . import autotune
            if self._hyper_stride:
                hl, hv = self._ls, self._var
            else:
                hl, hv = kernel.hypers(dev)
            f32 = dict(dtype=torch.float32, device=dev)
            choice = autotune.tune_schedule(
                off, self._desc, hl, hv, torch.as_tensor(self.dt, **f32), torch.zeros(self.num_data, **f32),
                torch.full((self.num_data,), 1e2, **f32), hyper_stride=self._hyper_stride, device=dev)
            chunk_len = int(choice["chunk_len"])
            self.autotune = choice
        self.plan = SeriesPlan(off, chunk_len=chunk_len, device=dev)
        P = self.plan
        self._dt = P.gather(torch.as_tensor(self.dt, dtype=torch.float32, device=dev))
        y32 = torch.as_tensor(self.Y[:, 0], dtype=torch.float32, device=dev)
        self._y = P.gather(y32, pad=float("nan"))[0]
        self._mask = P.gather_mask(torch.isnan(y32))
        # sites N(0, 1e2) (basemodels.py:130-133): nat1 = 0, nat2 = 1e-2
        self._nat1 = torch.zeros(P.n_slots, dtype=torch.float64, device=dev)
        self._nat2 = torch.full((P.n_slots,), 1e-2, dtype=torch.float64, device=dev)
        self._py = torch.empty(P.n_slots, dtype=torch.float32, device=dev)
        self._pr = torch.empty(P.n_slots, dtype=torch.float32, device=dev)
        nat.check(nat.lib().bn_site_reparametrise(P.handle, nat.ptr(self._nat1), nat.ptr(self._nat2),
                                                  nat.ptr(self._py), nat.ptr(self._pr), nat.stream_ptr()))
        self._pm = torch.zeros(P.n_slots, dtype=torch.float32, device=dev)
        self._pv = torch.ones(P.n_slots, dtype=torch.float32, device=dev)
        self._loglik = torch.zeros(self.num_series, dtype=torch.float64, device=dev)
        self._status = torch.zeros(self.num_series, dtype=torch.int32, device=dev)
        self._post_key = None  # state (hypers, sites) the stored posterior was computed from
        self._ll_key = None    # state the stored log-likelihood was computed from
        self._d1 = torch.empty(self.num_series, dtype=torch.float64, device=dev)
        self._d2 = torch.empty(self.num_series, dtype=torch.float64, device=dev)
        self.pseudo_likelihood = PseudoLikelihood(self)

    # ---- hyperparameters ---------------------------------------------------------
    def _hypers(self):
        if self._per_series:
            return self._ls, self._var
        key = tuple((c.lengthscale, c.variance) for c in self.kernel.components)
        if key != self._hyper_key:
            self._ls, self._var = self.kernel.hypers(self.device)
            self._hyper_key = key
            self._hyper_version += 1
        return self._ls, self._var

    def _state_key(self):
        """Identity of the (kernel hypers, sites) state a stored posterior or
        log-likelihood belongs to: two counters bumped whenever the hyper tensors
        are replaced or the sites change (tensor ids are reused by the allocator,
        so they cannot serve as the key)."""
        self._hypers()
        return (self._hyper_version, self._sites_version)

    _sites_version = 0
    _hyper_version = 0

    # ---- posterior -----------------------------------------------------------------
    # Replay the fused pass from a hipGraph instead of launching it eagerly: the pass's
    # launches (8 with the associative scan) become one graph launch.  Every buffer it
    # touches lives as long as the model; the graph is recorded again when the hyper
    # tensors are replaced.
    use_graph = False
    _graph = None
    _graph_key = None

    def _fused_pass(self):
        ls, var = self._hypers()
        ops.filter_smoother(self.plan, self._desc, ls, var, self._dt, self._py, self._pr, self._mask,
                            hyper_stride=self._hyper_stride, post_mean=self._pm, post_var=self._pv,
                            loglik=self._loglik, status=self._status)

    @property
    def status(self):
        """BN_STATUS_* bits of every series from the last posterior update (numpy int32
        [num_series]; 0 = no numerical failure): a non-PD innovation, a non-finite
        log-likelihood or filtered mean, a non-positive smoothed variance.  The affected
        series' outputs carry NaN as in the reference; the others are unaffected."""
        return self._status.cpu().numpy()

    def update_posterior(self):
        """basemodels.py:689-706: filter + smoother of the current sites."""
        self._hypers()
        if self.use_graph:
            if self._graph is None or self._graph_key != self._hyper_version:
                self.plan.workspace(self._desc)  # allocated before capture
                self._graph = ops.capture_graph(self._fused_pass)
                self._graph_key = self._hyper_version
            self._graph.replay()
        else:
            self._fused_pass()
        self._post_key = self._ll_key = self._state_key()

    def _ensure_posterior(self):
        # inference() starts with update_posterior() (inference.py:70); skipping it
        # when the stored posterior already belongs to the current state is exact.
        if self._post_key != self._state_key():
            self.update_posterior()

    def _ensure_loglik(self):
        key = self._state_key()
        if self._ll_key != key:
            ls, var = self._hypers()
            ops.filter_loglik(self.plan, self._desc, ls, var, self._dt, self._py, self._pr, self._mask,
                              hyper_stride=self._hyper_stride, loglik=self._loglik)
            self._ll_key = key

    def compute_log_lik(self, per_series=False):
        """basemodels.py:726-741: log normaliser of the pseudo model (filter log p(y))."""
        self._ensure_loglik()
        return self._loglik.clone() if per_series else self._loglik.sum()

    @property
    def posterior_mean(self):
        return self.plan.scatter(self._pm[None]).to(torch.float64).reshape(-1, 1, 1)

    @property
    def posterior_variance(self):
        return self.plan.scatter(self._pv[None]).to(torch.float64).reshape(-1, 1, 1)

    # ---- inference (inference.py:65-90) ----------------------------------------------
    def _batch_mask(self, batch_ind):
        """Tiled mask of a mini-batch (indices into the model's time-sorted data; for a
        model over several series, into their concatenation in series order), or None
        for the whole data set (inference.py:67-68)."""
        if batch_ind is None:
            return None
        idx = np.asarray(batch_ind).reshape(-1)
        if idx.shape[0] == self.num_data:
            return None
        if idx.size == 0:
            raise ValueError("empty mini-batch")
        if idx.min() < 0 or idx.max() >= self.num_data:
            raise IndexError("batch_ind out of range")
        m = torch.zeros(self.num_data, dtype=torch.float32, device=self.device)
        m[torch.as_tensor(idx, dtype=torch.int64, device=self.device)] = 1.0
        return self.plan.gather(m)[0]

    def inference(self, lr=1.0, batch_ind=None, cubature=None, ensure_psd=True, return_terms=True, **kwargs):
        """inference.py:65-90.  With batch_ind, only the sites of those data points are
        updated (the others keep their values, group_natural_params), diff1 / diff2 are
        means over all N sites, and the returned terms are those of the batch [B, 1, 1]."""
        batch = self._batch_mask(batch_ind)
        self._ensure_posterior()
        P = self.plan
        dev = self.device
        if return_terms:
            om = torch.empty(P.n_slots, dtype=torch.float32, device=dev)
            oj = torch.empty_like(om)
            oh = torch.empty_like(om)
        else:
            om = oj = oh = None
        lik = self.likelihood.desc()
        nat.check(nat.lib().bn_site_update(
            P.handle, ctypes.byref(self._desc), ctypes.byref(lik), self.method, cub.num_points(cubature), float(lr),
            self.power, 1 if ensure_psd else 0, nat.ptr(self._y), nat.ptr(self._pm), nat.ptr(self._pv),
            nat.ptr(batch), nat.ptr(self._nat1), nat.ptr(self._nat2), nat.ptr(self._py), nat.ptr(self._pr),
            nat.ptr(om),
            nat.ptr(oj), nat.ptr(oh), nat.ptr(self._d1), nat.ptr(self._d2), nat.ptr(P.workspace(self._desc)),
            nat.stream_ptr()))
        self._sites_version += 1
        self.update_posterior()
        diff1 = self._d1.sum() / self.num_data
        diff2 = self._d2.sum() / self.num_data
        if return_terms:
            terms = tuple(P.scatter(a[None]).to(torch.float64).reshape(-1, 1, 1) for a in (om, oj, oh))
            if batch is not None:
                sel = torch.as_tensor(np.asarray(batch_ind).reshape(-1), dtype=torch.int64, device=dev)
                terms = tuple(a[sel] for a in terms)
        else:
            terms = (None, None, None)
        return terms, (diff1, diff2)

    # ---- energy -------------------------------------------------------------------
    def energy_per_series(self, cubature=None, batch_ind=None):
        """Energy of every series (the reference's energy() for each one).  As in the
        reference, the stored posterior is used as is and only log Z (the filter
        log-likelihood of the sites under the current hypers) is recomputed.  With
        batch_ind the data term sums the batch and is scaled by N / B (inference.py:
        134-138, 203-207, 292-296); the pseudo terms and log Z keep every site."""
        batch = self._batch_mask(batch_ind)
        scale = 1.0 if batch is None else self.num_data / float(np.asarray(batch_ind).size)
        self._ensure_loglik()
        P = self.plan
        lik = self.likelihood.desc()
        t1 = torch.empty(self.num_series, dtype=torch.float64, device=self.device)
        t2 = torch.empty_like(t1)
        nat.check(nat.lib().bn_energy_terms(
            P.handle, ctypes.byref(self._desc), ctypes.byref(lik), self.method, cub.num_points(cubature), self.power,
            nat.ptr(self._y), nat.ptr(self._pm), nat.ptr(self._pv), nat.ptr(batch), nat.ptr(self._nat1),
            nat.ptr(self._nat2), nat.ptr(t1), nat.ptr(t2), nat.ptr(P.workspace(self._desc)), nat.stream_ptr()))
        logZ = self._loglik
        if self.method == nat.BN_EP:
            return -(logZ + (scale * t1 - t2) / self.power)
        return -(scale * t1 - (t2 - logZ))

    def energy(self, batch_ind=None, cubature=None, **kwargs):
        """inference.py:130-154 (Newton), 197-222 (VI), 286-325 (EP)."""
        return self.energy_per_series(cubature, batch_ind).sum()


    # ---- prediction (basemodels.py:766-816, 165-193) ----------------------------------------
    def _full_posterior(self):
        """Filtered, then smoothed full state of the current sites: means [d] and
        covariances / gains [d*d] as tiled FP64 planes (the filter and smoother of
        predict(), basemodels.py:779-792, with the smoother's shifted dt)."""
        key = self._state_key()
        if getattr(self, "_full_key", None) == key:
            return self._full
        ls, var = self._hypers()
        P = self.plan
        d = self.kernel.state_dim
        dev = self.device
        if getattr(self, "_dt_shift", None) is None:
            sh = np.zeros(self.num_data, dtype=np.float64)
            for b in range(self.num_series):
                a, e = self.offsets[b], self.offsets[b + 1]
                sh[a:e - 1] = self.dt[a + 1:e]
            self._dt_shift = P.gather(torch.as_tensor(sh, dtype=torch.float32, device=dev))
            self._x_train = torch.as_tensor(self.X[:, 0], dtype=torch.float64, device=dev)
            self._train_off = torch.as_tensor(self.offsets, dtype=torch.int64, device=dev)
        fm = torch.empty((d, P.n_slots), dtype=torch.float64, device=dev)
        fP = torch.empty((d * d, P.n_slots), dtype=torch.float64, device=dev)
        ll = torch.empty(self.num_series, dtype=torch.float64, device=dev)
        ws = P.workspace(self._desc)
        nat.check(nat.lib().bn_kalman_filter(P.handle, ctypes.byref(self._desc), nat.ptr(ls), nat.ptr(var),
                                             self._hyper_stride, nat.ptr(self._dt), nat.ptr(self._py),
                                             nat.ptr(self._pr), nat.ptr(self._mask), nat.ptr(fm), nat.ptr(fP),
                                             nat.ptr(ll), 0, None, nat.ptr(ws), nat.stream_ptr()))
        sm = torch.empty_like(fm)
        sP = torch.empty_like(fP)
        G = torch.empty_like(fP)
        nat.check(nat.lib().bn_rts_smoother(P.handle, ctypes.byref(self._desc), nat.ptr(ls), nat.ptr(var),
                                            self._hyper_stride, nat.ptr(self._dt_shift), nat.ptr(fm), nat.ptr(fP),
                                            nat.ptr(sm), nat.ptr(sP), nat.ptr(G), nat.BN_RETURN_FULL, nat.ptr(ws),
                                            nat.stream_ptr()))
        self._full, self._full_key = (sm, sP, G), key
        return self._full

    def _test_points(self, X):
        """Test times as one FP64 device array in series order plus CSR offsets: X is an
        array (single-series model, [n] or [n, 1]; only column 0, the time, is used) or a
        list with one array per series."""
        if self.num_series == 1 and not isinstance(X, (list, tuple)):
            X = [X]
        if len(X) != self.num_series:
            raise ValueError(f"expected test inputs for {self.num_series} series")
        xs = [np.asarray(x, dtype=np.float64) for x in X]
        xs = [x.reshape(len(x), -1)[:, 0] if x.size else np.zeros(0) for x in xs]  # a series may have none
        off = np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.int64)
        xt = np.concatenate(xs) if off[-1] > 0 else np.zeros(0)
        return (torch.as_tensor(xt, dtype=torch.float64, device=self.device),
                torch.as_tensor(off, dtype=torch.int64, device=self.device))

    def predict(self, X=None, R=None):
        """basemodels.py:766-816: predictive mean and variance of f at test times X
        (default: the training inputs), FP64 device tensors [n_test] in input order
        (series after series for a multi-series model)."""
        if R is not None:
            raise NotImplementedError("spatial inputs R are for the space-time models")
        if X is None:
            X = [self.X[self.offsets[b]:self.offsets[b + 1], 0] for b in range(self.num_series)]
        sm, sP, G = self._full_posterior()
        xt, off = self._test_points(X)
        n = int(xt.shape[0])
        mean = torch.empty(n, dtype=torch.float64, device=self.device)
        var = torch.empty_like(mean)
        ls, lv = self._hypers()
        nat.check(nat.lib().bn_predict(self.plan.handle, ctypes.byref(self._desc), nat.ptr(ls), nat.ptr(lv),
                                       self._hyper_stride, nat.ptr(self._train_off), nat.ptr(self._x_train),
                                       nat.ptr(sm), nat.ptr(sP), nat.ptr(G), nat.ptr(off), nat.ptr(xt), n,
                                       nat.ptr(mean), nat.ptr(var), nat.stream_ptr()))
        return mean, var

    def _lik_predict(self, mean_f, var_f, y=None, moments=True, cubature=None):
        n = int(mean_f.shape[0])
        my = torch.empty(n, dtype=torch.float64, device=self.device) if moments else None
        vy = torch.empty_like(my) if moments else None
        ld = torch.empty(n, dtype=torch.float64, device=self.device) if y is not None else None
        lik = self.likelihood.desc()
        nat.check(nat.lib().bn_lik_predict(ctypes.byref(lik), cub.num_points(cubature), n, nat.ptr(mean_f),
                                           nat.ptr(var_f), nat.ptr(y), nat.ptr(my), nat.ptr(vy), nat.ptr(ld),
                                           nat.stream_ptr()))
        return my, vy, ld

    def predict_y(self, X, R=None, cubature=None):
        """basemodels.py:165-175: predictive mean and variance of y at X."""
        mean_f, var_f = self.predict(X, R)
        my, vy, _ = self._lik_predict(mean_f, var_f, cubature=cubature)
        return my, vy

    def negative_log_predictive_density(self, X, Y, R=None, cubature=None):
        """basemodels.py:177-193: -nanmean log p(y* | data) over the test points (Y in the
        order of X; NaN entries are skipped)."""
        mean_f, var_f = self.predict(X, R)
        if self.num_series > 1 or isinstance(Y, (list, tuple)):
            Y = np.concatenate([np.asarray(v, dtype=np.float64).reshape(-1) for v in Y])
        y = torch.as_tensor(np.asarray(Y, dtype=np.float64).reshape(-1), dtype=torch.float64, device=self.device)
        if y.shape[0] != mean_f.shape[0]:
            raise ValueError("Y must hold one value per test input")
        _, _, ld = self._lik_predict(mean_f, var_f, y=y, moments=False, cubature=cubature)
        return float(-torch.nanmean(ld))

    # ---- hyperparameter gradients ----------------------------------------------------
    def _trainables(self):
        """(owner, attribute) of every trainable raw (softplus-inverse) hyperparameter, in
        the order objax lists the reference's TrainVars: per kernel component
        transformed_lengthscale then transformed_variance (kernels.py:80-88), then the
        likelihood's (likelihoods.py:696-706, 1019-1029)."""
        out = []
        if not self._per_series:
            for c in self.kernel.components:
                if not c.fix_lengthscale:
                    out.append((c, "transformed_lengthscale"))
                if not c.fix_variance:
                    out.append((c, "transformed_variance"))
        lik = self.likelihood
        for attr, fix in (("transformed_variance", "fix_variance"), ("transformed_scale", "fix_scale")):
            if hasattr(lik, attr) and not getattr(lik, fix, False):
                out.append((lik, attr))
        return out

    def trainable_names(self):
        return [f"{type(o).__name__}.{a}" for o, a in self._trainables()]

    def get_raw_hypers(self):
        return np.array([getattr(o, a) for o, a in self._trainables()], dtype=np.float64)

    def set_raw_hypers(self, raw):
        for (o, a), v in zip(self._trainables(), np.asarray(raw, dtype=np.float64).reshape(-1)):
            setattr(o, a, float(v))

    def energy_and_grad_per_series(self, cubature=None):
        """Energy of every series and its gradient w.r.t. the natural-scale hypers
        with the sites and posterior held fixed (the gradient objax.GradValues(energy)
        takes, README.md:56-76): (E [B], dE/d(ls_c, var_c) [B, 2 n_comp], dE/dphi [B])."""
        P = self.plan
        ls, var = self._hypers()
        nh = 2 * self._desc.n_comp
        gk = torch.empty((self.num_series, nh), dtype=torch.float64, device=self.device)
        gws = getattr(self, "_gws", None)
        if gws is None:
            nbytes = int(nat.lib().bn_grad_workspace_bytes(P.handle, ctypes.byref(self._desc)))
            gws = self._gws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
        nat.check(nat.lib().bn_filter_loglik_grad(
            P.handle, ctypes.byref(self._desc), nat.ptr(ls), nat.ptr(var), self._hyper_stride, nat.ptr(self._dt),
            nat.ptr(self._py), nat.ptr(self._pr), nat.ptr(self._mask), nat.ptr(self._loglik), nat.ptr(gk),
            nat.ptr(P.workspace(self._desc)), nat.ptr(gws), nat.stream_ptr()))
        self._ll_key = self._state_key()
        E = self.energy_per_series(cubature)
        gl = torch.empty(self.num_series, dtype=torch.float64, device=self.device)
        lik = self.likelihood.desc()
        nat.check(nat.lib().bn_lik_param_grad(
            P.handle, ctypes.byref(self._desc), ctypes.byref(lik), self.method, cub.num_points(cubature),
            self.power, nat.ptr(self._y), nat.ptr(self._pm), nat.ptr(self._pv), nat.ptr(self._nat1),
            nat.ptr(self._nat2), nat.ptr(gl), nat.ptr(P.workspace(self._desc)), nat.stream_ptr()))
        # energy = -(term1 - term2 + log Z): d/d kernel = -dlogZ, d/d likelihood = -dterm1 (for EP
        # term1 is the moment-matched log Z over the power, inference.py:286-325)
        return E, -gk, -gl

    def energy_and_grad(self, cubature=None):
        """(energy, dE/d raw hypers) for the trainable hyperparameters, in the order of
        trainable_names() (natural-scale gradients chained through softplus)."""
        E, gk, gl = self.energy_and_grad_per_series(cubature)
        gk = gk.sum(0).cpu().numpy()
        gl = float(gl.sum())
        out = []
        comps = list(self.kernel.components)
        for o, a in self._trainables():
            sig = 1.0 / (1.0 + np.exp(-getattr(o, a)))  # d softplus / d raw
            if o in comps:
                c = comps.index(o)
                g = gk[2 * c] if a == "transformed_lengthscale" else gk[2 * c + 1]
            else:
                g = gl
            out.append(g * sig)
        return float(E.sum()), np.array(out)


MarkovGP = MarkovGaussianProcess


class MarkovEnergy(torch.autograd.Function):
    """energy(raw hypers) as a differentiable torch function: forward sets the raw
    hyperparameters and returns the energy, backward returns the hand-written
    gradient (forward-mode tangents through the chunked filter scan).

    The reference's training step (README.md:56-76) runs inference() on the
    optimiser's current hypers, then GradValues(energy), then Adam.  The optimiser
    updates `raw` in place, so call ``model.set_raw_hypers(raw)`` (or
    ``MarkovEnergy.sync(raw, model)``) after ``opt.step()`` and before the next
    ``inference()``; otherwise that inference runs one step behind."""

    @staticmethod
    def sync(raw, model):
        """Push the optimiser's raw hypers into the model."""
        model.set_raw_hypers(raw.detach().cpu().numpy())

    @staticmethod
    def forward(ctx, raw, model, cubature=None):
        model.set_raw_hypers(raw.detach().cpu().numpy())
        E, g = model.energy_and_grad(cubature)
        ctx.save_for_backward(torch.as_tensor(g, dtype=raw.dtype, device=raw.device))
        return torch.as_tensor(E, dtype=raw.dtype, device=raw.device)

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        return gout * g, None, None


class MarkovMeanFieldGaussianProcess:
    """basemodels.py:1158-1175: the space-time Markov GP with the mean-field filter and
    smoother (bayesnewton_amd.spacetime.SpaceTimeMeanFieldGP)."""
    method = None

    def __new__(cls, kernel=None, likelihood=None, X=None, Y=None, R=None, parallel=None, device=None):
        if not isinstance(kernel, SpatioTemporalKernel) or R is None:
            raise NotImplementedError("the mean-field model is implemented for a SpatioTemporalKernel with "
                                      "locations R [T, n_loc, 2]")
        st_cls = {nat.BN_VI: spacetime.SpaceTimeVariationalMeanFieldGP,
                  nat.BN_NEWTON: spacetime.SpaceTimeLaplaceMeanFieldGP}.get(cls.method)
        if st_cls is None:
            raise NotImplementedError("mean-field models support VI and Newton / Laplace")
        return st_cls(kernel, likelihood, X, Y, R, parallel=parallel, device=device)


MarkovMeanFieldGP = MarkovMeanFieldGaussianProcess


class MarkovVariationalMeanFieldGP(MarkovMeanFieldGaussianProcess):
    """models.py:154-156"""
    method = nat.BN_VI


class MarkovNewtonMeanFieldGP(MarkovMeanFieldGaussianProcess):
    """models.py:307-309"""
    method = nat.BN_NEWTON


MarkovLaplaceGPMeanField = MarkovNewtonMeanFieldGP


class InfiniteHorizonGaussianProcess:
    """basemodels.py:1257-1301: the infinite-horizon space-time Markov GP
    (bayesnewton_amd.spacetime.SpaceTimeInfiniteHorizonGP)."""
    method = None

    def __new__(cls, kernel=None, likelihood=None, X=None, Y=None, R=None, dare_iters=20, parallel=None,
                device=None):
        if not isinstance(kernel, SpatioTemporalKernel) or R is None:
            raise NotImplementedError("the infinite-horizon model is implemented for a SpatioTemporalKernel with "
                                      "locations R [T, n_loc, 2]")
        st_cls = {nat.BN_VI: spacetime.SpaceTimeVariationalInfiniteHorizonGP,
                  nat.BN_NEWTON: spacetime.SpaceTimeLaplaceInfiniteHorizonGP}.get(cls.method)
        if st_cls is None:
            raise NotImplementedError("infinite-horizon models support VI and Newton / Laplace")
        return st_cls(kernel, likelihood, X, Y, R, dare_iters=dare_iters, parallel=parallel, device=device)


class InfiniteHorizonVariationalGP(InfiniteHorizonGaussianProcess):
    """models.py:162-164"""
    method = nat.BN_VI


class InfiniteHorizonNewtonGP(InfiniteHorizonGaussianProcess):
    """models.py:315-317"""
    method = nat.BN_NEWTON


InfiniteHorizonLaplaceGP = InfiniteHorizonNewtonGP


class MarkovVariationalGP(MarkovGaussianProcess):
    """models.py:118-132 (VariationalInference + MarkovGaussianProcess)."""
    method = nat.BN_VI


class MarkovNewtonGP(MarkovGaussianProcess):
    """models.py:297-299 (Newton + MarkovGaussianProcess)."""
    method = nat.BN_NEWTON


MarkovLaplaceGP = MarkovNewtonGP


class MarkovExpectationPropagationGP(MarkovGaussianProcess):
    """models.py:216-231 (ExpectationPropagation + MarkovGaussianProcess, power EP)."""
    method = nat.BN_EP

    def __init__(self, kernel, likelihood, X, Y, R=None, power=1.0, parallel=None, chunk_len=None, device=None):
        super().__init__(kernel, likelihood, X, Y, R=R, parallel=parallel, power=power, chunk_len=chunk_len,
                         device=device)
