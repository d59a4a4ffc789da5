"""Learned warm start of the space-time sites.

The reference initialises every site as N(0, 1e2 I) (basemodels.py:130-133) and has no
convergence loop.  Here a small network, trained on converged sites of plain runs,
proposes initial sites; the plain iteration (inference.py:65-90) then runs from that
start to the same convergence test, so it stays both the fallback and the check on the
network's error.

Network: features (BN_WS_FEATURES per site: the site's count, kernel-smoothed log
rates of a local space-time window, the hypers) -> 3 x 128 SiLU -> (m, log v), the
converged marginal q(f) at the site.  The initial site is the VI site step at that
marginal (bn_ws_sites).  Weights are BF16 with FP32 accumulation in every product;
training is mini-batch Adam on FP32 master weights, all in libbn_hip.so
(bn_mlp_grad_step).
"""
import ctypes
import math
import os
import time

import numpy as np
import torch

from . import _native as nat

F = nat.BN_WS_FEATURES
P = nat.BN_WS_PARAMS
# network trained on converged sites of 200 plain runs of the space-time Cox process
# (M = 64 inducing points on an 8 x 8 grid, T = 2000, 256 locations per step, Poisson
# counts with binsize 2, Matern32 (lengthscale 10) x Matern32 (lengthscale 0.2))
PRETRAINED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pretrained", "spacetime_cox_warmstart.npz")
H = 128
# flat parameter layout of include/bn_hip.h
OFF = {"W1": (0, (F, H)), "b1": (F * H, (H,)), "W2": (F * H + H, (H, H)), "b2": (F * H + H + H * H, (H,)),
       "W3": (F * H + 2 * H + H * H, (H, H)), "b3": (F * H + 2 * H + 2 * H * H, (H,)),
       "W4": (F * H + 3 * H + 2 * H * H, (H, 2)), "b4": (F * H + 3 * H + 2 * H * H + 2 * H, (2,))}


def site_features(model):
    """[K*T*R, F] features of every site of a space-time model (bn_ws_features)."""
    kt, ks = model.kernel.temporal_kernel, model.kernel.spatial_kernel
    binsize = getattr(model.likelihood, "binsize", 1.0)
    t = torch.tensor(model.X, dtype=torch.float64, device=model.device)
    X = torch.empty((model.K * model.T * model.n_loc, F), dtype=torch.float32, device=model.device)
    nat.check(nat.lib().bn_ws_features(ctypes.byref(model._dims), nat.ptr(model._r), nat.ptr(model._y), nat.ptr(t),
                                       kt.lengthscale, ks.lengthscale, kt.variance, float(binsize), nat.ptr(X),
                                       nat.stream_ptr()))
    return X


def site_marginals(model):
    """Converged-marginal targets [K*T*R, 2] = (mean_f, log var_f) of the model's posterior."""
    model._ensure_conditional()
    model._ensure_posterior()
    shape = (model.K, model.T, model.n_loc)
    mf = torch.empty(shape, dtype=torch.float64, device=model.device)
    vf = torch.empty_like(mf)
    nat.check(nat.lib().bn_st_marginals(ctypes.byref(model._dims), nat.ptr(model._B), nat.ptr(model._cdiag),
                                        nat.ptr(model._pm), nat.ptr(model._pP), nat.ptr(mf), nat.ptr(vf),
                                        nat.stream_ptr()))
    return torch.stack([mf.reshape(-1), torch.log(vf.reshape(-1))], 1).to(torch.float32)


class WarmStartMLP:
    def __init__(self, device=None, seed=0):
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        rng = np.random.default_rng(seed)
        p = np.zeros(P, dtype=np.float32)
        for name, fan_in in (("W1", F), ("W2", H), ("W3", H), ("W4", H)):
            o, shp = OFF[name]
            p[o:o + int(np.prod(shp))] = (rng.standard_normal(int(np.prod(shp))) * math.sqrt(1.0 / fan_in))
        self.params = torch.tensor(p, device=dev)
        self.wbf = torch.empty(P, dtype=torch.int16, device=dev)
        nat.check(nat.lib().bn_mlp_to_bf16(nat.ptr(self.params), nat.ptr(self.wbf), nat.stream_ptr()))
        self.norm = torch.tensor(np.concatenate([np.zeros(F), np.ones(F), np.zeros(2), np.ones(2)]),
                                 dtype=torch.float32, device=dev)
        self.m1 = torch.zeros(P, dtype=torch.float32, device=dev)
        self.m2 = torch.zeros(P, dtype=torch.float32, device=dev)
        self.step = 0

    # ---- normalisation ---------------------------------------------------------------
    def set_norm(self, X, Y):
        xmu, xsd = X.mean(0), X.std(0)
        ymu, ysd = Y.mean(0), Y.std(0)
        xsd = torch.where(xsd > 1e-6, xsd, torch.ones_like(xsd))  # constant features (the hypers)
        ysd = torch.where(ysd > 1e-6, ysd, torch.ones_like(ysd))
        self.norm = torch.cat([xmu, xsd, ymu, ysd]).to(torch.float32).contiguous()

    # ---- forward / gradient ------------------------------------------------------------
    def predict(self, X):
        out = torch.empty((X.shape[0], 2), dtype=torch.float32, device=self.device)
        nat.check(nat.lib().bn_mlp_forward(X.shape[0], nat.ptr(X), nat.ptr(self.wbf), nat.ptr(self.params),
                                           nat.ptr(self.norm), nat.ptr(out), nat.stream_ptr()))
        return out

    def _ws(self, n):
        nb = nat.lib().bn_mlp_grad_workspace_bytes(n)
        if getattr(self, "_wsb", None) is None or self._wsb.numel() < nb:
            self._wsb = torch.empty(int(nb), dtype=torch.uint8, device=self.device)
        return self._wsb

    def gradient(self, X, Y):
        """(dL/dparams [P], L) of L = sum (out - y)^2 / (2 n) on standardised outputs."""
        g = torch.empty(P, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        nat.check(nat.lib().bn_mlp_grad_step(X.shape[0], nat.ptr(X), nat.ptr(Y), nat.ptr(self.params),
                                             nat.ptr(self.wbf), nat.ptr(self.norm), None, None, 0, 0.0, nat.ptr(g),
                                             nat.ptr(loss), nat.ptr(self._ws(X.shape[0])), nat.stream_ptr()))
        return g, loss

    def fit(self, X, Y, epochs=4, batch=8192, lr=2e-3, seed=0, log=None):
        """Mini-batch Adam over (X, Y) (device tensors [n, F], [n, 2])."""
        self.set_norm(X, Y)
        n = X.shape[0]
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        ws = self._ws(batch)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen, device=self.device)
            tot, cnt = 0.0, 0
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                xb = X.index_select(0, idx).contiguous()
                yb = Y.index_select(0, idx).contiguous()
                self.step += 1
                nat.check(nat.lib().bn_mlp_grad_step(batch, nat.ptr(xb), nat.ptr(yb), nat.ptr(self.params),
                                                     nat.ptr(self.wbf), nat.ptr(self.norm), nat.ptr(self.m1),
                                                     nat.ptr(self.m2), self.step, float(lr), None, nat.ptr(loss),
                                                     nat.ptr(ws), nat.stream_ptr()))
                if cnt % 50 == 0:
                    tot += float(loss)
                cnt += 1
            hist.append(tot / max(1, (cnt + 49) // 50))
            if log is not None:
                log(f"[warmstart] epoch {ep}: loss {hist[-1]:.4f}")
        return hist

    # ---- use -------------------------------------------------------------------------
    def init_sites(self, model):
        """Replace the model's sites by the network's proposal (update_nat_params)."""
        X = site_features(model)
        pred = self.predict(X)
        lik = model.likelihood.desc()
        nat.check(nat.lib().bn_ws_sites(X.shape[0], ctypes.byref(lik), nat.ptr(model._y), nat.ptr(pred),
                                        nat.ptr(model._nat1), nat.ptr(model._nat2), nat.stream_ptr()))
        model._sites_version += 1

    @classmethod
    def pretrained(cls, path=PRETRAINED, device=None):
        """The shipped network (or another saved by save()); the file is read with
        numpy's loader without pickles.  Returns (net, record): record holds the
        training figures stored beside the weights."""
        with np.load(path, allow_pickle=False) as z:
            params = np.asarray(z["params"], np.float32)
            norm = np.asarray(z["norm"], np.float32)
            record = {k: z[k] for k in z.files if k not in ("params", "norm")}
        if params.shape != (P,) or norm.shape != (2 * F + 4,):
            raise ValueError(f"{path}: parameter vector {params.shape} / normalisation {norm.shape} do not match "
                             f"the network layout ({P},) / ({2 * F + 4},)")
        net = cls(device=device)
        net.load_state_dict({"params": params, "norm": norm})
        return net, record

    def save(self, path, **record):
        np.savez(path, **self.state_dict(), **record)

    def state_dict(self):
        return {"params": self.params.cpu().numpy(), "norm": self.norm.cpu().numpy()}

    def load_state_dict(self, d):
        self.params = torch.tensor(np.asarray(d["params"], np.float32), device=self.device)
        self.norm = torch.tensor(np.asarray(d["norm"], np.float32), device=self.device)
        nat.check(nat.lib().bn_mlp_to_bf16(nat.ptr(self.params), nat.ptr(self.wbf), nat.stream_ptr()))


def run_to_convergence(model, tol=1e-7, max_iter=200, lr=1.0, log=None, record=None):
    """inference(lr) + energy() until |E_k - E_{k-1}| <= tol |E_k| for every run.
    A run whose energy is no longer finite has diverged (the plain iteration from the
    N(0, 1e2) sites can overshoot past the range of exp on large counts): it leaves the
    test and keeps iterating with the others; callers read it from the non-finite
    energy.  Returns (iterations, final energies [K], wall seconds); the GPU is
    synchronised before the clock is read.  `log(it, energies)` is called after every
    iteration.  With a dict `record`, record["first_met"] holds per run the first
    iteration at which the test held (0: never) and record["diverged_at"] the first
    iteration with a non-finite energy (0: never)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prev = None
    it = 0
    first = torch.zeros(model.K, dtype=torch.int64, device=model.device)
    div = torch.zeros_like(first)
    while it < max_iter:
        model.inference(lr=lr, return_terms=False)
        e = model.energy_per_run()
        it += 1
        if log is not None:
            log(it, e)
        bad = ~torch.isfinite(e)
        div = torch.where((div == 0) & bad, torch.full_like(div, it), div)
        if prev is not None:
            met = (torch.abs(e - prev) <= tol * torch.abs(e)) & ~bad
            first = torch.where((first == 0) & met, torch.full_like(first, it), first)
            if bool(torch.all(met | bad)):
                break
        prev = e
    if record is not None:
        record["first_met"] = first.cpu().tolist()
        record["diverged_at"] = div.cpu().tolist()
    torch.cuda.synchronize()
    return it, e, time.perf_counter() - t0
