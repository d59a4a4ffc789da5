"""Benchmark: batched filter + smoother time-steps/s on one MI355X.

Workload (one "step" = one fused filter + RTS smoother pass over the batch,
i.e. MarkovGaussianProcess.update_posterior, bayesnewton/basemodels.py:689-706,
for every series): B = 4096 independent Markov-GP series of lengths log-uniform
in [256, 16384] (sum 15.74M time steps), Matern52 + Matern32 sum kernel (state
dim 5) with one lengthscale per series, synthetic GP draws + 0.2 t_3 noise as
pseudo observations with pseudo variance 0.04 (the Student-t scale^2).  Inputs
are resident in HBM before the timed region.

Prints one JSON line.  `value` = total time steps / second of one pass, from the
median of K hipGraph replays of the pass timed with hipEvents on the replay stream.
`roofline` is the pass-level HBM bound: 180 algorithmic bytes per time step over
the pass time (the pass's launches are recorded in one graph; per-kernel launch
times, timed live by libbn_hip.so on the launch stream, are listed beside it).
`cpu_baseline` is the oracle's FP64 C filter + smoother (oracle/kfrts.c built with
-march=native on the measuring host, OpenMP over series) on the same inputs.
`warm_start` is the second half of the metric: plain vs learned warm start to the
same convergence test on held-out space-time runs, with the shipped network.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
D = 5
# Algorithmic bytes per time step (FP32 words, d = 5): dt, y, r in, smoothed mean
# and variance out, filtered mean + packed covariance written once and read
# once: 4*(3+2) + 2*4*(d + d(d+1)/2) = 180 B per filter+smoother step (the
# sequential filter-then-smoother's own I/O).  Per kernel of this pass: the chunk
# reduction reads dt, y, r (12 B); the filter pass reads them and writes the
# smoother's FP32 rows H P- (d) and H m- (12 + 24 = 36 B); the Bryson-Frazier
# smoother reads inputs and rows and writes mean / variance (12 + 24 + 8 = 44 B);
# the two-pass RTS form (k_smoother_pass) reads the filtered moments (80 + 8 = 88 B).
PASS_BYTES_PER_STEP = 180.0
ALG_BYTES = {"k_filter_reduce": 12.0, "k_filter_pass": 36.0, "k_smoother_pass": 88.0, "k_smoother_mbf": 44.0}
# Algorithmic FP64 flops per whole filter+smoother step: the dense count of the
# reference recurrences (ops.py:159-171, 294-305), 14 1/3 d^3 + lower order terms =
# 2299 at d = 5, p = 1 (the kernels exploit the block-diagonal A of the sum kernel and
# execute fewer).
PASS_FLOPS_PER_STEP = 2299.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chunk", default="auto",
                    help="batched workload: chunk length, or 'auto' (bayesnewton_amd.autotune, before the timed passes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--retune", action="store_true",
                    help="batched workload: time the schedules even when the shipped table has this batch")
    ap.add_argument("--record-schedule", default=None,
                    help="batched workload: write the tuned schedule into this table (JSON)")
    ap.add_argument("--ab-runs", type=int, default=2,
                    help="default workload: held-out space-time runs (seeds --st-heldout, ...) of the warm-start "
                         "A/B reported beside the pass figures (0 = skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (reference dispatch for HBM counter scaling)")
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "spacetime", "spacetime-pass", "meanfield-seq", "meanfield-par",
                             "infinite-horizon", "gradient", "predict"],
                    help="batched: filter+smoother steps/s over 4096 uneven series (default); spacetime: "
                         "time to converge of the d=128 Cox process, learned warm start vs plain; spacetime-pass: "
                         "VI iterations (sites + filter + smoother + energy) over --st-runs d=128 runs; predict: "
                         "predictive mean / variance at one test time per training time of the 4096-series batch; gradient: "
                         "200 Adam steps on the energy of one 500k-step series")
    ap.add_argument("--st-train", type=int, default=200, help="spacetime: number of training runs")
    ap.add_argument("--st-runs", type=int, default=200, help="spacetime-pass: runs in the batch")
    ap.add_argument("--st-heldout", type=int, default=1000, help="spacetime: first held-out seed")
    ap.add_argument("--st-heldout-n", type=int, default=10, help="spacetime: number of held-out runs (one batch)")
    ap.add_argument("--st-max-iter", type=int, default=600, help="spacetime: iteration cap of every run")
    ap.add_argument("--st-save", default=None, help="spacetime: write the trained network and training record (.npz)")
    ap.add_argument("--st-train-only", action="store_true", help="spacetime: stop after training (with --st-save)")
    ap.add_argument("--st-load", default=None,
                    help="spacetime: skip training, load the network and training record written by --st-save")
    ap.add_argument("--st-steps", type=int, default=2000, help="spacetime: time steps per run")
    ap.add_argument("--st-tol", type=float, default=1e-7, help="spacetime: relative energy change to stop")
    ap.add_argument("--st-lr", type=float, default=0.2,
                    help="spacetime: inference step size (lr_newton of the reference's 2-D Cox demo)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_gradient(args):
    """One series of N = 500,000 sorted U[0, 5000] inputs, Matern52 (l=1, s2=1), y = GP draw
    + 0.1 eps, Gaussian(0.01), MarkovVariationalGP with the associative scan.  Times 200
    Adam steps (lr 0.1) of the reference's training loop (README.md:56-76): inference(lr=1),
    energy and its gradient (MarkovEnergy), optimiser step.  The first gradient is checked
    against central differences (h = 1e-4 on the raw hypers) of -log p(y) from the FP64 C
    oracle, which equals the VI energy at the exact Gaussian sites."""
    import numpy as np
    import torch
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M, workloads

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t0 = time.time()
    d = workloads.single_series_500k()
    log(f"[gradient] data generated in {time.time() - t0:.1f}s")
    x, y = d["x"], d["y"].astype(np.float32).astype(np.float64)
    model = M.MarkovVariationalGP(K.Matern52(variance=1.0, lengthscale=1.0), L.Gaussian(variance=0.01), x, y,
                                  parallel=True)
    model.inference(lr=1.0)
    E0, g0 = model.energy_and_grad()
    names = model.trainable_names()
    raw0 = np.array(model.get_raw_hypers())
    fd = None
    cpu_s = None
    if not args.no_cpu:
        from oracle import cfast
        dt = np.concatenate([[0.0], np.diff(x)]).astype(np.float32).astype(np.float64)

        def negll(raw):
            sp = np.log1p(np.exp(raw))
            c0 = time.perf_counter()
            _, _, ll = cfast.filter_smoother(["matern52"], [sp[0]], [sp[1]], [0, len(x)], dt, y,
                                             np.full(len(x), sp[2]))
            return -ll[0], time.perf_counter() - c0

        h = 1e-4
        fd = []
        for i in range(len(raw0)):
            e = np.zeros_like(raw0)
            e[i] = h
            fp, cpu_s = negll(raw0 + e)
            fm, _ = negll(raw0 - e)
            fd.append((fp - fm) / (2 * h))
        fd = np.array(fd)
    raw = torch.tensor(raw0, dtype=torch.float64, device=dev, requires_grad=True)
    opt = torch.optim.Adam([raw], lr=0.1)

    def step():
        model.inference(lr=1.0, return_terms=False)
        E = M.MarkovEnergy.apply(raw, model)
        opt.zero_grad()
        E.backward()
        opt.step()
        return E

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    n_steps = 200
    t0 = time.perf_counter()
    for _ in range(n_steps):
        E = step()
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    out = {
        "metric": "energy-gradient training steps/s (one 500k-step series, Matern52, 200 Adam steps)",
        "value": n_steps / sec,
        "unit": "Adam steps/s",
        "n_gpus": 1,
        "steps": n_steps,
        "ms_per_step": sec * 1e3 / n_steps,
        "higher_is_better": True,
        "dtype": "f64 recurrences and tangents, f32 per-step I/O",
        "data": "synthetic (seeded GP draw + 0.1 noise)",
        "config": {"workload": "N=500,000 sorted U[0,5000], Matern52 (l=1, s2=1), Gaussian(0.01), "
                               "inference(lr=1) + energy gradient + Adam(lr=0.1) per step"},
        "time_steps_per_s": len(x) * n_steps / sec,
        "first_gradient": {"names": names, "gpu": [float(v) for v in g0],
                           "fd_fp64_oracle": None if fd is None else [float(v) for v in fd],
                           "max_rel_err": None if fd is None else float(np.max(np.abs(g0 - fd) / np.abs(fd)))},
        "energy_first": float(E0), "energy_last": float(E.detach()),
        "hypers_last": [float(v) for v in np.log1p(np.exp(raw.detach().cpu().numpy()))],
        "cpu_baseline": None if cpu_s is None else {
            "value": 1.0 / cpu_s, "unit": "log-likelihood evaluations/s (C FP64 filter+smoother, 1 core)",
            "cores": 1, "kind": "port", "sample": "one sequential filter+smoother over the 500k steps"},
    }
    print(json.dumps(out), flush=True)


def bench_spacetime(args):
    """Space-time log-Gaussian Cox process (d = 128: M = 64 inducing points on an 8x8 grid,
    Matern32 in time (lengthscale 10) x Matern32 in space (lengthscale 0.2), T steps of
    R = 256 locations, Poisson(binsize 2)), MarkovVariationalGP.  Trains the warm-start
    network on the converged sites of --st-train plain runs (seeds 2000, 2001, ...) and
    compares plain vs warm start on --st-heldout-n held-out runs (seeds --st-heldout, ...),
    each side as one batch: iterations and wall time to the same convergence test
    (relative energy change <= --st-tol), final energies, diverged runs."""
    import numpy as np
    import torch
    from bayesnewton_amd import _native as nat, kernels as K, likelihoods as L, spacetime as S
    from bayesnewton_amd import warmstart as W, workloads

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    T = args.st_steps
    g = np.linspace(0.0, 1.0, 8)
    za, zb = np.meshgrid(g, g)
    z = np.stack([za.reshape(-1), zb.reshape(-1)], 1)

    def kern():
        return K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=10.0),
                                      K.Matern32(variance=1.0, lengthscale=0.2), z=z, sparse=True, conditional="Full")

    lik = L.Poisson(binsize=2.0)
    if args.st_load:
        net, rec = W.WarmStartMLP.pretrained(args.st_load)
        it_tr, sec_tr, fit_s, n_div = int(rec["it_tr"]), float(rec["sec_tr"]), float(rec["fit_s"]), int(rec["n_div"])
        hist = [float(v) for v in rec["hist"]]
        kern_ms = dict(zip([str(k) for k in rec["kern_names"]], [float(v) for v in rec["kern_ms"]]))
        steps_per_pass = int(rec["steps_per_pass"])
        log(f"[spacetime] network and training record loaded from {args.st_load}")
    else:
        t0 = time.perf_counter()
        train = []
        for k in range(args.st_train):
            train.append(workloads.spacetime_cox(seed=2000 + k, T=T, device=dev))
            if (k + 1) % 20 == 0:
                log(f"[spacetime] generated {k + 1}/{args.st_train} training runs ({time.perf_counter() - t0:.1f}s)")
        log(f"[spacetime] {args.st_train} training runs generated in {time.perf_counter() - t0:.1f}s")
        tm = S.SpaceTimeVariationalGP.from_runs(kern(), lik, [d["t"] for d in train], [d["r"] for d in train],
                                                [d["y"] for d in train])
        del train
        L_ = nat.lib()
        L_.bn_timing_enable(1)
        state = {"prev": None}

        def tr_log(i, e):
            ok = torch.isfinite(e)
            msg = f"[spacetime] training batch iteration {i}: {int(ok.sum())}/{e.numel()} runs finite"
            if state["prev"] is not None:
                rel = torch.abs(e - state["prev"]) / torch.abs(e)
                rel = rel[ok & torch.isfinite(rel)]
                if rel.numel():
                    msg += f", {int((rel <= args.st_tol).sum())} within tol, max rel change {float(rel.max()):.3e}"
            state["prev"] = e
            log(msg)

        it_tr, e_tr, sec_tr = W.run_to_convergence(tm, tol=args.st_tol, max_iter=args.st_max_iter, lr=args.st_lr,
                                                   log=tr_log)
        conv = torch.isfinite(e_tr)
        n_div = int((~conv).sum())
        log(f"[spacetime] training batch: {n_div} of {tm.K} plain runs diverged (non-finite energy)")
        kern_ms = {}
        for kid in range(L_.bn_num_kernel_ids()):
            tot, cnt = ctypes.c_double(), ctypes.c_int32()
            nat.check(L_.bn_timing_read(kid, ctypes.byref(tot), ctypes.byref(cnt)))
            if cnt.value:
                kern_ms[L_.bn_kernel_name(kid).decode()] = tot.value / cnt.value
        L_.bn_timing_enable(0)
        log(f"[spacetime] plain training batch: K={tm.K} converged in {it_tr} iterations, {sec_tr:.1f}s; "
            f"per-launch ms {kern_ms}")
        steps_per_pass = tm.K * tm.T
        X = W.site_features(tm)
        Y = W.site_marginals(tm)
        keep = conv.repeat_interleave(tm.T * tm.n_loc) & torch.isfinite(Y).all(1)  # converged runs only
        X, Y = X[keep], Y[keep]
        del tm
        torch.cuda.empty_cache()
        n_fit = min(X.shape[0], 4_000_000)
        gen = torch.Generator(device=dev)
        gen.manual_seed(0)
        idx = torch.randperm(X.shape[0], generator=gen, device=dev)[:n_fit]
        Xs, Ys = X.index_select(0, idx).contiguous(), Y.index_select(0, idx).contiguous()
        del X, Y
        net = W.WarmStartMLP(seed=0)
        t0 = time.perf_counter()
        hist = net.fit(Xs, Ys, epochs=3, batch=8192, lr=2e-3, log=log)
        torch.cuda.synchronize()
        fit_s = time.perf_counter() - t0
        del Xs, Ys
        torch.cuda.empty_cache()
        if args.st_save:
            sd = net.state_dict()
            np.savez(args.st_save, params=sd["params"], norm=sd["norm"], it_tr=it_tr, sec_tr=sec_tr, fit_s=fit_s, n_div=n_div,
                     hist=np.asarray(hist), kern_names=np.asarray(list(kern_ms.keys())),
                     kern_ms=np.asarray(list(kern_ms.values())), steps_per_pass=steps_per_pass)
            log(f"[spacetime] network and training record written to {args.st_save}")
        if args.st_train_only:
            print(json.dumps({"spacetime_training": {"runs": args.st_train, "diverged": n_div, "iterations": it_tr,
                                                     "seconds": sec_tr,
                                                     "fit_seconds": fit_s, "loss": hist,
                                                     "kernel_ms_per_launch": kern_ms}}), flush=True)
            return
    pass_ms = kern_ms.get("k_st_filter", 0.0) + kern_ms.get("k_st_smoother", 0.0)
    seeds = list(range(args.st_heldout, args.st_heldout + args.st_heldout_n))
    out = warmstart_ab(net, seeds, T, args.st_tol, args.st_max_iter, args.st_lr, dev)
    out["config"]["train_runs"] = args.st_train
    out["config"]["network_loaded_from"] = args.st_load
    out["train"] = {"plain_iterations": it_tr, "plain_seconds": sec_tr, "plain_diverged": n_div,
                    "fit_seconds": fit_s, "loss": hist}
    out["batched_pass"] = {"runs": steps_per_pass // T, "filter_plus_smoother_ms": pass_ms,
                           "steps_per_s": steps_per_pass / (pass_ms / 1e3) if pass_ms else None,
                           "kernel_ms_per_launch": kern_ms}
    print(json.dumps(out), flush=True)


def warmstart_ab(net, seeds, T, tol, max_iter, lr, dev):
    """Plain vs learned warm start on held-out runs of the space-time Cox process (d = 128:
    M = 64 inducing points on an 8x8 grid, Matern32(l=10) x Matern32(l=0.2), T steps of
    R = 256 locations, Poisson(binsize 2)), each start as one batch of the held-out seeds:
    iterations and wall time until every run meets the same convergence test (relative
    energy change <= tol), final energies, diverged runs.  Both wall times include
    everything after the data are on the GPU (the warm side includes the network)."""
    import numpy as np
    import torch
    from bayesnewton_amd import kernels as K, likelihoods as L, spacetime as S, warmstart as W, workloads

    g = np.linspace(0.0, 1.0, 8)
    za, zb = np.meshgrid(g, g)
    z = np.stack([za.reshape(-1), zb.reshape(-1)], 1)

    def kern():
        return K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=10.0),
                                      K.Matern32(variance=1.0, lengthscale=0.2), z=z, sparse=True, conditional="Full")

    lik = L.Poisson(binsize=2.0)
    held = [workloads.spacetime_cox(seed=sd, T=T, device=dev) for sd in seeds]

    def batch():
        return S.SpaceTimeVariationalGP.from_runs(kern(), lik, [d["t"] for d in held], [d["r"] for d in held],
                                                  [d["y"] for d in held])

    def ab_log(tag):
        def f(i, e):
            if i % 10 == 0 or i < 5:
                ok = torch.isfinite(e)
                log(f"[warm start A/B] {tag} iteration {i}: {int(ok.sum())}/{e.numel()} finite, "
                    f"energies {[round(float(v), 4) for v in e[:4]]}")
        return f

    plain = batch()
    rec_p = {}
    it_p, e_p, sec_p = W.run_to_convergence(plain, tol=tol, max_iter=max_iter, lr=lr, log=ab_log("plain"),
                                            record=rec_p)
    log(f"[warm start A/B] plain: {it_p} iterations, {sec_p:.2f}s")
    del plain
    warm = batch()
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    net.init_sites(warm)
    torch.cuda.synchronize()
    init_s = time.perf_counter() - c0
    rec_w = {}
    it_w, e_w, sec_w = W.run_to_convergence(warm, tol=tol, max_iter=max_iter, lr=lr, log=ab_log("warm"),
                                            record=rec_w)
    sec_w += init_s
    log(f"[warm start A/B] warm: {it_w} iterations, {sec_w:.2f}s (network {init_s:.3f}s)")
    ep, ew = e_p.cpu().numpy(), e_w.cpu().numpy()
    runs = []
    for i, sd in enumerate(seeds):
        runs.append({"seed": sd, "plain_iterations": rec_p["first_met"][i] or None,
                     "plain_diverged_at": rec_p["diverged_at"][i] or None,
                     "warm_iterations": rec_w["first_met"][i] or None,
                     "warm_diverged_at": rec_w["diverged_at"][i] or None,
                     "plain_energy": float(ep[i]), "warm_energy": float(ew[i]),
                     "energy_rel_diff": float(abs(ew[i] - ep[i]) / abs(ep[i])) if np.isfinite(ep[i]) else None})
    both = [r for r in runs if r["plain_iterations"] and r["warm_iterations"]]
    per_it_p, per_it_w = sec_p / it_p, (sec_w - init_s) / it_w
    return {
        "metric": "time-to-converge, learned warm start vs plain (space-time Cox process, d=128)",
        "value": sec_w / sec_p,
        "unit": "warm-start wall time / plain wall time (held-out batch, same GPU, same test)",
        "higher_is_better": False,
        "dtype": "f64 filter/smoother/sites, bf16 weights + f32 accumulation in the network",
        "data": "synthetic (seeded prior draws of the model, Poisson counts)",
        "config": {"workload": f"space-time Cox process: T={T}, R=256 locations/step, M=64 inducing (8x8), "
                               "Matern32(l=10) x Matern32(l=0.2), Poisson(binsize=2), MarkovVariationalGP",
                   "held_out_seeds": list(seeds), "tol": tol, "lr": lr, "max_iter": max_iter},
        "plain": {"iterations": it_p, "seconds": sec_p, "seconds_per_iteration": per_it_p,
                  "diverged": sum(1 for r in runs if r["plain_diverged_at"])},
        "warm": {"iterations": it_w, "seconds": sec_w, "network_seconds": init_s, "seconds_per_iteration": per_it_w,
                 "diverged": sum(1 for r in runs if r["warm_diverged_at"])},
        "converged_in_both": len(both),
        "energy_rel_diff_max": max((r["energy_rel_diff"] for r in both), default=None),
        "runs": runs,
    }


def st_flops_per_step(D, M, R):
    """Algorithmic FP64 flops per time step of the d = 128 kernels (mul + add = 2):
    smoother: sweep inverse of P- (D pivots x D^2 updates) + three D^3 products;
    filter: U = P-[:, :M] L, M x M products L'U and the sweep of Mm, V = U Mm^-1, P -= V U';
    sites (one launch): C P (R x M x M) and B' diag(n2) B (M x M x R)."""
    return {"k_st_smoother": 2.0 * D ** 3 + 3 * 2.0 * D ** 3,
            "k_st_filter": 2.0 * (2 * D * M * M + 2 * M ** 3 + D * D * M),
            "k_st_sites": 2.0 * (2 * R * M * M)}


def mf_flops_per_step(M):
    """Algorithmic FP64 flops per time step of the mean-field kernels: the sweep inverse of
    an M x M matrix (M pivots x M^2 updates) and the M x M product S^-1 v; k_mf_pseudo also
    forms L L' (M^3 / 3 multiply-adds).  The per-block filters and smoother are O(M KT^3)."""
    sweep = 2.0 * M ** 3 + 2.0 * M * M
    return {"k_mf_filter": sweep, "k_mf_sdiag": sweep, "k_mf_pseudo": sweep + 2.0 * M ** 3 / 3}


# kernels of one filter + smoother pass and their launches per pass
MF_PASS_KERNELS = {"seq": {"k_mf_pseudo": 1, "k_mf_filter": 1, "k_mf_smoother": 1},
                   "par": {"k_mf_pseudo": 1, "k_mf_sdiag": 2, "k_mf_blocks": 1, "k_mf_smoother": 1},
                   "ih": {"k_mf_pseudo": 1, "k_ih_tied": 1, "k_ih_dare": 1, "k_ih_sinv": 1, "k_ih_filter": 1,
                          "k_ih_smoother": 1}}


def bench_spacetime_pass(args, mean_field=None):
    """Throughput of the d = 128 space-time path: --st-runs runs of the Cox-process shape
    (T = --st-steps, R = 256, 8x8 inducing grid, Matern32 x Matern32), one VI iteration per
    step (site update, projection, filter, smoother, energy), lr = --st-lr.  `value` is
    filter + smoother time-steps per second over the batch; `roofline` is the dominant
    kernel's algorithmic FP64 rate (st_flops_per_step) against MI355X's FP64 peak.
    mean_field "seq" / "par": the same model with the mean-field filter and smoother
    (MarkovVariationalMeanFieldGP; sequential or associative form); "ih": the
    infinite-horizon model (InfiniteHorizonVariationalGP, 20 Riccati iterations per call).
    Their filter + smoother time includes forming the pseudo noise (k_mf_pseudo) and, for
    "ih", the tied noise and the Riccati iterations."""
    import numpy as np
    import torch
    from bayesnewton_amd import _native as nat, kernels as K, likelihoods as L, spacetime as S, workloads

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    T = args.st_steps
    g = np.linspace(0.0, 1.0, 8)
    za, zb = np.meshgrid(g, g)
    z = np.stack([za.reshape(-1), zb.reshape(-1)], 1)

    def kern():
        return K.SpatioTemporalKernel(K.Matern32(variance=1.0, lengthscale=10.0),
                                      K.Matern32(variance=1.0, lengthscale=0.2), z=z, sparse=True, conditional="Full")

    runs = [workloads.spacetime_cox(seed=2000 + k, T=T, device=dev) for k in range(args.st_runs)]
    if mean_field is None:
        m = S.SpaceTimeVariationalGP.from_runs(kern(), L.Poisson(binsize=2.0), [d["t"] for d in runs],
                                               [d["r"] for d in runs], [d["y"] for d in runs])
    elif mean_field == "ih":
        m = S.SpaceTimeVariationalInfiniteHorizonGP.from_runs(kern(), L.Poisson(binsize=2.0), [d["t"] for d in runs],
                                                              [d["r"] for d in runs], [d["y"] for d in runs])
    else:
        m = S.SpaceTimeVariationalMeanFieldGP.from_runs(kern(), L.Poisson(binsize=2.0), [d["t"] for d in runs],
                                                        [d["r"] for d in runs], [d["y"] for d in runs],
                                                        parallel=(mean_field == "par"))
    del runs
    log(f"[spacetime-pass] {m.K} runs x {T} steps ready (mean field: {mean_field})")

    def step():
        m.inference(lr=args.st_lr, return_terms=False)
        m.energy_per_run()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    L_ = nat.lib()
    L_.bn_timing_enable(1)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"[spacetime-pass] iteration {i + 1}/{args.steps}")
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    kern_ms = {}
    for kid in range(L_.bn_num_kernel_ids()):
        tot, cnt = ctypes.c_double(), ctypes.c_int32()
        nat.check(L_.bn_timing_read(kid, ctypes.byref(tot), ctypes.byref(cnt)))
        if cnt.value:
            kern_ms[L_.bn_kernel_name(kid).decode()] = (tot.value / cnt.value, cnt.value)
    L_.bn_timing_enable(0)
    steps = m.K * m.T
    if mean_field is None:
        pass_kernels = {"k_st_filter": 1, "k_st_smoother": 1}
        flops = st_flops_per_step(m.D, m.M, m.n_loc)
    else:
        pass_kernels = MF_PASS_KERNELS[mean_field]
        flops = {"k_st_sites": st_flops_per_step(m.D, m.M, m.n_loc)["k_st_sites"], **mf_flops_per_step(m.M)}
        if mean_field == "ih":
            flops = {"k_st_sites": flops["k_st_sites"], "k_ih_sinv": flops["k_mf_sdiag"],
                     "k_mf_pseudo": flops["k_mf_pseudo"]}
    fs_ms = sum(kern_ms[k][0] * n for k, n in pass_kernels.items() if k in kern_ms)
    cands = [k for k in flops if k in kern_ms]
    dom = max(cands, key=lambda k: kern_ms[k][0] * kern_ms[k][1])
    dom_ms = kern_ms[dom][0]
    achieved = flops[dom] * steps / (dom_ms / 1e3) / 1e12
    cpu = None
    if not args.no_cpu:
        # the oracle (FP64 NumPy) on a bounded sample: one run, the first 40 steps
        from oracle import kernels as OK, likelihoods as OL, models as OM
        d = workloads.spacetime_cox(seed=2000, T=40)
        ok = OK.SpatioTemporalKernel(OK.Matern32(1.0, 10.0), OK.Matern32(1.0, 0.2), z, sparse=True,
                                     conditional="Full")
        om = OM.MarkovGP(ok, OL.Poisson(2.0), d["t"][:, None], d["y"], R=d["r"], inference="VI",
                         mean_field=mean_field in ("seq", "par"), parallel=(mean_field == "par"),
                         infinite_horizon=(mean_field == "ih"))
        c0 = time.perf_counter()
        n_it = 3
        for _ in range(n_it):
            om.inference(lr=args.st_lr)
            om.energy()
        csec = (time.perf_counter() - c0) / n_it
        cpu = {"value": 40 / csec, "unit": "time-steps/s (whole VI iteration incl. energy)",
               "cores": cpu_threads(), "kind": "port",
               "sample": f"oracle/models.py MarkovGP (FP64 NumPy, BLAS threads = cores), one run of 40 steps, "
                         f"{n_it} VI iterations + energy, {csec:.2f}s per iteration"}
    out = {
        "metric": ("space-time VI iteration: time-steps/s over a batch of d=128 runs (filter + smoother kernels)"
                   + {None: "", "seq": ", mean-field sequential form", "par": ", mean-field associative form",
                      "ih": ", infinite-horizon"}[mean_field]),
        "value": steps / (fs_ms / 1e3),
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded prior draws of the model, Poisson counts)",
        "config": {"workload": f"space-time Cox process: {m.K} runs, T={T}, R={m.n_loc}, M={m.M} (8x8), d={m.D}, "
                               "Matern32(l=10) x Matern32(l=0.2), Poisson(binsize=2), one VI iteration per step"
                               + {None: "", "seq": ", mean-field filter/smoother (sequential)",
                                  "par": ", mean-field filter/smoother (associative)",
                                  "ih": ", infinite-horizon filter/smoother (20 Riccati iterations per call)"}[
                                   mean_field],
                   "runs": m.K, "time_steps": T, "state_dim": m.D},
        "iteration_time_steps_per_s": steps / sec,
        "roofline": {"bound": "mfma", "compute": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFS, "traffic": None, "kernel": dom,
                     "alg_flops_per_step": flops[dom], "steps_per_launch": steps, "avg_launch_ms": dom_ms},
        "kernel_ms_per_launch": {k: v[0] for k, v in kern_ms.items()},
        "kernel_tflops": {k: flops[k] * steps / (kern_ms[k][0] / 1e3) / 1e12 for k in flops if k in kern_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


# Algorithmic bytes per predicted test time (d = 5): the test time and both outputs
# (8 + 16), the two neighbouring smoothed states (2 x (5 + 25) doubles) and the left
# gain (25 doubles): 24 + 480 + 200 = 704 B.
PRED_BYTES = 704.0


def bench_predict(args):
    """Prediction over the 4096-series batch (MarkovGaussianProcess.predict, basemodels.py:766-816):
    4096 series fitted with one VI step, one test time per training time (uniform over
    each series' span, 0.5 beyond both ends).  `value` is test times per second through
    bn_predict with the test times resident on the GPU and the smoothed full state
    computed beforehand (that pass, once per site update, is timed beside it)."""
    import numpy as np
    import torch
    from bayesnewton_amd import _native as nat, kernels as K, likelihoods as L, models as M, workloads

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    data = workloads.batched_4096_series()
    off = data["offsets"]
    B = len(off) - 1
    xs = [data["x"][off[b]:off[b + 1]] for b in range(B)]
    ys = [data["y"][off[b]:off[b + 1]].astype(np.float32).astype(np.float64) for b in range(B)]
    kern = K.Sum([K.Matern52(), K.Matern32()])
    m = M.MarkovVariationalGP.from_series(kern, L.StudentsT(scale=0.2, df=3.0), xs, ys, lengthscale=data["ls"],
                                          variance=data["var"], chunk_len=args.chunk)
    m.inference(lr=1.0, return_terms=False)
    rng = np.random.default_rng(7)
    tests = [rng.uniform(x[0] - 0.5, x[-1] + 0.5, len(x)) for x in xs]
    xt, toff = m._test_points(tests)
    n = int(xt.shape[0])
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    sm, sP, G = m._full_posterior()
    torch.cuda.synchronize()
    full_s = time.perf_counter() - c0
    mean = torch.empty(n, dtype=torch.float64, device=dev)
    var = torch.empty_like(mean)
    ls, lv = m._hypers()

    def launch():
        nat.check(nat.lib().bn_predict(m.plan.handle, ctypes.byref(m._desc), nat.ptr(ls), nat.ptr(lv),
                                       m._hyper_stride, nat.ptr(m._train_off), nat.ptr(m._x_train), nat.ptr(sm),
                                       nat.ptr(sP), nat.ptr(G), nat.ptr(toff), nat.ptr(xt), n, nat.ptr(mean),
                                       nat.ptr(var), nat.stream_ptr()))

    for _ in range(args.warmup):
        launch()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        launch()
    e1.record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    kern_ms = e0.elapsed_time(e1) / args.steps
    gbs = PRED_BYTES * n / (kern_ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import kernels as OK, likelihoods as OL, models as OM
        b = int(np.argmin(np.abs(np.diff(off) - 2000)))
        o = OM.MarkovGP(OK.Sum([OK.Matern52(1.0, data["ls"][b, 0]), OK.Matern32(1.0, data["ls"][b, 1])]),
                        OL.StudentsT(0.2, 3.0), xs[b], ys[b], inference="VI")
        o.inference(lr=1.0)
        c0 = time.perf_counter()
        o.predict(tests[b][:, None])
        csec = time.perf_counter() - c0
        cpu = {"value": len(tests[b]) / csec, "unit": "test times/s", "cores": 1, "kind": "port",
               "sample": f"oracle/models.py predict (FP64 NumPy: filter, smoother, per-point conditional) on series "
                         f"{b} ({len(xs[b])} steps, {len(tests[b])} test times), {csec:.2f}s"}
    out = {
        "metric": "prediction: test times/s (4096-series batch, predictive mean and variance of f)",
        "value": n / (kern_ms / 1e3),
        "unit": "test times/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kern_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (the 4096-series batch, one VI step of Student-t sites)",
        "config": {"workload": f"predict at {n} test times over {B} series (one per training time), d=5",
                   "series": B, "test_times": n},
        "wall_ms_per_call": wall * 1e3,
        "full_state_filter_smoother_s": full_s,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "scope": "one bn_predict call: k_predict_locate, radix sort by state slot, "
                     "k_predict", "alg_bytes_per_test_time": PRED_BYTES, "test_times_per_call": n,
                     "ms_per_call": kern_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def copy_calibration():
    """A plain 1 GiB device copy: FETCH_SIZE / WRITE_SIZE of its dispatch against the
    known 1 GiB read + 1 GiB written give the factor that turns those counters into bytes."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}), flush=True)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        n = min(n, int(omp))
    return max(1, n)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: bayesnewton_amd runs on one GPU; --gpus N with N > 1 is not supported", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "spacetime":
        bench_spacetime(args)
        return
    if args.workload == "spacetime-pass":
        bench_spacetime_pass(args)
        return
    if args.workload in ("meanfield-seq", "meanfield-par"):
        bench_spacetime_pass(args, mean_field=args.workload.split("-")[1])
        return
    if args.workload == "infinite-horizon":
        bench_spacetime_pass(args, mean_field="ih")
        return
    if args.workload == "gradient":
        bench_gradient(args)
        return
    if args.workload == "predict":
        bench_predict(args)
        return
    import numpy as np
    import torch
    from bayesnewton_amd import _native as nat, kernels as K, ops, workloads
    from bayesnewton_amd.plan import SeriesPlan

    t_gen = time.time()
    data = workloads.batched_4096_series()
    off = data["offsets"]
    N = int(off[-1])
    B = len(off) - 1
    dt = data["dt"].astype(np.float32)
    y = data["y"].astype(np.float32)
    r = np.full(N, data["scale"] ** 2, dtype=np.float32)
    print(f"[bench] batch generated in {time.time() - t_gen:.1f}s: B={B} N={N}", file=sys.stderr, flush=True)

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    kern = K.Sum([K.Matern52(), K.Matern32()])
    desc = kern.desc()
    ls = torch.tensor(data["ls"], dtype=torch.float64, device=dev)
    var = torch.tensor(data["var"], dtype=torch.float64, device=dev)
    tune = None
    if str(args.chunk) == "auto":
        from bayesnewton_amd import autotune
        t_tune = time.time()
        tune = autotune.tune_schedule(off, desc, ls, var, torch.from_numpy(dt).to(dev), torch.from_numpy(y).to(dev),
                                      torch.from_numpy(r).to(dev), hyper_stride=2, device=dev,
                                      table=not args.retune)
        tune["seconds"] = time.time() - t_tune
        chunk = int(tune["chunk_len"])
        print(f"[bench] schedule {tune['schedule']} chunk_len {chunk} ({tune['source']}): "
              f"{tune.get('ms_per_pass')}", file=sys.stderr, flush=True)
        if args.record_schedule:
            rec = {k: v for k, v in tune.items() if k not in ("source", "signature", "seconds")}
            autotune.record_schedule(tune["signature"], rec, path=args.record_schedule)
    else:
        chunk = int(args.chunk)
    plan = SeriesPlan(off, chunk_len=chunk, device=dev)
    dt_t = plan.gather(torch.from_numpy(dt).to(dev))
    y_t = plan.gather(torch.from_numpy(y).to(dev))
    r_t = plan.gather(torch.from_numpy(r).to(dev), pad=1.0)
    pm = torch.empty(plan.n_slots, dtype=torch.float32, device=dev)
    pv = torch.empty(plan.n_slots, dtype=torch.float32, device=dev)
    ll = torch.empty(B, dtype=torch.float64, device=dev)
    plan.workspace(desc)

    def step():
        ops.filter_smoother(plan, desc, ls, var, dt_t, y_t, r_t, None, hyper_stride=2, post_mean=pm,
                            post_var=pv, loglik=ll)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # per-kernel launch times (hipEvents recorded by the library on the launch stream),
    # from eager passes outside the timed region
    L = nat.lib()
    L.bn_timing_enable(1)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    kern_ms = {}
    for kid in range(L.bn_num_kernel_ids()):
        tot = ctypes.c_double()
        cnt = ctypes.c_int32()
        nat.check(L.bn_timing_read(kid, ctypes.byref(tot), ctypes.byref(cnt)))
        if cnt.value:
            kern_ms[L.bn_kernel_name(kid).decode()] = tot.value / cnt.value
    L.bn_timing_enable(0)
    # eager passes, timed back to back between two events (launch overhead included)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    e1.synchronize()
    eager_ms = e0.elapsed_time(e1) / args.steps
    # the timed region: one pass captured as a hipGraph, W warm-up replays, then K
    # replays, each between two events on the replay stream; value = median
    graph_ms, wall_ms = graph_timing(step, args.steps, args.warmup)
    ms_per_step = float(np.median(graph_ms))
    value = N / (ms_per_step / 1e3)
    log(f"[bench] pass: graph median {ms_per_step:.4f} ms (min {min(graph_ms):.4f}, max {max(graph_ms):.4f}), "
        f"eager {eager_ms:.4f} ms, wall {wall_ms:.4f} ms per replay")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "batched_4096_series_pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if tj.get("chunk_len") == chunk:
            traffic = sum(v["hbm_bytes_per_launch"] for k, v in tj.get("kernels", {}).items() if k in kern_ms)
    pass_gbs = PASS_BYTES_PER_STEP * N / (ms_per_step / 1e3) / 1e9
    fp64_tfs = PASS_FLOPS_PER_STEP * N / (ms_per_step / 1e3) / 1e12
    roof = {"bound": "hbm", "achieved": pass_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": pass_gbs / HBM_PEAK_GBS, "traffic": traffic,
            "scope": "whole filter+smoother pass (all its launches, one hipGraph replay)",
            "alg_bytes_per_step": PASS_BYTES_PER_STEP, "steps_per_pass": N, "ms_per_pass": ms_per_step,
            "fp64_valu": {"alg_flops_per_step": PASS_FLOPS_PER_STEP, "achieved_TFs": fp64_tfs,
                          "peak_TFs": FP64_PEAK_TFS, "frac": fp64_tfs / FP64_PEAK_TFS,
                          "note": "dense count of the reference recurrences, not the executed instructions"},
            "kernels": {k: {"avg_launch_ms": v, "alg_bytes_per_step": ALG_BYTES.get(k),
                            "hbm_frac": (ALG_BYTES[k] * N / (v / 1e3) / 1e9 / HBM_PEAK_GBS) if k in ALG_BYTES else None}
                        for k, v in kern_ms.items()}}

    # accuracy vs the FP64 oracle on the same (FP32) inputs
    post_mean = plan.scatter(pm).cpu().numpy().reshape(-1)
    post_var = plan.scatter(pv).cpu().numpy().reshape(-1)
    gpu_ll = ll.cpu().numpy()
    cpu = None
    err = None
    if not args.no_cpu:
        from oracle import cfast
        nth = cpu_threads()
        lib_path, prov = cfast.build_native(os.path.join(ROOT, "gpurun_out", "oracle_native"))
        c0 = time.perf_counter()
        om, ov, oll = cfast.filter_smoother(["matern52", "matern32"], data["ls"], data["var"], off,
                                            dt.astype(np.float64), y.astype(np.float64), r.astype(np.float64),
                                            nthreads=nth, lib_path=lib_path)
        csec = time.perf_counter() - c0
        cpu = {"value": N / csec, "unit": "time-steps/s", "cores": nth, "kind": "port",
               "sample": f"full workload: one FP64 filter+smoother pass over all {B} series ({N} steps), "
                         f"oracle/kfrts.c, OpenMP dynamic over series, {csec:.2f}s",
               **prov}
        err = {"mean_max_rel": float(np.max(np.abs(post_mean - om) / np.maximum(np.abs(om), 1.0))),
               "mean_max_abs": float(np.max(np.abs(post_mean - om))),
               "mean_median_abs": float(np.median(np.abs(post_mean - om))),
               "var_max_rel": float(np.max(np.abs(post_var - ov) / np.abs(ov))),
               "var_median_rel": float(np.median(np.abs(post_var - ov) / np.abs(ov))),
               "loglik_max_rel": float(np.max(np.abs(gpu_ll - oll) / np.abs(oll)))}

    warm = None
    if args.ab_runs > 0:
        from bayesnewton_amd import warmstart as W
        net, _ = W.WarmStartMLP.pretrained()
        seeds = list(range(args.st_heldout, args.st_heldout + args.ab_runs))
        try:
            warm = warmstart_ab(net, seeds, args.st_steps, args.st_tol, args.st_max_iter, args.st_lr, dev)
            warm["config"]["network_loaded_from"] = "bayesnewton_amd/pretrained/spacetime_cox_warmstart.npz"
        except Exception as exc:  # reported in the line, never hides the pass figures
            warm = {"error": f"{type(exc).__name__}: {exc}"}

    out = {
        "metric": "filter+smoother steps/s vs CPU; time-to-converge learned-warm-start vs plain",
        "value": value,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded GP draws + Student-t noise)",
        "config": {"workload": "batched uneven series: B=4096 Markov-GP series, lengths logU[256,16384], Matern52+Matern32 (d=5), "
                               "one fused filter+RTS-smoother pass per step",
                   "series": B, "time_steps": N, "chunk_len": chunk, "state_dim": 5,
                   "io": "f32 per-step inputs and posterior outputs; f64 recurrences, scan elements and log-likelihood"},
        "timing": {"method": "median of K hipGraph replays of one pass, hipEvents on the replay stream",
                   "graph_ms_median": ms_per_step, "graph_ms_min": min(graph_ms), "graph_ms_max": max(graph_ms),
                   "wall_ms_per_replay": wall_ms, "eager_ms_per_pass": eager_ms},
        "autotune": tune,
        "roofline": roof,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64_oracle": err,
        "warm_start": warm,
    }
    print(json.dumps(out), flush=True)


def graph_timing(step, steps, warmup):
    """Capture one call of `step` as a hipGraph (torch.cuda.graph: the library launches
    onto torch's current stream, which is the capture stream there), replay it `warmup`
    times, then time `steps` replays each between two events on the replay stream.
    Returns ([ms per replay], wall-clock ms per replay over the timed replays); the
    device is synchronised on both sides of the timed region."""
    import torch
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        g.replay()
        b.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return [a.elapsed_time(b) for a, b in ev], wall


if __name__ == "__main__":
    main()
