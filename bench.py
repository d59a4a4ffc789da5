"""Benchmark: batched filter + smoother time-steps/s on one MI355X.

Workload (one "step" = one fused filter + RTS smoother pass over the batch,
i.e. MarkovGaussianProcess.update_posterior, bayesnewton/basemodels.py:689-706,
for every series): B = 4096 independent Markov-GP series of lengths log-uniform
in [256, 16384] (sum 15.74M time steps), Matern52 + Matern32 sum kernel (state
dim 5) with one lengthscale per series, synthetic GP draws + 0.2 t_3 noise as
pseudo observations with pseudo variance 0.04 (the Student-t scale^2).  Inputs
are resident in HBM before the timed region.

Prints one JSON line.  `value` = total time steps / second over the K timed
passes.  `roofline` is the dominant kernel's algorithmic bytes per launch over
its average launch time, timed live with hipEvents recorded by libbn_hip.so on
the kernel's own stream.  `cpu_baseline` is the oracle's FP64 C filter +
smoother (oracle/kfrts.c, OpenMP over series) on the same inputs.
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
# Algorithmic bytes per time step (FP32 words, d = 5): the filter side moves
# dt, y, r in + filtered mean and packed covariance out = 4*3 + 4*(5+15) = 92 B;
# the smoother side reads the filtered moments back and writes the smoothed
# mean / variance = 4*(5+15) + 4*2 = 88 B; 180 B per filter+smoother step.
ALG_BYTES = {"k_filter_pass": 92.0, "k_smoother_pass": 88.0}
PASS_BYTES_PER_STEP = 180.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


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
        print("bench.py: this build runs on one GPU; --gpus N with N > 1 is not supported", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from bayesnewton_amd import _native as nat, kernels as K, ops, workloads
    from bayesnewton_amd.plan import SeriesPlan

    t_gen = time.time()
    data = workloads.c3()
    off = data["offsets"]
    N = int(off[-1])
    B = len(off) - 1
    dt = data["dt"].astype(np.float32)
    y = data["y"].astype(np.float32)
    r = np.full(N, data["scale"] ** 2, dtype=np.float32)
    print(f"[bench] C3 generated in {time.time() - t_gen:.1f}s: B={B} N={N}", file=sys.stderr, flush=True)

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    plan = SeriesPlan(off, chunk_len=args.chunk, device=dev)
    kern = K.Sum([K.Matern52(), K.Matern32()])
    desc = kern.desc()
    dt_t = plan.gather(torch.from_numpy(dt).to(dev))
    y_t = plan.gather(torch.from_numpy(y).to(dev))
    r_t = plan.gather(torch.from_numpy(r).to(dev), pad=1.0)
    ls = torch.tensor(data["ls"], dtype=torch.float64, device=dev)
    var = torch.tensor(data["var"], dtype=torch.float64, device=dev)
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
    L = nat.lib()
    L.bn_timing_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kern_ms = {}
    for kid in range(L.bn_num_kernel_ids()):
        tot = ctypes.c_double()
        cnt = ctypes.c_int32()
        nat.check(L.bn_timing_read(kid, ctypes.byref(tot), ctypes.byref(cnt)))
        if cnt.value:
            kern_ms[L.bn_kernel_name(kid).decode()] = (tot.value, cnt.value)
    L.bn_timing_enable(0)
    value = N / (ms_per_step / 1e3)

    dom = max(kern_ms, key=lambda k: kern_ms[k][0])
    dom_avg_ms = kern_ms[dom][0] / kern_ms[dom][1]
    dom_bytes = ALG_BYTES.get(dom, 0.0) * N
    achieved = dom_bytes / (dom_avg_ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom)

    # accuracy vs the FP64 oracle on the same (FP32) inputs
    post_mean = plan.scatter(pm).cpu().numpy().reshape(-1)
    post_var = plan.scatter(pv).cpu().numpy().reshape(-1)
    gpu_ll = ll.cpu().numpy()
    cpu = None
    err = None
    if not args.no_cpu:
        from oracle import cfast
        nth = cpu_threads()
        c0 = time.perf_counter()
        om, ov, oll = cfast.filter_smoother(["matern52", "matern32"], data["ls"], data["var"], off,
                                            dt.astype(np.float64), y.astype(np.float64), r.astype(np.float64),
                                            nthreads=nth)
        csec = time.perf_counter() - c0
        cpu = {"value": N / csec, "unit": "time-steps/s", "cores": nth, "kind": "port",
               "sample": f"full workload: one FP64 filter+smoother pass over all {B} series ({N} steps), "
                         f"oracle/kfrts.c -O3, OpenMP dynamic over series, {csec:.2f}s"}
        scale = np.maximum(np.abs(om), 1e-3)
        err = {"mean_max_rel": float(np.max(np.abs(post_mean - om) / np.maximum(np.abs(om), 1.0))),
               "mean_max_abs": float(np.max(np.abs(post_mean - om))),
               "var_max_rel": float(np.max(np.abs(post_var - ov) / np.abs(ov))),
               "var_median_rel": float(np.median(np.abs(post_var - ov) / np.abs(ov))),
               "loglik_max_rel": float(np.max(np.abs(gpu_ll - oll) / np.abs(oll)))}
        del scale

    out = {
        "metric": "filter+smoother time-steps/s (batched uneven series, config 3)",
        "value": value,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state recurrences, f32 per-step I/O",
        "data": "synthetic (seeded GP draws + Student-t noise)",
        "config": {"workload": "C3: B=4096 Markov-GP series, lengths logU[256,16384], Matern52+Matern32 (d=5), "
                               "one fused filter+RTS-smoother pass per step",
                   "series": B, "time_steps": N, "chunk_len": args.chunk, "state_dim": 5},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_step": ALG_BYTES.get(dom, 0.0), "avg_launch_ms": dom_avg_ms},
        "pass_roofline": {"alg_bytes_per_step": PASS_BYTES_PER_STEP,
                          "achieved_GBs": PASS_BYTES_PER_STEP * N / (ms_per_step / 1e3) / 1e9,
                          "frac": PASS_BYTES_PER_STEP * N / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_step": {k: v[0] / v[1] for k, v in kern_ms.items()},
        "cpu_baseline": cpu,
        "accuracy_vs_fp64_oracle": err,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
