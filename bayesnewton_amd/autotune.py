"""Schedule autotuner for the fused filter + smoother pass over a batch of series.

The pass (update_posterior, bayesnewton/basemodels.py:689-706, as one launch sequence
per call) cuts every series into chunks of `chunk_len` steps, one lane per chunk, and
scans the chunks' elements.  Short chunks give more lanes but more scan work and more
chunk-boundary state; long chunks the reverse, and where the balance lies depends on
the batch (series count and lengths) and on the kernel's state dimension.  The tuner
times a few candidate lengths on the caller's own inputs with device events and keeps
the fastest; results are identical across chunk lengths up to rounding.  The
sequential schedule (one lane per whole series, no scans) competes where it can win
(tune_schedule), and tuned choices are kept in a table shipped with the package
(tuned/schedules.json) so a known batch shape needs no timing.
"""
import hashlib
import json
import os

import numpy as np
import torch

from . import _native as nat
from . import ops
from .plan import SeriesPlan

# Tuned choices committed with the package, keyed by batch_signature(); written by
# record_schedule() from a tuning run on the device they name.
SCHEDULE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "schedules.json")

CANDIDATES = (32, 64, 128, 256)
# Waves per SIMD that the chunk-walking kernels (k_filter_reduce, k_filter_pass) reach
# at their register budget; used only to propose the fill candidates below.
WAVES_PER_SIMD = 2


def fill_chunk_len(lengths, n_cu, rounds=1, waves_per_simd=WAVES_PER_SIMD, lo=16, hi=4096):
    """Smallest chunk length whose chunks fit `rounds` rounds of resident wavefronts.

    The chunk-walking kernels put 64 chunks on a wavefront and every wavefront walks
    chunk_len steps, so their time is about chunk_len x the number of rounds of
    resident wavefronts (n_cu x 4 SIMDs x waves_per_simd) the chunks need.  The
    smallest length that still fits in `rounds` rounds leaves no partly filled round.
    Returns None when no length in [lo, hi] fits."""
    n = np.asarray(lengths, dtype=np.int64)
    cap = int(n_cu) * 4 * int(waves_per_simd) * int(rounds) * 64
    nchunks = lambda c: int(np.sum(-(-n // c)))  # noqa: E731
    if nchunks(hi) > cap:
        return None
    a, b = lo, hi
    while a < b:
        m = (a + b) // 2
        if nchunks(m) <= cap:
            b = m
        else:
            a = m + 1
    return a


def tune_chunk_len(offsets, kernel_desc, lengthscale, variance, dt, y, r, hyper_stride=0, candidates=CANDIDATES,
                   reps=10, device=None):
    """Chunk length of the fastest fused pass among `candidates`.

    offsets: int64 [B+1] series boundaries; dt, y, r: [N] FP32 device tensors in
    series-contiguous order (the layout the model gathers from); lengthscale / variance
    and hyper_stride as for ops.filter_smoother.  Each candidate runs one untimed pass,
    then `reps` passes between two events on the current stream.  With the default
    candidates, the fill lengths for one and two rounds of resident wavefronts on this
    device (fill_chunk_len) are timed as well.  Returns (best, {chunk_len: ms per
    pass}); candidates the plan rejects (a series longer than 128 x 128 chunks) are
    skipped."""
    dev = device if device is not None else dt.device
    offsets = np.asarray(offsets, dtype=np.int64)
    if candidates is CANDIDATES:
        n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
        extra = [fill_chunk_len(np.diff(offsets), n_cu, rounds=k) for k in (1, 2)]
        candidates = sorted(set(CANDIDATES) | {c for c in extra if c is not None})
    times = {}
    for c in candidates:
        try:
            plan = SeriesPlan(offsets, chunk_len=int(c), device=dev)
        except nat.NativeError:
            continue
        dt_t = plan.gather(dt)
        y_t = plan.gather(y)
        r_t = plan.gather(r, pad=1.0)
        pm = torch.empty(plan.n_slots, dtype=torch.float32, device=dev)
        pv = torch.empty(plan.n_slots, dtype=torch.float32, device=dev)
        ll = torch.empty(plan.n_series, dtype=torch.float64, device=dev)

        def run():
            ops.filter_smoother(plan, kernel_desc, lengthscale, variance, dt_t, y_t, r_t, None,
                                hyper_stride=hyper_stride, post_mean=pm, post_var=pv, loglik=ll)

        run()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        e1.synchronize()
        times[int(c)] = e0.elapsed_time(e1) / reps
        del plan, dt_t, y_t, r_t, pm, pv, ll
    if not times:
        raise nat.NativeError("no candidate chunk length fits this batch")
    best = min(times, key=times.get)
    return best, times


def batch_signature(lengths, kernel_desc, device=None):
    """Key of a batch for the schedule table: device architecture and compute-unit
    count, the kernel's component kinds, the series count, the total and longest
    length and a hash of the log2 length histogram (batches that share it cost the same
    to within the tuner's resolution)."""
    n = np.asarray(lengths, dtype=np.int64)
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    props = torch.cuda.get_device_properties(dev)
    arch = getattr(props, "gcnArchName", props.name).split(":")[0]
    kinds = "-".join(str(kernel_desc.kind[i]) for i in range(kernel_desc.n_comp))
    hist = np.bincount(np.floor(np.log2(np.maximum(n, 1))).astype(np.int64), minlength=32)
    h = hashlib.sha1(hist.tobytes()).hexdigest()[:12]
    return f"{arch}/cu{props.multi_processor_count}/kinds{kinds}/B{n.size}/N{int(n.sum())}/max{int(n.max(initial=0))}/h{h}"


def load_table(path=SCHEDULE_TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def record_schedule(signature, choice, path=SCHEDULE_TABLE):
    """Add (or replace) a tuned choice in the schedule table."""
    table = load_table(path)
    table[signature] = choice
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def tune_schedule(offsets, kernel_desc, lengthscale, variance, dt, y, r, hyper_stride=0, device=None, table=True,
                  reps=10):
    """Schedule of the fused pass for this batch: 'sequential' (one lane walks a whole
    series: the sequential filter and smoother, no scans) or 'chunked' (series cut into
    chunks of chunk_len steps, one lane per chunk, associative scans over the chunks'
    elements), with the chunk length.  With table=True a choice recorded for this
    batch's signature (schedule table shipped with the package) is used as is;
    otherwise the candidates are timed on the caller's inputs (tune_chunk_len) and the
    sequential schedule beside them when every lane-per-series wavefront can be
    resident at once and the longest series is at most 8 x the longest chunk length
    timed.  Returns a dict {"schedule", "chunk_len", "ms_per_pass", "source"}."""
    offsets = np.asarray(offsets, dtype=np.int64)
    lengths = np.diff(offsets)
    sig = batch_signature(lengths, kernel_desc, device)
    if table:
        hit = load_table().get(sig)
        if hit is not None:
            return dict(hit, source="table", signature=sig)
    best, times = tune_chunk_len(offsets, kernel_desc, lengthscale, variance, dt, y, r, hyper_stride=hyper_stride,
                                 reps=reps, device=device)
    out = {"schedule": "chunked", "chunk_len": int(best), "ms_per_pass": {str(k): v for k, v in times.items()}}
    dev = device if device is not None else dt.device
    n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
    nmax = int(lengths.max(initial=0))
    if lengths.size <= n_cu * 4 * 2 * 64 and nmax <= 8 * max(times):
        seq, seq_t = tune_chunk_len(offsets, kernel_desc, lengthscale, variance, dt, y, r, hyper_stride=hyper_stride,
                                    candidates=(max(nmax, 1),), reps=max(1, reps // 5), device=device)
        out["ms_per_pass"]["sequential"] = seq_t[seq]
        if seq_t[seq] < times[best]:
            out.update(schedule="sequential", chunk_len=int(seq))
    out["source"] = "timed"
    out["signature"] = sig
    return out
