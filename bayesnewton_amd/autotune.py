"""Chunk-length autotuner for the fused filter + smoother pass over a batch of series.

The pass (update_posterior, bayesnewton/basemodels.py:689-706, as one launch sequence
per call) cuts every series into chunks of `chunk_len` steps, one lane per chunk, and
scans the chunks' elements.  Short chunks give more lanes but more scan work and more
chunk-boundary state; long chunks the reverse, and where the balance lies depends on
the batch (series count and lengths) and on the kernel's state dimension.  The tuner
times a few candidate lengths on the caller's own inputs with device events and keeps
the fastest; results are identical across chunk lengths up to rounding.
"""
import numpy as np
import torch

from . import _native as nat
from . import ops
from .plan import SeriesPlan

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
