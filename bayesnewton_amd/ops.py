"""Drop-in replacements for bayesnewton/ops.py's filter and smoother.

`kalman_filter` and `rauch_tung_striebel_smoother` keep the reference
signatures (bayesnewton/ops.py:256-285 and 357-380) and return FP64 tensors of
the reference's shapes; the work runs in libbn_hip.so.  `parallel=False`
selects the sequential scan (one lane walks the whole series), `parallel=True`
the chunked associative scan; both compute the same quantities.

`filter_smoother` is the fused path of MarkovGaussianProcess.update_posterior
(bayesnewton/basemodels.py:689-706) on a SeriesPlan: FP32 tiled pseudo
observations in, FP32 tiled posterior mean/variance and FP64 per-series
log-likelihoods out, no filtered state leaving the device scratch.
"""
import numpy as np
import torch

from . import _native as nat
from .plan import SeriesPlan


def _device():
    if not torch.cuda.is_available():
        raise nat.NativeError("bayesnewton_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _as_tensor(x, dtype, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype)
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


def discretise(kernel, dt):
    """vmap(kernel.state_transition)(dt) and Q = Pinf - A Pinf A' (ops.py:274-279).

    Returns (A [N,d,d], Q [N,d,d]) as FP32 device tensors computed in FP64.
    """
    dev = _device()
    dt = _as_tensor(dt, torch.float32, dev).reshape(-1).contiguous()
    n = dt.shape[0]
    d = kernel.state_dim
    A = torch.empty((n, d, d), dtype=torch.float32, device=dev)
    Q = torch.empty((n, d, d), dtype=torch.float32, device=dev)
    desc = kernel.desc()
    ls, var = kernel.hypers(dev)
    import ctypes
    nat.check(nat.lib().bn_discretise(ctypes.byref(desc), nat.ptr(ls), nat.ptr(var), n, nat.ptr(dt), nat.ptr(A),
                                      nat.ptr(Q), nat.stream_ptr()))
    return A, Q


def _single_plan(n, parallel, dev):
    return SeriesPlan(np.array([0, n], dtype=np.int64), parallel=parallel, device=dev)


def kalman_filter(dt, kernel, y, noise_cov, mask=None, parallel=False, return_predict=False, plan=None,
                  status=None):
    """ops.py:256-285.  dt [N(,1)], y [N,1,1], noise_cov [N,1,1], mask [N,1,1] bool.

    Returns (ell, (means [N,d,1], covs [N,d,d])), FP64 tensors on the GPU.
    
    status: optional int32 device tensor [n_series] that receives the BN_STATUS_* bits
    (a non-PD innovation or a non-finite log-likelihood; the outputs carry NaN as in
    the reference).
    """
    import ctypes
    dev = _device()
    dt = _as_tensor(dt, torch.float32, dev).reshape(-1)
    N = dt.shape[0]
    y = _as_tensor(y, torch.float32, dev).reshape(N, -1)
    r = _as_tensor(noise_cov, torch.float32, dev).reshape(N, -1)
    if y.shape[1] != 1 or r.shape[1] != 1:
        raise NotImplementedError("the temporal HIP path takes scalar observations (y [N,1,1])")
    if plan is None:
        plan = _single_plan(N, parallel, dev)
    d = kernel.state_dim
    desc = kernel.desc()
    ls, var = kernel.hypers(dev)
    dt_t, y_t, r_t = plan.gather(dt), plan.gather(y), plan.gather(r, pad=1.0)
    mk = plan.gather_mask(_as_tensor(mask, torch.bool, dev)) if mask is not None else None
    means = torch.empty((d, plan.n_slots), dtype=torch.float64, device=dev)
    covs = torch.empty((d * d, plan.n_slots), dtype=torch.float64, device=dev)
    ll = torch.empty(plan.n_series, dtype=torch.float64, device=dev)
    flags = nat.BN_RETURN_PREDICT if return_predict else 0
    nat.check(nat.lib().bn_kalman_filter(plan.handle, ctypes.byref(desc), nat.ptr(ls), nat.ptr(var), 0,
                                         nat.ptr(dt_t), nat.ptr(y_t), nat.ptr(r_t), nat.ptr(mk), nat.ptr(means),
                                         nat.ptr(covs), nat.ptr(ll), flags, nat.ptr(status),
                                         nat.ptr(plan.workspace(desc)), nat.stream_ptr()))
    m = plan.scatter(means).reshape(N, d, 1)
    P = plan.scatter(covs).reshape(N, d, d)
    return ll.sum(), (m, P)


def rauch_tung_striebel_smoother(dt, kernel, filter_mean, filter_cov, return_full=False, parallel=False, plan=None):
    """ops.py:357-380.  dt is the smoother's (shifted) step vector.

    Returns (means, covs, gains): [N,p,1], [N,p,p] (or [N,d,1], [N,d,d] with
    return_full) and gains [N,d,d], FP64 tensors on the GPU.
    """
    import ctypes
    dev = _device()
    dt = _as_tensor(dt, torch.float32, dev).reshape(-1)
    N = dt.shape[0]
    d = kernel.state_dim
    fm = _as_tensor(filter_mean, torch.float64, dev).reshape(N, d)
    fP = _as_tensor(filter_cov, torch.float64, dev).reshape(N, d * d)
    if plan is None:
        plan = _single_plan(N, parallel, dev)
    desc = kernel.desc()
    ls, var = kernel.hypers(dev)
    dt_t, fm_t, fP_t = plan.gather(dt), plan.gather(fm), plan.gather(fP)
    w_m, w_P = (d, d * d) if return_full else (1, 1)
    om = torch.empty((w_m, plan.n_slots), dtype=torch.float64, device=dev)
    oP = torch.empty((w_P, plan.n_slots), dtype=torch.float64, device=dev)
    G = torch.empty((d * d, plan.n_slots), dtype=torch.float64, device=dev)
    flags = nat.BN_RETURN_FULL if return_full else 0
    nat.check(nat.lib().bn_rts_smoother(plan.handle, ctypes.byref(desc), nat.ptr(ls), nat.ptr(var), 0,
                                        nat.ptr(dt_t), nat.ptr(fm_t), nat.ptr(fP_t), nat.ptr(om), nat.ptr(oP),
                                        nat.ptr(G), flags, nat.ptr(plan.workspace(desc)), nat.stream_ptr()))
    if return_full:
        means = plan.scatter(om).reshape(N, d, 1)
        covs = plan.scatter(oP).reshape(N, d, d)
    else:
        means = plan.scatter(om).reshape(N, 1, 1)
        covs = plan.scatter(oP).reshape(N, 1, 1)
    return means, covs, plan.scatter(G).reshape(N, d, d)


def filter_smoother(plan, kernel_desc, lengthscale, variance, dt, y, r, mask=None, hyper_stride=0,
                    post_mean=None, post_var=None, loglik=None, status=None, noalloc=False):
    """Fused filter + smoother on tiled FP32 inputs (update_posterior, basemodels.py:689-706).

    lengthscale / variance: FP64 device tensors [n_comp] (hyper_stride 0) or
    [B, n_comp] (hyper_stride n_comp).  status: optional int32 [n_series] tensor for the
    BN_STATUS_* bits.  noalloc: run bn_filter_smoother_noalloc (scratch inside the library;
    outputs must then be given to allocate nothing).  Returns (post_mean, post_var, loglik).
    """
    import ctypes
    dev = plan.device
    if post_mean is None:
        post_mean = torch.empty(plan.n_slots, dtype=torch.float32, device=dev)
    if post_var is None:
        post_var = torch.empty(plan.n_slots, dtype=torch.float32, device=dev)
    if loglik is None:
        loglik = torch.empty(plan.n_series, dtype=torch.float64, device=dev)
    args = (plan.handle, ctypes.byref(kernel_desc), nat.ptr(lengthscale), nat.ptr(variance), hyper_stride, nat.ptr(dt),
            nat.ptr(y), nat.ptr(r), nat.ptr(mask), nat.ptr(post_mean), nat.ptr(post_var), nat.ptr(loglik),
            nat.ptr(status))
    if noalloc:
        nat.check(nat.lib().bn_filter_smoother_noalloc(*args, nat.stream_ptr()))
    else:
        nat.check(nat.lib().bn_filter_smoother(*args, nat.ptr(plan.workspace(kernel_desc)), nat.stream_ptr()))
    return post_mean, post_var, loglik


def capture_graph(fn):
    """Record the launches of fn() (library calls on torch's current stream) as one
    hipGraph: fn runs once on a side stream first, then once under capture.  Returns
    the torch.cuda.CUDAGraph; g.replay() re-runs every launch on the current stream
    with the same pointers, so fn must only touch tensors that outlive the graph
    (the library's entry points allocate nothing on the device and record no
    memset, so a fused pass is capturable as it stands)."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    return g


def filter_loglik(plan, kernel_desc, lengthscale, variance, dt, y, r, mask=None, hyper_stride=0, loglik=None):
    """Per-series filter log-likelihood on tiled FP32 inputs (compute_log_lik, basemodels.py:726-741)."""
    import ctypes
    if loglik is None:
        loglik = torch.empty(plan.n_series, dtype=torch.float64, device=plan.device)
    nat.check(nat.lib().bn_filter_loglik(plan.handle, ctypes.byref(kernel_desc), nat.ptr(lengthscale),
                                         nat.ptr(variance), hyper_stride, nat.ptr(dt), nat.ptr(y), nat.ptr(r),
                                         nat.ptr(mask), nat.ptr(loglik), nat.ptr(plan.workspace(kernel_desc)),
                                         nat.stream_ptr()))
    return loglik


# ---- mean-field space-time filter / smoother (ops.py:581-608, 681-706) -------------------

def _mf_setup(dt, kernel, n):
    from .kernels import SpatioTemporalKernel
    if not isinstance(kernel, SpatioTemporalKernel):
        raise TypeError("the mean-field filter needs a SpatioTemporalKernel")
    dev = _device()
    kt = kernel.temporal_kernel
    dims = nat.StDims(1, n, 1, kernel.M, kt.kind)
    f64 = dict(dtype=torch.float64, device=dev)
    ls = torch.tensor([kt.lengthscale], **f64)
    var = torch.tensor([kt.variance], **f64)
    hyp = nat.StHypers(nat.ptr(ls), nat.ptr(var), nat.ptr(var), 0)
    dt = _as_tensor(dt, torch.float64, dev).reshape(-1).contiguous()
    if dt.shape[0] != n:
        raise ValueError("dt must have one entry per step")
    return dev, dims, hyp, (ls, var), dt, kt.state_dim


def kalman_filter_meanfield(dt, kernel, y, noise_cov, mask=None, parallel=False, block_index=None):
    """ops.py:581-608.  y [N,M,1], noise_cov [N,M,M] for a SpatioTemporalKernel with M
    inducing points.  Returns (ell, (means [N,M,KT,1], covs [N,M,KT,KT])), FP64 tensors on
    the GPU.  block_index is implied by the kernel (kernels.py:573-580) and ignored.
    Masks are not supported (the space-time models never mask their pseudo observations)."""
    import ctypes
    if mask is not None and bool(np.any(np.asarray(mask.cpu() if isinstance(mask, torch.Tensor) else mask))):
        raise NotImplementedError("masked observations are not supported by the mean-field filter")
    n = int(y.shape[0])
    dev, dims, hyp, keep, dt, KT = _mf_setup(dt, kernel, n)
    M = kernel.M
    yv = _as_tensor(y, torch.float64, dev).reshape(n, M).contiguous()
    R = _as_tensor(noise_cov, torch.float64, dev).reshape(n, M, M).contiguous()
    f64 = dict(dtype=torch.float64, device=dev)
    fm = torch.empty((n, M, KT), **f64)
    fP = torch.empty((n, M, KT, KT), **f64)
    ell = torch.empty(n, **f64)
    ll = torch.empty(1, **f64)
    ws = None
    if parallel:
        nb = nat.lib().bn_mf_workspace_bytes(ctypes.byref(dims))
        nat.check(0 if nb > 0 else -1)
        ws = torch.empty(int(nb), dtype=torch.uint8, device=dev)
    nat.check(nat.lib().bn_mf_filter(ctypes.byref(dims), ctypes.byref(hyp), 1 if parallel else 0, nat.ptr(dt),
                                     nat.ptr(yv), nat.ptr(R), nat.ptr(fm), nat.ptr(fP), nat.ptr(ell), nat.ptr(ll),
                                     nat.ptr(ws), nat.stream_ptr()))
    return ll[0], (fm[..., None], fP)


def rauch_tung_striebel_smoother_meanfield(dt, kernel, filter_mean, filter_cov, return_full=False, parallel=False,
                                           block_index=None):
    """ops.py:681-706 (dt is the smoother's step vector, as the reference takes it).
    filter_mean [N,M,KT(,1)], filter_cov [N,M,KT,KT].  Returns (means, covs, gains) with the
    reference's shapes: return_full -> [N,M*KT,1], [N,M*KT,M*KT] and block-diagonal gains;
    else H m [N,M,1] and H P H' [N,M,M] (diagonal).  Both forms give the same result."""
    import ctypes
    n = int(filter_cov.shape[0])
    dev, dims, hyp, keep, dt, KT = _mf_setup(dt, kernel, n)
    M = kernel.M
    fm = _as_tensor(filter_mean, torch.float64, dev).reshape(n, M, KT).contiguous()
    fP = _as_tensor(filter_cov, torch.float64, dev).reshape(n, M, KT, KT).contiguous()
    f64 = dict(dtype=torch.float64, device=dev)
    gains = torch.empty((n, M, KT, KT), **f64)
    if return_full:
        sm = torch.empty((n, M, KT), **f64)
        sP = torch.empty((n, M, KT, KT), **f64)
        nat.check(nat.lib().bn_mf_smoother(ctypes.byref(dims), ctypes.byref(hyp), nat.ptr(dt), nat.ptr(fm),
                                           nat.ptr(fP), None, None, 0, 0, nat.ptr(sm), nat.ptr(sP), nat.ptr(gains),
                                           nat.stream_ptr()))
        return sm.reshape(n, M * KT, 1), _block_diag(sP), _block_diag(gains)
    hm = torch.empty((n, M), **f64)
    hv = torch.zeros((n, M, M), **f64)
    nat.check(nat.lib().bn_mf_smoother(ctypes.byref(dims), ctypes.byref(hyp), nat.ptr(dt), nat.ptr(fm), nat.ptr(fP),
                                       nat.ptr(hm), nat.ptr(hv), M * M, M + 1, None, None, nat.ptr(gains),
                                       nat.stream_ptr()))
    return hm[..., None], hv, _block_diag(gains)


def _block_diag(blocks):
    """[N,M,d,d] -> [N,M d,M d] block-diagonal (build_block_diag, ops.py:644-647)."""
    n, M, d = blocks.shape[0], blocks.shape[1], blocks.shape[2]
    out = torch.zeros((n, M * d, M * d), dtype=blocks.dtype, device=blocks.device)
    idx = torch.arange(M, device=blocks.device)
    for a in range(d):
        for b in range(d):
            out[:, idx * d + a, idx * d + b] = blocks[:, :, a, b]
    return out


# ---- infinite-horizon space-time filter / smoother (ops.py:877-953, 1018-1068) ------------

def _ih_setup(kernel, n, step):
    from .kernels import SpatioTemporalKernel
    from .spacetime import state_perm, _pinf_temporal
    if not isinstance(kernel, SpatioTemporalKernel):
        raise TypeError("the infinite-horizon filter here needs a SpatioTemporalKernel")
    dev = _device()
    kt = kernel.temporal_kernel
    M, KT = kernel.M, kt.state_dim
    dims = nat.StDims(1, n, 1, M, kt.kind)
    f64 = dict(dtype=torch.float64, device=dev)
    ls = torch.tensor([kt.lengthscale], **f64)
    var = torch.tensor([kt.variance], **f64)
    hyp = nat.StHypers(nat.ptr(ls), nat.ptr(var), nat.ptr(var), 0)
    p = torch.as_tensor(state_perm(M, KT), device=dev)
    pinf = torch.tensor(np.kron(np.eye(M), _pinf_temporal(kt)), **f64)
    st = torch.tensor([float(step)], **f64)
    return dev, dims, hyp, (ls, var), p, torch.argsort(p), pinf, st, M, KT


def kalman_filter_infinite_horizon(dt, kernel, y, noise_cov, mask=None, parallel=False, heteroscedastic=False,
                                   noise_cov_tied=None, dare_iters=20, dare_init=None):
    """ops.py:877-953 for a SpatioTemporalKernel (state in the reference's order).
    y [N,M,1], noise_cov [N,M,M], noise_cov_tied [M,M], dare_init [D,D] (default Pinf).
    Returns (ell, (means [N,D,1], (Pdare, cov))), FP64 tensors on the GPU.  Masks are not
    supported; both of the reference's forms compute the same means."""
    import ctypes
    if mask is not None and bool(np.any(np.asarray(mask.cpu() if isinstance(mask, torch.Tensor) else mask))):
        raise NotImplementedError("masked observations are not supported by the infinite-horizon filter")
    if noise_cov_tied is None:
        raise ValueError("noise_cov_tied is required")
    dtv = np.asarray(dt.cpu() if isinstance(dt, torch.Tensor) else dt, dtype=np.float64).reshape(-1)
    n = dtv.shape[0]
    dev, dims, hyp, keep, p, ip, pinf, st, M, KT = _ih_setup(kernel, n, dtv[1])
    D = M * KT
    f64 = dict(dtype=torch.float64, device=dev)
    yv = _as_tensor(y, torch.float64, dev).reshape(n, M).contiguous()
    R = _as_tensor(noise_cov, torch.float64, dev).reshape(n, M, M).contiguous() if heteroscedastic else None
    Rt = _as_tensor(noise_cov_tied, torch.float64, dev).reshape(M, M).contiguous()
    X0 = pinf if dare_init is None else _as_tensor(dare_init, torch.float64, dev).reshape(D, D)
    X = X0[p][:, p].contiguous()
    cov = torch.empty((D, D), **f64)
    fm = torch.empty((n, D), **f64)
    ell = torch.empty(n, **f64)
    ll = torch.empty(1, **f64)
    nb = nat.lib().bn_ih_workspace_bytes(ctypes.byref(dims), 1 if heteroscedastic else 0)
    nat.check(0 if nb > 0 else -1)
    ws = torch.empty(int(nb), dtype=torch.uint8, device=dev)
    nat.check(nat.lib().bn_ih_filter(ctypes.byref(dims), ctypes.byref(hyp), nat.ptr(st), int(dare_iters),
                                     1 if heteroscedastic else 0, nat.ptr(yv), nat.ptr(R), nat.ptr(Rt), nat.ptr(X),
                                     nat.ptr(cov), nat.ptr(fm), nat.ptr(ell), nat.ptr(ll), nat.ptr(ws),
                                     nat.stream_ptr()))
    back = lambda A: A[ip][:, ip]  # noqa: E731
    return ll[0], (fm[:, ip][..., None], (back(X), back(cov)))


def rauch_tung_striebel_smoother_infinite_horizon(dt, kernel, filter_mean, filter_cov, return_full=False,
                                                  parallel=False, dare_iters=20, dare_init=None):
    """ops.py:1018-1068 for a SpatioTemporalKernel (dt is the smoother's step vector; the
    step size is its first entry).  filter_cov = (Pdare, cov) as the filter returns it.
    Returns (means, covs, gains, dare_cov) with the reference's shapes."""
    import ctypes
    dtv = np.asarray(dt.cpu() if isinstance(dt, torch.Tensor) else dt, dtype=np.float64).reshape(-1)
    n = dtv.shape[0]
    dev, dims, hyp, keep, p, ip, pinf, st, M, KT = _ih_setup(kernel, n, dtv[0])
    D = M * KT
    f64 = dict(dtype=torch.float64, device=dev)
    fm = _as_tensor(filter_mean, torch.float64, dev).reshape(n, D)[:, p].contiguous()
    Pd = _as_tensor(filter_cov[0], torch.float64, dev).reshape(D, D)[p][:, p].contiguous()
    fc = _as_tensor(filter_cov[1], torch.float64, dev).reshape(D, D)[p][:, p].contiguous()
    X0 = pinf if dare_init is None else _as_tensor(dare_init, torch.float64, dev).reshape(D, D)
    Xs = X0[p][:, p].contiguous()
    G = torch.empty((D, D), **f64)
    sm = torch.empty((n, D), **f64)
    nb = nat.lib().bn_ih_workspace_bytes(ctypes.byref(dims), 0)
    nat.check(0 if nb > 0 else -1)
    ws = torch.empty(int(nb), dtype=torch.uint8, device=dev)
    nat.check(nat.lib().bn_ih_smoother(ctypes.byref(dims), ctypes.byref(hyp), nat.ptr(st), int(dare_iters),
                                       nat.ptr(fm), nat.ptr(Pd), nat.ptr(fc), nat.ptr(Xs), nat.ptr(G), None, None,
                                       nat.ptr(sm), nat.ptr(ws), nat.stream_ptr()))
    back = lambda A: A[ip][:, ip]  # noqa: E731
    dare_cov = back(Xs)
    gain = back(G)
    if return_full:
        means, cov = sm[:, ip][..., None], dare_cov
    else:
        means, cov = sm[:, :M][..., None], Xs[:M, :M]
    return means, cov.expand(n, *cov.shape).contiguous(), gain.expand(n, D, D).contiguous(), dare_cov


# ---- Kalman filter over pairs of states (ops.py:383-426) ----------------------------------

def kalman_filter_pairs(dt, kernel, y, noise_cov, mask=None, parallel=False):
    """ops.py:383-426 (the filter of the sparse Markov model): y [N,2d,1], noise_cov
    [N,2d,2d].  Returns (ell, (means [N-1,d,1], covs [N-1,d,d])), FP64 tensors on the GPU:
    the first state of each pair after the first step.  Both of the reference's forms give
    this result, so `parallel` is accepted for the signature and both run the one
    sequential kernel (k_pairs_filter); the associative pairs form with its 1e-32
    jitter (ops.py:405-426) is not built.  Masks are not supported."""
    import ctypes
    if mask is not None and bool(np.any(np.asarray(mask.cpu() if isinstance(mask, torch.Tensor) else mask))):
        raise NotImplementedError("masked observations are not supported by the pairs filter")
    dev = _device()
    d = kernel.state_dim
    dtv = _as_tensor(dt, torch.float64, dev).reshape(-1).contiguous()
    n = dtv.shape[0]
    yv = _as_tensor(y, torch.float64, dev).reshape(n, 2 * d).contiguous()
    R = _as_tensor(noise_cov, torch.float64, dev).reshape(n, 2 * d, 2 * d).contiguous()
    f64 = dict(dtype=torch.float64, device=dev)
    off = torch.tensor([0, n], dtype=torch.int64, device=dev)
    mo = torch.empty((n, 2 * d), **f64)
    Po = torch.empty((n, 2 * d, 2 * d), **f64)
    ll = torch.empty(1, **f64)
    desc = kernel.desc()
    ls, var = kernel.hypers(dev)
    nat.check(nat.lib().bn_pairs_filter(ctypes.byref(desc), nat.ptr(ls), nat.ptr(var), 0, 1, nat.ptr(off),
                                        nat.ptr(dtv), nat.ptr(yv), nat.ptr(R), nat.ptr(mo), nat.ptr(Po), nat.ptr(ll),
                                        nat.stream_ptr()))
    return ll[0], (mo[1:, :d, None], Po[1:, :d, :d])
