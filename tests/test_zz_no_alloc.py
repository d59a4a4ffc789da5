"""bn_filter_smoother_noalloc (scratch in the library's own device array) against
bn_filter_smoother (caller-provided workspace), called through the C ABI: two calls
in a row on different plans, the second with a smaller input, results bit-identical.
Plans whose scratch would not fit are refused on the host."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _series(lens, seed):
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for n in lens:
        x = np.sort(rng.uniform(0.0, 0.05 * n + 1.0, n))
        y = np.sin(0.7 * x) + 0.3 * rng.standard_normal(n)
        if n > 10:
            y[rng.choice(n, n // 10, replace=False)] = np.nan
        xs.append(x)
        ys.append(y)
    return xs, ys


def _run(lens, seed, chunk_len):
    import torch
    from bayesnewton_amd import _native as nat, kernels as K
    from bayesnewton_amd.plan import SeriesPlan
    dev = torch.device("cuda", 0)
    xs, ys = _series(lens, seed)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    dt = np.concatenate([np.concatenate([[0.0], np.diff(x)]) if len(x) else np.zeros(0) for x in xs])
    y = np.concatenate(ys)
    rng = np.random.default_rng(seed + 1)
    r = rng.uniform(0.05, 0.5, off[-1])
    plan = SeriesPlan(off, chunk_len=chunk_len, device=dev)
    f32 = dict(dtype=torch.float32, device=dev)
    dt_t = plan.gather(torch.as_tensor(dt, **f32))
    y_t = plan.gather(torch.as_tensor(np.nan_to_num(y), **f32))
    r_t = plan.gather(torch.as_tensor(r, **f32), pad=1.0)
    mask = plan.gather_mask(torch.as_tensor(np.isnan(y), device=dev))
    kern = K.Sum([K.Matern52(variance=1.0, lengthscale=1.3), K.Matern32(variance=0.4, lengthscale=4.0)])
    desc = kern.desc()
    ls = torch.tensor([[1.3, 4.0]] * len(lens), dtype=torch.float64, device=dev)
    ls[:, 0] *= torch.linspace(0.8, 1.2, len(lens), dtype=torch.float64, device=dev)
    var = torch.tensor([[1.0, 0.4]] * len(lens), dtype=torch.float64, device=dev)
    L = nat.lib()
    outs = []
    for noalloc in (False, True):
        pm = torch.full((plan.n_slots,), float("nan"), **f32)
        pv = torch.full((plan.n_slots,), float("nan"), **f32)
        ll = torch.full((len(lens),), float("nan"), dtype=torch.float64, device=dev)
        st = torch.full((len(lens),), -1, dtype=torch.int32, device=dev)
        args = (plan.handle, ctypes.byref(desc), nat.ptr(ls), nat.ptr(var), 2, nat.ptr(dt_t), nat.ptr(y_t),
                nat.ptr(r_t), nat.ptr(mask), nat.ptr(pm), nat.ptr(pv), nat.ptr(ll), nat.ptr(st))
        print(f"[no_alloc] lens={list(lens)} chunk_len={chunk_len} noalloc={noalloc}", flush=True)
        if noalloc:
            nat.check(L.bn_filter_smoother_noalloc(*args, nat.stream_ptr()))
        else:
            nat.check(L.bn_filter_smoother(*args, nat.ptr(plan.workspace(desc)), nat.stream_ptr()))
        torch.cuda.synchronize()
        outs.append((plan.scatter(pm).cpu(), plan.scatter(pv).cpu(), ll.cpu()))
        assert torch.equal(st.cpu(), torch.zeros(len(lens), dtype=torch.int32)), st
    return outs


def test_noalloc_matches_workspace_entry_on_two_calls():
    import torch
    # first call: long ragged series over many chunks (both scans run); second: a
    # smaller batch on another plan, reusing the scratch the first call filled
    for lens, seed, chunk in (([5000, 1200, 0, 300, 1, 2600], 11, 32), ([700, 90, 3], 12, 16)):
        (pm, pv, ll), (pm2, pv2, ll2) = _run(lens, seed, chunk)
        assert torch.isfinite(pm).all() and torch.isfinite(pv).all() and torch.isfinite(ll).all()
        assert torch.equal(pm, pm2) and torch.equal(pv, pv2) and torch.equal(ll, ll2), lens


def test_noalloc_refuses_a_plan_beyond_its_scratch():
    import torch
    from bayesnewton_amd import _native as nat, kernels as K
    from bayesnewton_amd.plan import SeriesPlan
    dev = torch.device("cuda", 0)
    n = 600_000
    plan = SeriesPlan(np.array([0, n], dtype=np.int64), chunk_len=64, device=dev)
    desc = K.Sum([K.Matern52(), K.Matern32()]).desc()
    need = int(nat.lib().bn_workspace_bytes(plan.handle, ctypes.byref(desc)))
    assert need > nat.BN_NOALLOC_SCRATCH_BYTES
    buf = torch.zeros(plan.n_slots, dtype=torch.float32, device=dev)
    hyp = torch.ones(2, dtype=torch.float64, device=dev)
    ll = torch.zeros(1, dtype=torch.float64, device=dev)
    rc = nat.lib().bn_filter_smoother_noalloc(plan.handle, ctypes.byref(desc), nat.ptr(hyp), nat.ptr(hyp), 0,
                                              nat.ptr(buf), nat.ptr(buf), nat.ptr(buf), None, nat.ptr(buf),
                                              nat.ptr(buf), nat.ptr(ll), None, nat.stream_ptr())
    assert rc == -1 and b"scratch" in nat.lib().bn_last_error()
