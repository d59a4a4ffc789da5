"""Per-series status bits of the fused pass and the filter (BN_STATUS_*) on the GPU:
forced failures in one batch (a negative noise variance, i.e. a negative site precision,
and a NaN observation on an unmasked step) set the bits of exactly those series and
propagate NaN as the reference does; the healthy series keep status 0 and their oracle
values.  The same checks run on the host build of the lane bodies under the sanitizers
(tests/test_cpu_lanes.py)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_status_bits_fire_on_forced_failures():
    import torch
    from bayesnewton_amd import _native as nat, kernels as K, ops
    from bayesnewton_amd.plan import SeriesPlan
    from oracle import cfast
    rng = np.random.default_rng(5)
    lens = np.array([500, 400, 300, 600])
    off = np.concatenate([[0], np.cumsum(lens)])
    N = int(off[-1])
    dt = np.concatenate([np.concatenate([[0.0], rng.uniform(0.01, 0.2, n - 1)]) for n in lens]).astype(np.float32)
    y = (np.sin(np.arange(N) * 0.05) + 0.3 * rng.standard_normal(N)).astype(np.float32)
    r = rng.uniform(0.01, 0.3, N).astype(np.float32)
    r[off[1] + 123] = -5.0
    y[off[2] + 77] = np.nan
    dev = torch.device("cuda", 0)
    kern = K.Sum([K.Matern52(variance=1.0, lengthscale=1.5), K.Matern32(variance=0.5, lengthscale=3.0)])
    ls, var = kern.hypers(dev)
    for chunk in (16, 4096):
        plan = SeriesPlan(off, chunk_len=chunk, device=dev)
        t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
        dt_t, y_t, r_t = plan.gather(t(dt)), plan.gather(t(y)), plan.gather(t(r), pad=1.0)
        st = torch.full((4,), -1, dtype=torch.int32, device=dev)
        pm, pv, ll = ops.filter_smoother(plan, kern.desc(), ls, var, dt_t, y_t, r_t, None, status=st)
        st = st.cpu().numpy()
        print(f"[status] chunk {chunk}: fused {st}", flush=True)
        assert st[0] == 0 and st[3] == 0
        assert st[1] & nat.BN_STATUS_INNOVATION_NONPD
        assert st[2] & nat.BN_STATUS_NONFINITE
        ll = ll.cpu().numpy()
        assert np.isfinite(ll[[0, 3]]).all() and not np.isfinite(ll[2])
        om, ov, oll = cfast.filter_smoother(["matern52", "matern32"], [1.5, 3.0], [1.0, 0.5], off,
                                            dt.astype(np.float64), y.astype(np.float64), r.astype(np.float64))
        pm = plan.scatter(pm).cpu().numpy().reshape(-1)
        for b in (0, 3):
            s = slice(off[b], off[b + 1])
            assert np.max(np.abs(pm[s] - om[s])) <= 1e-4 * np.max(np.abs(om[s]))
            assert abs(ll[b] - oll[b]) <= 1e-5 * abs(oll[b])
        st2 = torch.full((4,), -1, dtype=torch.int32, device=dev)
        ops.kalman_filter(dt, kern, y[:, None, None], r[:, None, None], plan=plan, status=st2)
        st2 = st2.cpu().numpy()
        assert st2[0] == 0 and st2[3] == 0 and st2[1] & nat.BN_STATUS_INNOVATION_NONPD
        assert st2[2] & nat.BN_STATUS_NONFINITE


def test_model_status_is_zero_on_healthy_batches():
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    rng = np.random.default_rng(6)
    xs = [np.sort(rng.uniform(0, 20, n)) for n in (300, 40, 900)]
    ys = [np.sin(x) + 0.2 * rng.standard_normal(len(x)) for x in xs]
    g = M.MarkovVariationalGP.from_series(K.Matern32(), L.Gaussian(variance=0.1), xs, ys)
    g.inference(lr=1.0)
    assert (g.status == 0).all()
