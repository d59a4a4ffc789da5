"""GPU parity of the filter over pairs of states (ops.py:383-426, the sparse Markov
model's filter) against the FP64 oracle (itself pinned against dense Gaussian
conditioning in tests/test_oracle_pairs.py).  FP64 throughout: 1e-9 relative on the log
likelihood, 1e-8 on the moments."""
import numpy as np
import pytest

from oracle import kernels as OK, ops as OP

pytestmark = pytest.mark.gpu


def _kernels(name):
    from bayesnewton_amd import kernels as K
    if name == "m12":
        return K.Matern12(variance=0.8, lengthscale=1.0), OK.Matern12(0.8, 1.0)
    if name == "m32":
        return K.Matern32(variance=1.3, lengthscale=2.0), OK.Matern32(1.3, 2.0)
    if name == "m52":
        return K.Matern52(variance=1.1, lengthscale=3.0), OK.Matern52(1.1, 3.0)
    if name == "sum52_32":
        return (K.Sum([K.Matern52(variance=1.0, lengthscale=3.0), K.Matern32(variance=0.5, lengthscale=0.7)]),
                OK.Sum([OK.Matern52(1.0, 3.0), OK.Matern32(0.5, 0.7)]))
    return (K.Sum([K.Matern52(variance=1.0, lengthscale=3.0), K.Matern52(variance=0.4, lengthscale=0.9)]),
            OK.Sum([OK.Matern52(1.0, 3.0), OK.Matern52(0.4, 0.9)]))


@pytest.mark.parametrize("name", ["m12", "m32", "m52", "sum52_32", "sum52_52"])
@pytest.mark.parametrize("parallel", [False, True])
def test_pairs_filter_matches_oracle(name, parallel):
    from bayesnewton_amd import ops
    gk, ok = _kernels(name)
    d = ok.stationary_covariance().shape[0]
    rng = np.random.default_rng(3)
    N = 60
    dt = np.concatenate([[2e10], rng.uniform(0.2, 1.5, N - 2), [2e10]])
    y = rng.standard_normal((N, 2 * d, 1))
    G = rng.standard_normal((N, 2 * d, 2 * d)) * 0.3
    R = G @ np.swapaxes(G, 1, 2) + 0.2 * np.eye(2 * d)
    print(f"[pairs] kernel={name} parallel={parallel} N={N} d={d}", flush=True)
    ell, (fm, fP) = ops.kalman_filter_pairs(dt, gk, y, R, parallel=parallel)
    oell, (ofm, ofP) = OP.kalman_filter_pairs(dt, ok, y, R, parallel=False)
    assert abs(float(ell) - oell) <= 1e-9 * abs(oell), (float(ell), oell)
    fm, fP = fm.cpu().numpy(), fP.cpu().numpy()
    assert fm.shape == ofm.shape and fP.shape == ofP.shape
    assert np.max(np.abs(fm - ofm)) <= 1e-8 * np.max(np.abs(ofm))
    assert np.max(np.abs(fP - ofP)) <= 1e-8 * np.max(np.abs(ofP))
