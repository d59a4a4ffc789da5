"""The oracle reproduces the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py from seeded inputs): a regression pin on the oracle, so
that the fixtures the GPU tests read stay the oracle's outputs."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _load(name):
    z = np.load(os.path.join(GOLDEN, name))
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["kf_rts_sum_m32_m52_n400.npz", "vi_probit_n300.npz",
                                  "spacetime_poisson_vi_g3_t24.npz", "pairs_sum_m52_m32_n60.npz",
                                  "predict_studentt_vi_n300.npz"])
def test_oracle_reproduces_golden(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    make_inputs, make_outputs = mg.FIXTURES[name]
    z = _load(name)
    d = make_inputs()
    for k, v in d.items():
        np.testing.assert_array_equal(z[k], v, err_msg=k)
    res = make_outputs(d)
    for k, v in res.items():
        np.testing.assert_allclose(z[k], v, rtol=1e-10, atol=1e-12, err_msg=k)


def test_golden_kf_rts_invariants():
    """Properties the stored filter/smoother outputs must have whatever produced them:
    smoothed variances below the prior variance (1.5) and positive, filtered
    covariances symmetric positive definite, the log-likelihood finite and negative
    of order n."""
    z = _load("kf_rts_sum_m32_m52_n400.npz")
    sv = z["smooth_var"].reshape(-1)
    assert np.all(sv > 0) and np.all(sv < 1.5)
    P = z["filter_cov"]
    assert np.allclose(P, np.transpose(P, (0, 2, 1)), atol=1e-12)
    assert np.all(np.linalg.eigvalsh(P) > -1e-12)
    assert np.isfinite(z["loglik"]) and abs(float(z["loglik"])) < 10 * len(z["dt"])


@pytest.mark.parametrize("name", ["spacetime_cox_full_seed1005.npz", "spacetime_cox_diverging_seeds.npz"])
def test_large_fixtures_hold_their_generator_inputs(name):
    """The full-size space-time fixtures take minutes of oracle time, so only their
    inputs (seeds, sizes, step sizes) are checked here against the generator; their
    outputs are pinned by the GPU parity tests that read them."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    make_inputs, _ = mg.LARGE_FIXTURES[name]
    z = _load(name)
    for k, v in make_inputs().items():
        np.testing.assert_array_equal(z[k], v, err_msg=k)
    for k in z:
        assert np.all(np.isfinite(z[k]) | (z[k].dtype == bool)), k
