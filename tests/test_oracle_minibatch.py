"""Pins the oracle's mini-batched inference and energy (inference.py:65-90 with
group_natural_params, basemodels.py:195-199; energies inference.py:130-325 with
scale = N / B).

* A batch holding every index (in any order) is the full step and the full energy.
* Site map: after a batch step the batch's sites equal those a full step gives at the
  same points, and every other site keeps its value; diff1 / diff2 are the batch's
  absolute changes summed and divided by N.
* Energy: over the two halves of a partition into equal batches, the mean of the
  batch energies is the full energy (the data term is scaled by N / B, the KL / pseudo
  terms are not).  A scale on the wrong term, or none, fails it.
"""
import numpy as np
import pytest

from oracle import kernels as OK, likelihoods as OL, models as OM


def _model(kind, seed=3, n=60):
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(0, 30, n))
    f = np.sin(0.3 * x)
    if kind == "EP":
        y, lik = (rng.uniform(size=n) < 0.5 * (1 + np.tanh(f))).astype(float), OL.Bernoulli()
    else:
        y, lik = f + 0.3 * rng.standard_t(3.0, size=n), OL.StudentsT(0.3, 3.0)
    y[[5, 22]] = np.nan
    return OM.MarkovGP(OK.Matern52(1.1, 4.0), lik, x[:, None], y[:, None], inference=kind, power=0.7)


@pytest.mark.parametrize("kind", ["VI", "Newton", "EP"])
def test_full_batch_equals_full_step(kind):
    a, b = _model(kind), _model(kind)
    perm = np.random.default_rng(0).permutation(a.num_data)
    for lr in (1.0, 0.4):
        _, (d1a, d2a) = a.inference(lr=lr)
        _, (d1b, d2b) = b.inference(lr=lr, batch_ind=perm)
        np.testing.assert_array_equal(a.nat1, b.nat1)
        np.testing.assert_array_equal(a.nat2, b.nat2)
        assert d1a == d1b and d2a == d2b
    # the permuted batch sums the data term in another order
    assert abs(a.energy() - b.energy(batch_ind=perm)) <= 1e-14 * abs(a.energy())


@pytest.mark.parametrize("kind", ["VI", "Newton", "EP"])
def test_batch_step_touches_only_the_batch(kind):
    a, b = _model(kind), _model(kind)
    a.inference(lr=1.0)
    b.inference(lr=1.0)
    batch = np.random.default_rng(1).choice(a.num_data, 17, replace=False)
    n1_old, n2_old = b.nat1.copy(), b.nat2.copy()
    a.inference(lr=1.0)
    _, (d1, d2) = b.inference(lr=1.0, batch_ind=batch)
    np.testing.assert_allclose(b.nat1[batch], a.nat1[batch], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(b.nat2[batch], a.nat2[batch], rtol=1e-13, atol=1e-13)
    rest = np.setdiff1d(np.arange(a.num_data), batch)
    np.testing.assert_array_equal(b.nat1[rest], n1_old[rest])
    np.testing.assert_array_equal(b.nat2[rest], n2_old[rest])
    assert abs(d1 - np.sum(np.abs(b.nat1 - n1_old)) / a.num_data) < 1e-14 * max(1.0, d1)
    assert abs(d2 - np.sum(np.abs(b.nat2 - n2_old)) / a.num_data) < 1e-14 * max(1.0, d2)


@pytest.mark.parametrize("kind", ["VI", "Newton", "EP"])
def test_mean_of_half_batch_energies_is_full_energy(kind):
    m = _model(kind)
    m.inference(lr=1.0)
    m.inference(lr=0.5)
    perm = np.random.default_rng(2).permutation(m.num_data)
    h = m.num_data // 2
    e_half = 0.5 * (m.energy(batch_ind=perm[:h]) + m.energy(batch_ind=perm[h:]))
    e = m.energy()
    assert abs(e_half - e) <= 1e-12 * abs(e)
    assert abs(m.energy(batch_ind=perm[:h]) - e) > 1e-6 * abs(e)  # the halves differ
