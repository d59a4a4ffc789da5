"""The temporal model's posterior / log-likelihood cache key (host logic only).

A stored posterior is reused only while neither the kernel hypers nor the sites
changed.  The key must differ after every hyper change even when the allocator
hands the new hyper tensors the addresses (and Python ids) of the old ones.
"""
import numpy as np

from bayesnewton_amd import kernels as K
from bayesnewton_amd import models as M


def _bare_model(kernel):
    g = object.__new__(M.MarkovVariationalGP)
    g.kernel = kernel
    g.device = "cpu"
    g._per_series = False
    g._hyper_key = None
    return g


def test_state_key_changes_with_every_hyper_change():
    kern = K.Sum([K.Matern52(variance=1.0, lengthscale=1.0), K.Matern32(variance=0.5, lengthscale=2.0)])
    g = _bare_model(kern)
    seen = {g._state_key()}
    rng = np.random.default_rng(0)
    for i in range(200):
        c = kern.components[i % 2]
        c.lengthscale = float(rng.uniform(0.5, 5.0))
        if i % 3 == 0:
            c.variance = float(rng.uniform(0.1, 2.0))
        # two changes with no refresh in between, as hyper-only Adam steps make them
        k1 = g._state_key()
        kern.components[0].variance = float(rng.uniform(0.1, 2.0))
        k2 = g._state_key()
        assert k1 not in seen and k2 not in seen and k1 != k2, i
        seen.update((k1, k2))


def test_state_key_stable_without_changes_and_tracks_sites():
    g = _bare_model(K.Matern52(variance=1.0, lengthscale=1.0))
    k0 = g._state_key()
    assert g._state_key() == k0
    g._sites_version += 1
    k1 = g._state_key()
    assert k1 != k0 and g._state_key() == k1
    # setting a hyper to its current value is not a change
    g.kernel.components[0].lengthscale = 1.0
    assert g._state_key() == k1
