"""MI355X-native Markov-GP inference hot path (drop-in for BayesNewton's ops /
Markov models).  Compute runs in libbn_hip.so (HIP, gfx950); this package is the
host-side mirror of the reference interface."""
from . import _native, kernels, ops, plan  # noqa: F401

__version__ = "0.1.0"
