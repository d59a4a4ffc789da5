"""ctypes binding of the oracle's FP64 C filter/smoother (oracle/kfrts.c).

Test infrastructure only: used by tests/ as the checker for long series and by
bench.py's cpu_baseline leg.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libkfrts.so")
_lib = None

KIND = {"matern12": 1, "matern32": 2, "matern52": 3}


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle/libkfrts.so is not built (run `make -C oracle`)")
        _lib = ctypes.CDLL(_LIB)
        _lib.oracle_filter_smoother.restype = ctypes.c_int
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def filter_smoother(kinds, ls, var, offsets, dt, y, r, mask=None, nthreads=1):
    """Batched FP64 sequential filter + RTS smoother (ops.py:154-180, 288-311).

    kinds: list of component names; ls/var: [n_comp] shared or [B, n_comp] per series.
    Returns (post_mean [N], post_var [N], loglik [B]).
    """
    lib = load()
    k = np.array([KIND[s] for s in kinds], dtype=np.int32)
    ls = np.ascontiguousarray(ls, dtype=np.float64)
    var = np.ascontiguousarray(var, dtype=np.float64)
    stride = 0 if ls.ndim == 1 else len(kinds)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    B = len(offsets) - 1
    dt = np.ascontiguousarray(dt, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    N = dt.shape[0]
    pm = np.empty(N)
    pv = np.empty(N)
    ll = np.empty(B)
    mk = None
    if mask is not None:
        mk = np.ascontiguousarray(mask, dtype=np.uint8)
    rc = lib.oracle_filter_smoother(
        ctypes.c_int(len(kinds)), _p(k, ctypes.c_int), _p(ls, ctypes.c_double), _p(var, ctypes.c_double),
        ctypes.c_int(stride), _p(offsets, ctypes.c_int64), ctypes.c_int(B), _p(dt, ctypes.c_double),
        _p(y, ctypes.c_double), _p(r, ctypes.c_double),
        _p(mk, ctypes.c_uint8) if mk is not None else None,
        _p(pm, ctypes.c_double), _p(pv, ctypes.c_double), _p(ll, ctypes.c_double), ctypes.c_int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_filter_smoother failed ({rc})")
    return pm, pv, ll
