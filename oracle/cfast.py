"""ctypes binding of the oracle's FP64 C filter/smoother (oracle/kfrts.c).

Test infrastructure only: used by tests/ as the checker for long series and by
bench.py's cpu_baseline leg.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libkfrts.so")
_libs = {}

KIND = {"matern12": 1, "matern32": 2, "matern52": 3}
NATIVE_CFLAGS = ["-O3", "-march=native", "-fPIC", "-fopenmp", "-std=c99"]


def load(path=None):
    path = _LIB if path is None else path
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run `make -C oracle`)")
        lib = ctypes.CDLL(path)
        lib.oracle_filter_smoother.restype = ctypes.c_int
        _libs[path] = lib
    return _libs[path]


def build_native(outdir):
    """Compile kfrts.c for the host it runs on (-march=native) into outdir, for timing
    the CPU baseline on the machine that measures the GPU.  Returns (library path,
    provenance dict: compiler, flags, CPU model, logical CPUs)."""
    import platform
    import subprocess
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libkfrts_native.so")
    cc = os.environ.get("CC", "gcc")
    subprocess.run([cc] + NATIVE_CFLAGS + ["-shared", "-o", out, os.path.join(_HERE, "kfrts.c"), "-lm"], check=True)
    ver = subprocess.run([cc, "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return out, {"compiler": ver, "flags": " ".join(NATIVE_CFLAGS), "cpu_model": model,
                 "logical_cpus": os.cpu_count()}


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def filter_smoother(kinds, ls, var, offsets, dt, y, r, mask=None, nthreads=1, lib_path=None):
    """Batched FP64 sequential filter + RTS smoother (ops.py:154-180, 288-311).

    kinds: list of component names; ls/var: [n_comp] shared or [B, n_comp] per series.
    lib_path: another build of kfrts.c (build_native), default the in-tree one.
    Returns (post_mean [N], post_var [N], loglik [B]).
    """
    lib = load(lib_path)
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
