"""ctypes binding of libbn_hip.so (the C ABI declared in include/bn_hip.h).

The library is the only compute path: if it is missing, every entry point
raises.  There is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libbn_hip.so")

BN_MATERN12, BN_MATERN32, BN_MATERN52 = 1, 2, 3
BN_RETURN_PREDICT = 0x1
BN_RETURN_FULL = 0x2
BN_NOALLOC_SCRATCH_BYTES = 64 * 1024 * 1024
BN_STATUS_INNOVATION_NONPD, BN_STATUS_NONFINITE, BN_STATUS_SMOOTHER_NONPD = 1, 2, 4
BN_SPARSE_UPDATE, BN_SPARSE_ENERGY = 0, 1

vp = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64
u32 = ctypes.c_uint32
f32 = ctypes.c_float
f64 = ctypes.c_double
u8 = ctypes.c_uint8


class KernelDesc(ctypes.Structure):
    _fields_ = [("n_comp", i32), ("kind", i32 * 2)]


class LikDesc(ctypes.Structure):
    _fields_ = [("kind", i32), ("param0", f64), ("param1", f64)]


class StDims(ctypes.Structure):
    _fields_ = [("n_runs", i32), ("n_steps", i32), ("n_loc", i32), ("n_inducing", i32), ("temporal_kind", i32)]


class StHypers(ctypes.Structure):
    _fields_ = [("ls_t", vp), ("var_t", vp), ("var_s", vp), ("stride", i32)]


BN_ST_UPDATE, BN_ST_ENERGY, BN_ST_PROJECT = 0, 1, 2
BN_WS_FEATURES, BN_WS_PARAMS = 8, 34434

BN_LIK_GAUSSIAN, BN_LIK_BERNOULLI_PROBIT, BN_LIK_POISSON, BN_LIK_STUDENTS_T = 1, 2, 3, 4
BN_VI, BN_NEWTON, BN_EP = 0, 1, 2


# name -> (restype, argtypes); every symbol include/bn_hip.h declares
SIGNATURES = {
    "bn_last_error": (ctypes.c_char_p, []),
    "bn_version": (i32, []),
    "bn_plan_create": (i32, [vp, i32, i32, ctypes.POINTER(vp)]),
    "bn_plan_destroy": (i32, [vp]),
    "bn_plan_num_slots": (i64, [vp]),
    "bn_plan_num_chunks": (i32, [vp]),
    "bn_plan_num_tiles": (i32, [vp]),
    "bn_workspace_bytes": (i64, [vp, vp]),
    "bn_plan_gather_f32": (i32, [vp, vp, vp, i32, f32, vp]),
    "bn_plan_scatter_f32": (i32, [vp, vp, vp, i32, vp]),
    "bn_plan_gather_f64": (i32, [vp, vp, vp, i32, f64, vp]),
    "bn_plan_scatter_f64": (i32, [vp, vp, vp, i32, vp]),
    "bn_plan_gather_u8": (i32, [vp, vp, vp, u8, vp]),
    "bn_discretise": (i32, [vp, vp, vp, i64, vp, vp, vp, vp]),
    "bn_filter_smoother": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_filter_smoother_noalloc": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_kalman_filter": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, u32, vp, vp, vp]),
    "bn_filter_loglik": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp]),
    "bn_grad_workspace_bytes": (i64, [vp, vp]),
    "bn_filter_loglik_grad": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_rts_smoother": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, u32, vp, vp]),
    "bn_predict": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp]),
    "bn_lik_predict": (i32, [vp, i32, i64, vp, vp, vp, vp, vp, vp, vp]),
    "bn_site_update": (i32, [vp, vp, vp, i32, i32, f64, f64, i32] + [vp] * 15),
    "bn_energy_terms": (i32, [vp, vp, vp, i32, i32, f64] + [vp] * 10),
    "bn_site_reparametrise": (i32, [vp, vp, vp, vp, vp, vp]),
    "bn_lik_param_grad": (i32, [vp, vp, vp, i32, i32, f64, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_gauss_hermite": (i32, [i32, vp, vp]),
    "bn_st_conditional": (i32, [vp, vp, i32, f64, vp, vp, vp, vp, vp, vp]),
    "bn_st_sites": (i32, [vp, vp, i32, i32, i32, f64, i32] + [vp] * 17),
    "bn_st_ep_sites": (i32, [vp, vp, i32, i32, f64, i32, f64] + [vp] * 18),
    "bn_st_workspace_bytes": (i64, [vp]),
    "bn_st_filter": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_st_smoother": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_st_smoother_full": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_st_predict": (i32, [vp, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_st_sum": (i32, [i32, i32, vp, vp, vp]),
    "bn_st_marginals": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_pairs_filter": (i32, [vp, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_sparse_sites": (i32, [vp, vp, vp, vp, i32, i32, i32, f64, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                              vp, vp, vp, vp, vp]),
    "bn_mf_workspace_bytes": (i64, [vp]),
    "bn_mf_pseudo": (i32, [vp, vp, vp, vp, vp, vp]),
    "bn_mf_filter": (i32, [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_mf_smoother": (i32, [vp, vp, vp, vp, vp, vp, vp, i64, i64, vp, vp, vp, vp]),
    "bn_ih_workspace_bytes": (i64, [vp, i32]),
    "bn_ih_tied": (i32, [vp, vp, vp, vp, vp]),
    "bn_ih_filter": (i32, [vp, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_ih_smoother": (i32, [vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bn_ws_features": (i32, [vp, vp, vp, vp, f64, f64, f64, f64, vp, vp]),
    "bn_mlp_forward": (i32, [i64, vp, vp, vp, vp, vp, vp]),
    "bn_mlp_grad_workspace_bytes": (i64, [i64]),
    "bn_mlp_grad_step": (i32, [i64, vp, vp, vp, vp, vp, vp, vp, i32, f32, vp, vp, vp, vp]),
    "bn_mlp_to_bf16": (i32, [vp, vp, vp]),
    "bn_ws_sites": (i32, [i64, vp, vp, vp, vp, vp, vp]),
    "bn_timing_enable": (i32, [i32]),
    "bn_timing_read": (i32, [i32, ctypes.POINTER(f64), ctypes.POINTER(i32)]),
    "bn_kernel_name": (ctypes.c_char_p, [i32]),
    "bn_num_kernel_ids": (i32, []),
}

_lib = None


class NativeError(RuntimeError):
    pass


def lib():
    """Load libbn_hip.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(
                f"{LIB_PATH} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().bn_last_error().decode(errors="replace")
        raise NativeError(f"libbn_hip error {rc}: {msg}")


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
