"""The C ABI library loads and exports every function include/bn_hip.h declares,
and the ctypes binding covers all of them (no compute: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bn_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(bn_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ["bn_filter_smoother", "bn_kalman_filter", "bn_rts_smoother", "bn_discretise", "bn_plan_create",
                 "bn_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from bayesnewton_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libbn_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from bayesnewton_amd import _native
    assert sorted(_native.SIGNATURES) == declared_functions()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from bayesnewton_amd import _native
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(_native.NativeError):
        _native.lib()


def test_gauss_hermite_golub_welsch_matches_numpy():
    """bn_gauss_hermite (host-side Golub-Welsch in the library) == numpy.hermgauss,
    rescaled as cubature.py:76-84 does (x sqrt(2), w / sqrt(pi))."""
    import numpy as np
    from bayesnewton_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libbn_hip.so not built")
    L = _native.lib()
    for n in [1, 2, 5, 20, 41, 64]:
        x = np.zeros(n)
        w = np.zeros(n)
        _native.check(L.bn_gauss_hermite(n, x.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p)))
        hx, hw = np.polynomial.hermite.hermgauss(n)
        np.testing.assert_allclose(x, np.sqrt(2.0) * hx, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(w, hw / np.sqrt(np.pi), rtol=1e-9, atol=1e-300)
    assert L.bn_gauss_hermite(0, None, None) != 0


def test_space_time_entries_reject_bad_arguments_before_any_device_work():
    """Argument checks of the mean-field, infinite-horizon, prediction and pairs-filter
    entries run on the host and return BN_EINVAL / BN_EUNSUPPORTED with a message (no
    device pointer is touched, so this runs without a GPU)."""
    from bayesnewton_amd import _native as nat
    if not os.path.exists(nat.LIB_PATH):
        pytest.skip("libbn_hip.so not built")
    L = nat.lib()
    ok = nat.StDims(2, 10, 4, 9, 2)
    too_many = nat.StDims(2, 10, 4, 65, 2)
    bad_kind = nat.StDims(2, 10, 4, 9, 3)
    fake = ctypes.c_void_p(16)  # never dereferenced: every call below fails its checks first
    hyp = nat.StHypers(fake, fake, fake, 0)
    r = ctypes.byref
    assert L.bn_mf_filter(r(too_many), r(hyp), 0, fake, fake, fake, fake, fake, fake, fake, None, None) != 0
    assert L.bn_mf_filter(r(bad_kind), r(hyp), 0, fake, fake, fake, fake, fake, fake, fake, None, None) != 0
    assert L.bn_mf_filter(r(ok), r(hyp), 1, fake, fake, fake, fake, fake, fake, fake, None, None) == -1
    assert b"workspace" in L.bn_last_error()
    assert L.bn_mf_filter(r(ok), None, 0, fake, fake, fake, fake, fake, fake, fake, None, None) == -1
    assert L.bn_mf_smoother(r(ok), r(hyp), fake, fake, fake, None, None, 0, 0, None, None, None, None) == -1
    assert L.bn_mf_smoother(r(ok), r(hyp), fake, fake, fake, None, fake, 1, 1, None, None, None, None) == -1
    assert b"overlap" in L.bn_last_error()
    assert L.bn_ih_filter(r(ok), r(hyp), fake, -1, 1, fake, fake, fake, fake, fake, fake, fake, fake, fake, None) == -1
    assert L.bn_ih_filter(r(ok), r(hyp), fake, 20, 1, fake, None, fake, fake, fake, fake, fake, fake, fake, None) == -1
    assert L.bn_ih_smoother(r(ok), r(hyp), fake, -1, fake, fake, fake, fake, None, None, None, None, fake, None) == -1
    assert L.bn_ih_tied(r(ok), fake, fake, None, None) == -1
    assert L.bn_st_predict(r(ok), r(hyp), fake, 0, 5, fake, fake, fake, fake, fake, fake, None, fake, fake, None) == -1
    assert L.bn_st_predict(r(ok), r(hyp), fake, 3, 5, fake, fake, fake, fake, None, fake, None, fake, fake, None) == -1
    assert L.bn_st_smoother_full(r(ok), r(hyp), fake, fake, fake, fake, fake, fake, None, None, fake, None) == -1
    assert b"together" in L.bn_last_error()
    desc = nat.KernelDesc() if hasattr(nat, "KernelDesc") else None
    if desc is not None:
        desc.n_comp, desc.kind[0] = 1, 2
        assert L.bn_pairs_filter(r(desc), fake, fake, 0, 0, fake, fake, fake, fake, fake, fake, fake, None) == -1


def test_plan_create_rejects_bad_offsets_before_any_device_call():
    """bn_plan_create validates the CSR offsets on the host and returns -1 with a
    message before it touches the device (so this runs without a GPU)."""
    import numpy as np
    from bayesnewton_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libbn_hip.so not built")
    L = _native.lib()
    cases = [
        ([0, 0, 0], 2, 16, "no steps"),
        ([0, 5, 3], 2, 16, "non-decreasing"),
        ([1, 5], 1, 16, "offsets[0]"),
        ([0, 5], 0, 16, "n_series"),
        ([0, 5], 1, 0, "chunk_len"),
        ([0, 128 * 128 * 4 + 1], 1, 4, "chunks"),
    ]
    for off, ns, chunk, msg in cases:
        o = np.asarray(off, dtype=np.int64)
        h = ctypes.c_void_p()
        rc = L.bn_plan_create(o.ctypes.data_as(ctypes.c_void_p), ns, chunk, ctypes.byref(h))
        assert rc != 0 and h.value is None, off
        assert msg in L.bn_last_error().decode(), (off, L.bn_last_error())
