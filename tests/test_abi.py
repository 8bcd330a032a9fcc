"""The C ABI library loads, exports every symbol include/gpk.h declares, and
validates arguments on the host -- no GPU needed (no compute calls)."""

import ctypes
import re
from pathlib import Path

import pytest

from gpjax_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "gpk.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpk_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_functions()
    for must in ("gpk_gram", "gpk_cross_covariance", "gpk_mll_forward", "gpk_mll_backward",
                 "gpk_mll_workspace_bytes", "gpk_predict", "gpk_last_error", "gpk_version"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_and_workspace_sizes():
    lib = _lib.lib()
    assert b"gfx950" in lib.gpk_version()
    b64 = lib.gpk_mll_workspace_bytes(65536, 16, 0)
    b32 = lib.gpk_mll_workspace_bytes(65536, 16, 1)
    # two padded N x N factors dominate: 2 * 65536^2 * 8 bytes; mixed: two float32
    # factors plus Sigma's lower triangle in float64 (one more N^2 * 4 bytes)
    assert 2 * 65536 ** 2 * 8 <= b64 < 2.4 * 65536 ** 2 * 8
    assert 3 * 65536 ** 2 * 4 <= b32 < 3.6 * 65536 ** 2 * 4
    assert lib.gpk_mll_workspace_bytes(0, 1, 0) == 0
    assert lib.gpk_predict_workspace_bytes(2000, 1, 500, 0, 1) > lib.gpk_mll_workspace_bytes(2000, 1, 0)


@pytest.mark.parametrize("kind,prec,n,d,nls,code", [
    (9, 0, 10, 2, 1, 1),     # unknown kernel
    (0, 7, 10, 2, 1, 1),     # unknown precision
    (0, 0, 0, 2, 1, 1),      # n < 1
    (0, 0, 10, 99, 1, 1),    # d above the kernels' limit (64)
    (0, 0, 10, 3, 2, 1),     # nls neither 1 nor d
])
def test_argument_validation_without_device(kind, prec, n, d, nls, code):
    lib = _lib.lib()
    dummy = ctypes.c_void_p(0x1000)
    rc = lib.gpk_mll_forward(kind, prec, dummy, dummy, n, d, dummy, nls, 1e-6, -1, dummy,
                             1 << 40, dummy, None)
    assert rc == code
    assert lib.gpk_last_error().decode()


def test_workspace_too_small_is_reported():
    lib = _lib.lib()
    dummy = ctypes.c_void_p(0x1000)
    rc = lib.gpk_mll_forward(0, 0, dummy, dummy, 100, 2, dummy, 1, 1e-6, -1, dummy, 16, dummy,
                             None)
    assert rc == 5
    assert "workspace" in lib.gpk_last_error().decode()


def test_loocv_workspace_and_validation():
    lib = _lib.lib()
    for prec in (0, 1):
        base = lib.gpk_mll_workspace_bytes(3000, 3, prec)
        loo = lib.gpk_loocv_workspace_bytes(3000, 3, prec)
        # five padded vectors and the result slots on top of the exact-GP workspace
        assert base + 5 * 3072 * 8 <= loo <= base + 5 * 3072 * 8 + 6 * 256
    assert lib.gpk_loocv_workspace_bytes(0, 3, 0) == 0
    dummy = ctypes.c_void_p(0x1000)
    rc = lib.gpk_loocv_forward(0, 0, dummy, dummy, 100, 2, dummy, 1, 1e-6, -1, dummy, 16, dummy,
                               None)
    assert rc == 5 and "workspace" in lib.gpk_last_error().decode()
    rc = lib.gpk_loocv_backward(0, 0, dummy, 100, 2, dummy, 3, 1.0, dummy, 1 << 40, dummy, None)
    assert rc == 1 and "nls" in lib.gpk_last_error().decode()


def test_elbo_predict_workspace_and_validation():
    lib = _lib.lib()
    base = lib.gpk_elbo_workspace_bytes(5000, 300, 4, 0)
    full = lib.gpk_elbo_predict_workspace_bytes(5000, 300, 4, 1000, 1)
    diag = lib.gpk_elbo_predict_workspace_bytes(5000, 300, 4, 1000, 0)
    # Kzt and the two solves (3 x 512 x 1024 doubles), plus 1024^2 doubles for Ktt
    assert base + 3 * 512 * 1024 * 8 <= diag < full
    assert full - diag >= 1024 * 1024 * 8 - 256
    assert lib.gpk_elbo_predict_workspace_bytes(5000, 300, 4, 0, 1) == 0
    dummy = ctypes.c_void_p(0x1000)
    rc = lib.gpk_elbo_predict(0, 0, dummy, dummy, 100, dummy, 10, 2, dummy, 5, dummy, 1, 1e-6,
                              1, dummy, 16, dummy, dummy, None)
    assert rc == 5 and "workspace" in lib.gpk_last_error().decode()
    rc = lib.gpk_elbo_predict(0, 1, dummy, dummy, 100, dummy, 10, 2, dummy, 5, dummy, 1, 1e-6,
                              1, dummy, 1 << 40, dummy, dummy, None)
    assert rc == 1   # the collapsed path is float64 only
    rc = lib.gpk_elbo_predict(0, 0, dummy, dummy, 100, dummy, 10, 2, None, 5, dummy, 1, 1e-6,
                              1, dummy, 1 << 40, dummy, dummy, None)
    assert rc == 1 and "xt" in lib.gpk_last_error().decode()


def test_elbo_precision_codes():
    """gpk_elbo_forward takes GPK_PREC_FP64, GPK_PREC_MIXED and
    GPK_PREC_MIXED_ARITH(p, t, l) with p, t in {0, 3, 4, 6} (checked before any
    device work: a valid code gets as far as the workspace check), rejects every
    other code, and needs the same workspace for all of them."""
    lib = _lib.lib()
    dummy = ctypes.c_void_p(0x1000)
    size = lib.gpk_elbo_workspace_bytes(5000, 300, 4, 0)
    good = [0, 1] + [_lib.mixed_arith(p, t, 6) for p in (0, 3, 4, 6) for t in (0, 3, 4, 6)]
    for code in good:
        assert lib.gpk_elbo_workspace_bytes(5000, 300, 4, code) == size
        rc = lib.gpk_elbo_forward(0, code, dummy, dummy, 100, dummy, 10, 2, dummy, 1, 1e-6,
                                  dummy, 16, dummy, None)
        assert rc == 5 and "workspace" in lib.gpk_last_error().decode(), code
    bad = [2, 0 | _lib.ARITH_FLAG, 1 | (1 << 9),
           1 | _lib.ARITH_FLAG | (5 << 12) | (6 << 16) | (6 << 20),
           1 | _lib.ARITH_FLAG | (6 << 12) | (7 << 16) | (6 << 20)]
    for code in bad:
        rc = lib.gpk_elbo_forward(0, code, dummy, dummy, 100, dummy, 10, 2, dummy, 1, 1e-6,
                                  dummy, 1 << 40, dummy, None)
        assert rc == 1 and "prec" in lib.gpk_last_error().decode(), code


def test_mixed_arith_is_a_per_call_precision_code():
    """The stage arithmetic travels in the call's `prec` argument
    (GPK_PREC_MIXED_ARITH); there is no process-wide setting."""
    lib = _lib.lib()
    assert not hasattr(lib, "gpk_set_mixed_arith")
    plain = lib.gpk_mll_workspace_bytes(3000, 3, 1)
    for stages in ((6, 3, 6), (4, 3, 4), (0, 3, 0), (6, 6, 6)):
        code = _lib.mixed_arith(*stages)
        assert code & 0xFF == 1 and code & _lib.ARITH_FLAG
        assert lib.gpk_mll_workspace_bytes(3000, 3, code) == plain
    with pytest.raises(ValueError):
        _lib.mixed_arith(5, 3, 6)
    bad = 1 | _lib.ARITH_FLAG | (5 << 12) | (3 << 16) | (6 << 20)   # 5 products: no such split
    for code in (bad, 0 | _lib.ARITH_FLAG, 1 | (1 << 9), 2):
        assert lib.gpk_mll_workspace_bytes(3000, 3, code) == 0
        dummy = ctypes.c_void_p(0x1000)
        rc = lib.gpk_mll_forward(0, code, dummy, dummy, 100, 2, dummy, 1, 1e-6, -1, dummy,
                                 1 << 40, dummy, None)
        assert rc == 1 and "precision" in lib.gpk_last_error().decode()


def test_backward_entry_points_check_workspace_alignment():
    lib = _lib.lib()
    odd = ctypes.c_void_p(0x1008)
    dummy = ctypes.c_void_p(0x1000)
    rc = lib.gpk_mll_backward(0, 0, dummy, 100, 2, dummy, 1, 1.0, odd, 1 << 40, dummy, None)
    assert rc == 1 and "aligned" in lib.gpk_last_error().decode()
    rc = lib.gpk_loocv_backward(0, 1, dummy, 100, 2, dummy, 1, 1.0, odd, 1 << 40, dummy, None)
    assert rc == 1 and "aligned" in lib.gpk_last_error().decode()
    rc = lib.gpk_elbo_backward(0, 0, dummy, dummy, 100, dummy, 10, 2, dummy, 1, 1.0, odd,
                               1 << 40, dummy, dummy, None)
    assert rc == 1 and "aligned" in lib.gpk_last_error().decode()


def test_adam_step_validation():
    lib = _lib.lib()
    d = ctypes.c_void_p(0x1000)
    assert lib.gpk_adam_step(0, d, d, d, d, d, d, 0.1, 0.9, 0.999, 1e-8, d, d, d, 1, None) == 1
    assert lib.gpk_adam_step(3, d, d, d, d, d, d, 0.0, 0.9, 0.999, 1e-8, d, d, d, 1, None) == 1
    assert lib.gpk_adam_step(3, d, d, d, d, d, d, 0.1, 1.0, 0.999, 1e-8, d, d, d, 1, None) == 1
    assert lib.gpk_adam_step(3, None, d, d, d, d, d, 0.1, 0.9, 0.999, 1e-8, d, d, d, 1, None) == 1


def _offsets(*rows):
    arr = (ctypes.c_int64 * (len(rows) + 1))()
    acc = 0
    arr[0] = 0
    for i, r in enumerate(rows):
        acc += r
        arr[i + 1] = acc
    return arr


def test_batched_workspace_groups_by_padded_size():
    lib = _lib.lib()
    d = 4
    off = _offsets(10, 290, 10, 256)           # padded sizes 128, 384, 128, 256
    meta = 256 + 256                            # member map and row offsets, 256-byte aligned
    # each member: the single-problem layout, plus (FP64) room for Sigma's packed
    # float64 lower tiles and their matvec partials, which the mixed layout has
    for prec in (0, 1):
        got = lib.gpk_batched_workspace_bytes(off, 4, d, prec)
        per = {}
        for np_ in (128, 256, 384):
            t = np_ // 128
            ntile = t * (t + 1) // 2
            extra = (ntile * 128 * 128 * 8 + 2 * ntile * 128 * 8) if prec == 0 else 0
            per[np_] = lib.gpk_mll_workspace_bytes(np_, d, prec) + extra
        assert got == 2 * per[128] + per[256] + per[384] + meta
    # the fp16 split is refused in a batch (its scales come from one problem)
    assert lib.gpk_batched_workspace_bytes(off, 4, d, _lib.mixed_arith(6, 3, 6)) > 0
    assert lib.gpk_batched_workspace_bytes(off, 4, d, _lib.mixed_arith(4, 3, 6)) == 0
    assert "fp16" in lib.gpk_last_error().decode()


@pytest.mark.parametrize("rows,count,d", [
    ((5, 0, 7), 3, 2),       # a problem without rows
    ((5, 7), 0, 2),          # no problems
    ((5, 7), 2, 0),          # d < 1
    ((5, 7), 2, 99),         # d above the kernels' limit (64)
])
def test_batched_validation_without_device(rows, count, d):
    lib = _lib.lib()
    off = _offsets(*rows)
    assert lib.gpk_batched_workspace_bytes(off, count, d, 0) == 0
    dummy = ctypes.c_void_p(0x1000)
    rc = lib.gpk_batched_mll_grad(0, 0, dummy, dummy, off, count, d, dummy, 1, 1e-6, -1, dummy,
                                  1 << 40, dummy, dummy, None)
    assert rc == 1
    assert "batched" in lib.gpk_last_error().decode()


def test_batched_workspace_and_offsets_checks():
    lib = _lib.lib()
    dummy = ctypes.c_void_p(0x1000)
    off = _offsets(5, 7)
    rc = lib.gpk_batched_plan(off, 2, 3, 0, dummy, 64)
    assert rc == 5 and "workspace" in lib.gpk_last_error().decode()
    rc = lib.gpk_batched_mll_grad(0, 0, dummy, dummy, off, 2, 3, dummy, 1, 1e-6, -1, dummy, 64,
                                  dummy, dummy, None)
    assert rc == 5
    bad = _offsets(5, 7)
    bad[0] = 1
    assert lib.gpk_batched_plan(bad, 2, 3, 0, dummy, 1 << 40) == 1
    assert lib.gpk_batched_plan(off, 2, 3, 99, dummy, 1 << 40) == 1
    # a workspace that no plan has written is refused before any launch
    rc = lib.gpk_batched_mll_grad(0, 0, dummy, dummy, off, 2, 3, dummy, 1, 1e-6, -1,
                                  ctypes.c_void_p(0x7f0000), 1 << 40, dummy, dummy, None)
    assert rc == 1 and "plan" in lib.gpk_last_error().decode()
    # the autotuner runs the same checks before it launches anything
    rc = lib.gpk_batched_tune(0, 0, dummy, dummy, off, 2, 3, dummy, 1, 1e-6, -1,
                              ctypes.c_void_p(0x7f0000), 1 << 40, dummy, dummy, None, 0, None)
    assert rc == 1 and "batched_tune" in lib.gpk_last_error().decode()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libgpk.so")
    with pytest.raises(RuntimeError, match="no host fallback"):
        _lib.lib()


def test_linalg_entry_points_validate_on_the_host():
    lib = _lib.lib()
    d = ctypes.c_void_p(0x1000)
    assert lib.gpk_linalg_workspace_bytes(0, 0, 0) == 0
    assert lib.gpk_linalg_workspace_bytes(300, 0, 7) == 0
    w64 = lib.gpk_linalg_workspace_bytes(300, 0, 0)
    w32 = lib.gpk_linalg_workspace_bytes(300, 0, 1)
    assert 2 * 384 * 384 * 8 <= w64 and 2 * 384 * 384 * 4 <= w32 < w64
    # right-hand sides add two padded n x m buffers
    assert lib.gpk_linalg_workspace_bytes(300, 5, 0) >= w64 + 2 * 384 * 128 * 8
    assert lib.gpk_potrf(0, d, 10, 5, d, 10, None, None, d, 1 << 40, None) == 1   # lda < n
    assert lib.gpk_potrf(3, d, 10, 10, d, 10, None, None, d, 1 << 40, None) == 1  # precision
    assert lib.gpk_potrf(0, d, 10, 10, d, 10, None, None, d, 16, None) == 5       # workspace
    assert lib.gpk_potrf(0, d, 10, 10, d, 10, None, None, ctypes.c_void_p(0x1008), 1 << 40,
                         None) == 1                                               # alignment
    assert lib.gpk_trsm(d, 10, 10, d, 0, 1, d, 1 << 40, None) == 1               # m < 1
    assert lib.gpk_potri(d, 10, 10, None, 10, d, 1 << 40, None) == 1
    assert lib.gpk_logdet_tri(d, 0, 1, d, None) == 1
    assert lib.gpk_mvn_log_prob(d, d, 10, 10, 0, d, d, d, 16, None) == 5
    assert lib.gpk_mvn_workspace_bytes(10, 1) == 256


def test_noalloc_limits_checked_on_the_host():
    """gpk_mll_forward_noalloc refuses n above its fixed scratch (1,024 rows)
    before touching the device; the FP64 layout at its limit fits the 20 MiB."""
    lib = _lib.lib()
    d = ctypes.c_void_p(0x1000)
    rc = lib.gpk_mll_forward_noalloc(0, 0, d, d, 1025, 4, d, 1, 1e-6, -1, d, None, None)
    assert rc == 1 and "1024" in lib.gpk_last_error().decode()
    assert lib.gpk_mll_workspace_bytes(1024, 32, 0) <= 20 << 20
