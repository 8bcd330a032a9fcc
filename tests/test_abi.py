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
    # two padded N x N factors dominate: 2 * 65536^2 * 8 bytes, half for mixed
    assert 2 * 65536 ** 2 * 8 <= b64 < 2.4 * 65536 ** 2 * 8
    assert 2 * 65536 ** 2 * 4 <= b32 < 2.6 * 65536 ** 2 * 4
    assert lib.gpk_mll_workspace_bytes(0, 1, 0) == 0
    assert lib.gpk_predict_workspace_bytes(2000, 1, 500, 0, 1) > lib.gpk_mll_workspace_bytes(2000, 1, 0)


@pytest.mark.parametrize("kind,prec,n,d,nls,code", [
    (9, 0, 10, 2, 1, 1),     # unknown kernel
    (0, 7, 10, 2, 1, 1),     # unknown precision
    (0, 0, 0, 2, 1, 1),      # n < 1
    (0, 0, 10, 99, 1, 1),    # d above the fused-kernel limit
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


def test_mixed_arith_option_validation():
    lib = _lib.lib()
    assert lib.gpk_set_mixed_arith(4, 3, 6) == 1
    assert lib.gpk_set_mixed_arith(6, 3, 6) == 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libgpk.so")
    with pytest.raises(RuntimeError, match="no host fallback"):
        _lib.lib()
