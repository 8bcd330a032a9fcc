"""ctypes binding of libgpk's C ABI (``include/gpk.h``).

The library is required: importing the compute path raises if ``libgpk.so``
has not been built, and there is no host fallback.  Device memory and streams
come from PyTorch (``tensor.data_ptr()``, ``torch.cuda.current_stream()``).
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).with_name("libgpk.so")

GPK_OK = 0
STATUS = {0: "OK", 1: "BAD_ARG", 3: "NO_DEVICE", 4: "HIP", 5: "WORKSPACE"}
KERNEL_IDS = {"rbf": 0, "matern12": 1, "matern32": 2, "matern52": 3}
PRECISIONS = {"fp64": 0, "mixed": 1}
ARITH_FLAG = 0x100


def mixed_arith(potrf: int, trtri: int, lauum: int) -> int:
    """The ``prec`` code of the mixed path with the products of each stage named
    (GPK_PREC_MIXED_ARITH in include/gpk.h): 6, 3, 4 (fp16 split) or 0 (float64
    products of the float32 operands)."""
    for v in (potrf, trtri, lauum):
        if v not in (0, 3, 4, 6):
            raise ValueError(f"stage products must be 0, 3, 4 or 6, got {v}")
    return PRECISIONS["mixed"] | ARITH_FLAG | (potrf << 12) | (trtri << 16) | (lauum << 20)


RESULT_LEN = 10
RES_MLL, RES_LOGDET, RES_QUAD, RES_RESID, RES_INFO = 0, 1, 2, 3, 4
RES_PROBE, RES_TRACE, RES_FLAG, RES_SINV, RES_MLL_ERR = 5, 6, 7, 8, 9

_c = ctypes
_P = _c.c_void_p
_I = _c.c_int
_L = _c.c_int64
_D = _c.c_double
_S = _c.c_size_t

# name -> (restype, argtypes); mirrors include/gpk.h
SIGNATURES = {
    "gpk_version": (_c.c_char_p, []),
    "gpk_last_error": (_c.c_char_p, []),
    "gpk_cross_covariance": (_I, [_I, _P, _L, _P, _L, _L, _P, _I, _P, _L, _P]),
    "gpk_gram": (_I, [_I, _P, _L, _L, _P, _I, _D, _I, _P, _L, _P]),
    "gpk_mll_workspace_bytes": (_S, [_L, _L, _I]),
    "gpk_mll_forward": (_I, [_I, _I, _P, _P, _L, _L, _P, _I, _D, _I, _P, _S, _P, _P]),
    "gpk_mll_backward": (_I, [_I, _I, _P, _L, _L, _P, _I, _D, _P, _S, _P, _P]),
    "gpk_mll_forward_noalloc": (_I, [_I, _I, _P, _P, _L, _L, _P, _I, _D, _I, _P, _P, _P]),
    "gpk_mfma_selftest": (_I, [_P, _P, _P]),
    "gpk_elbo_workspace_bytes": (_S, [_L, _L, _L, _I]),
    "gpk_elbo_forward": (_I, [_I, _I, _P, _P, _L, _P, _L, _L, _P, _I, _D, _P, _S, _P, _P]),
    "gpk_elbo_backward": (_I, [_I, _I, _P, _P, _L, _P, _L, _L, _P, _I, _D, _P, _S, _P, _P, _P]),
    "gpk_loocv_workspace_bytes": (_S, [_L, _L, _I]),
    "gpk_loocv_forward": (_I, [_I, _I, _P, _P, _L, _L, _P, _I, _D, _I, _P, _S, _P, _P]),
    "gpk_loocv_backward": (_I, [_I, _I, _P, _L, _L, _P, _I, _D, _P, _S, _P, _P]),
    "gpk_elbo_predict_workspace_bytes": (_S, [_L, _L, _L, _L, _I]),
    "gpk_elbo_predict": (_I, [_I, _I, _P, _P, _L, _P, _L, _L, _P, _L, _P, _I, _D, _I, _P, _S,
                              _P, _P, _P]),
    "gpk_predict_workspace_bytes": (_S, [_L, _L, _L, _I, _I]),
    "gpk_predict": (_I, [_I, _I, _P, _P, _L, _L, _P, _L, _P, _I, _D, _D, _I, _I, _P, _S, _P, _P,
                         _P, _P]),
    "gpk_batched_workspace_bytes": (_S, [_P, _L, _L, _I]),
    "gpk_batched_plan": (_I, [_P, _L, _L, _I, _P, _S]),
    "gpk_batched_mll_grad": (_I, [_I, _I, _P, _P, _P, _L, _L, _P, _I, _D, _I, _P, _S, _P, _P,
                                  _P]),
    "gpk_batched_tune": (_I, [_I, _I, _P, _P, _P, _L, _L, _P, _I, _D, _I, _P, _S, _P, _P, _P,
                              _I, _P]),
    "gpk_linalg_workspace_bytes": (_S, [_L, _L, _I]),
    "gpk_potrf": (_I, [_I, _P, _L, _L, _P, _L, _P, _P, _P, _S, _P]),
    "gpk_trsm": (_I, [_P, _L, _L, _P, _L, _L, _P, _S, _P]),
    "gpk_potri": (_I, [_P, _L, _L, _P, _L, _P, _S, _P]),
    "gpk_logdet_tri": (_I, [_P, _L, _L, _P, _P]),
    "gpk_mvn_workspace_bytes": (_S, [_L, _I]),
    "gpk_mvn_log_prob": (_I, [_P, _P, _L, _L, _I, _P, _P, _P, _S, _P]),
    "gpk_adam_step": (_I, [_L, _P, _P, _P, _P, _P, _P, _D, _D, _D, _D, _P, _P, _P, _L, _P]),
    "gpk_profile_enable": (_I, [_I]),
    "gpk_profile_collect": (_I, [_P, _P]),
}

PHASES = ("gram", "potrf", "trtri", "solve", "adjoint")


def profile_enable(on: bool) -> None:
    call("gpk_profile_enable", 1 if on else 0)


def profile_collect() -> dict:
    """Per-phase (total ms, count) since the last collect; synchronises."""
    ms = (ctypes.c_double * len(PHASES))()
    cnt = (ctypes.c_int * len(PHASES))()
    call("gpk_profile_collect", ctypes.cast(ms, _P), ctypes.cast(cnt, _P))
    return {p: (ms[i], cnt[i]) for i, p in enumerate(PHASES)}


class GPKError(RuntimeError):
    """A libgpk call returned a non-OK status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with {STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """Load libgpk.so (once) and declare its signatures."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python gpjax_amd/csrc/build.py` "
                "(or __graft_entry__.build()); gpjax_amd has no host fallback")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def call(name: str, *args) -> int:
    """Call a status-returning entry point; raise GPKError on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != GPK_OK:
        raise GPKError(name, rc, lib().gpk_last_error().decode())
    return rc


def ptr(t) -> int:
    """Device address of a contiguous torch tensor."""
    if not t.is_contiguous():
        raise ValueError("libgpk needs contiguous tensors")
    return t.data_ptr()


def stream_ptr(device) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
