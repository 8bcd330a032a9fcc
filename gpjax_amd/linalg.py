"""Dense linear algebra on the libgpk path, mirroring ``gpjax/linalg/operations.py``.

GPJax dispatches ``lower_cholesky`` / ``solve`` / ``logdet`` on operator types
(``operations.py:22-181``); its call sites on the hot path hand them ``Dense``
(PSD) and ``Triangular`` operators.  Here the operators are float64 CUDA
tensors and the type is named by the function:

  lower_cholesky(A)         Dense A -> lower L, L L^T = A   (operations.py:54-55)
  solve_triangular(L, b)    Triangular lower solve          (operations.py:105-107)
  logdet(A)                 Dense PD A: log|A|              (operations.py:163-165)
  logdet_triangular(L)      sum log L_ii                    (operations.py:145-147)
  inverse_from_cholesky(L)  (L L^T)^-1, the dense inverse   (objectives.py:171)

All of them run in ``libgpk.so`` (recursive Cholesky and triangular inverse on
the matrix cores); a non-positive-definite ``A`` gives NaN (and ``info``), where
the reference's LU ``slogdet`` returns a finite value with the sign dropped.
"""

from __future__ import annotations

import torch

from . import _lib

PRECISION_IDS = {"fp64": 0, "mixed": 1}


def _check_square(A: torch.Tensor, what: str) -> int:
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError(f"{what}: expected a square matrix, got shape {tuple(A.shape)}")
    if not A.is_cuda or A.dtype != torch.float64:
        raise ValueError(f"{what}: expected a float64 CUDA tensor")
    return A.shape[0]


def _workspace(n: int, m: int, prec: int, dev) -> tuple[torch.Tensor, int]:
    nbytes = _lib.lib().gpk_linalg_workspace_bytes(n, m, prec)
    return torch.empty(nbytes, dtype=torch.uint8, device=dev), nbytes


def cholesky_with_info(A: torch.Tensor, precision: str = "fp64"):
    """(L, logdet, info): L = chol(A) (symmetrised input, as jnp.linalg.cholesky),
    log|A| = 2 sum log L_ii and info = 0 or 1 + the first non-positive pivot."""
    n = _check_square(A, "lower_cholesky")
    A = A.contiguous()
    prec = PRECISION_IDS[precision]
    ws, nbytes = _workspace(n, 0, prec, A.device)
    L = torch.empty_like(A)
    ld = torch.empty((), dtype=torch.float64, device=A.device)
    info = torch.empty((), dtype=torch.int32, device=A.device)
    _lib.call("gpk_potrf", prec, _lib.ptr(A), n, n, _lib.ptr(L), n, _lib.ptr(ld), _lib.ptr(info),
              _lib.ptr(ws), nbytes, _lib.stream_ptr(A.device))
    return L, ld, info


def lower_cholesky(A: torch.Tensor, precision: str = "fp64") -> torch.Tensor:
    """Lower Cholesky factor of a PD matrix (``operations.py:22-70``, Dense case)."""
    return cholesky_with_info(A, precision)[0]


def logdet(A: torch.Tensor) -> torch.Tensor:
    """log|A| of a PD matrix (the Dense case of ``operations.py:123-181``)."""
    return cholesky_with_info(A)[1]


def logdet_triangular(L: torch.Tensor) -> torch.Tensor:
    """sum log L_ii (the Triangular case, ``operations.py:145-147``)."""
    n = _check_square(L, "logdet")
    L = L.contiguous()
    out = torch.empty((), dtype=torch.float64, device=L.device)
    _lib.call("gpk_logdet_tri", _lib.ptr(L), n, n, _lib.ptr(out), _lib.stream_ptr(L.device))
    return out


def solve_triangular(L: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """x with L x = b for a lower-triangular L (``operations.py:105-107``);
    b is [N] or [N, M], as in ``solve`` (``operations.py:91-120``)."""
    n = _check_square(L, "solve")
    if b.shape[0] != n:
        raise ValueError("solve: right-hand side does not match the operator")
    squeeze = b.ndim == 1
    B = (b[:, None] if squeeze else b).to(torch.float64).contiguous().clone()
    m = B.shape[1]
    ws, nbytes = _workspace(n, m, 0, L.device)
    _lib.call("gpk_trsm", _lib.ptr(L.contiguous()), n, n, _lib.ptr(B), m, m, _lib.ptr(ws),
              nbytes, _lib.stream_ptr(L.device))
    return B[:, 0] if squeeze else B


def inverse_from_cholesky(L: torch.Tensor) -> torch.Tensor:
    """(L L^T)^-1 as a dense symmetric matrix."""
    n = _check_square(L, "inverse")
    out = torch.empty((n, n), dtype=torch.float64, device=L.device)
    ws, nbytes = _workspace(n, 0, 0, L.device)
    _lib.call("gpk_potri", _lib.ptr(L.contiguous()), n, n, _lib.ptr(out), n, _lib.ptr(ws),
              nbytes, _lib.stream_ptr(L.device))
    return out


def inverse(A: torch.Tensor) -> torch.Tensor:
    """Dense inverse of a PD matrix through its Cholesky factor."""
    return inverse_from_cholesky(lower_cholesky(A))


def mvn_log_prob(loc: torch.Tensor, scale: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """[log N(y; loc, scale), log|scale|, quadratic term, entropy] (device float64);
    scale is a dense [N, N] covariance or a [N] vector of variances."""
    n = loc.shape[-1]
    diag = 1 if scale.ndim == 1 else 0
    if diag and scale.shape[0] != n or not diag and tuple(scale.shape) != (n, n):
        raise ValueError("log_prob: scale does not match loc")
    dev = loc.device
    loc = loc.reshape(-1).to(torch.float64).contiguous()
    y = y.reshape(-1).to(torch.float64).contiguous()
    if y.shape[0] != n:
        raise ValueError("log_prob: value does not match the event shape")
    S = scale.to(torch.float64).contiguous()
    nbytes = _lib.lib().gpk_mvn_workspace_bytes(n, diag)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    out = torch.empty(4, dtype=torch.float64, device=dev)
    _lib.call("gpk_mvn_log_prob", _lib.ptr(loc), _lib.ptr(S), n, n, diag, _lib.ptr(y),
              _lib.ptr(out), _lib.ptr(ws), nbytes, _lib.stream_ptr(dev))
    return out
