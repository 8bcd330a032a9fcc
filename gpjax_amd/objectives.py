"""Objectives on the libgpk path.

``conjugate_mll(posterior, data)`` has the signature and meaning of
``gpjax/objectives.py:36-107``: the log marginal likelihood
-1/2 [n log 2pi + log|Sigma| + r^T Sigma^-1 r], Sigma = K + jitter I + obs_sd^2 I
(``objectives.py:96-103``), r = y - m(x).  It returns a float64 torch scalar whose
backward is the hand-written adjoint (``gpk_mll_backward``) -- the role
``jax.custom_vjp`` plays for a JAX caller -- giving gradients w.r.t. the
lengthscale(s), variance, obs_stddev and a trainable constant mean in
constrained coordinates.

Divergence from the reference, by design: the reference evaluates log|Sigma|
with an LU ``slogdet`` and discards the sign (``linalg/operations.py:164``), so
an indefinite Sigma still yields a finite number there.  The Cholesky path here
returns NaN and reports the failing pivot (``info``) instead.
"""

from __future__ import annotations

import math
import warnings

import torch

from . import _lib, config
from .mean_functions import Constant


class MixedPrecisionError(RuntimeError):
    """The mixed-precision path's a-posteriori check flagged its result."""


def _is_mixed(prec_id: int) -> bool:
    return (prec_id & 0xFF) == _lib.PRECISIONS["mixed"]


def mixed_check_says_recompute(res: torch.Tensor, prec_id: int, n: int, what: str,
                               elbo: bool = False) -> bool:
    """Apply ``config.mixed_check`` to a mixed-precision result buffer: True when
    the caller should recompute in float64.  Reads the check's result slots (a
    host sync) unless the check is off, the path is float64, or a graph is being
    captured.  ``elbo``: the buffer is gpk_elbo_forward's."""
    if not _is_mixed(prec_id) or config.mixed_check == "off":
        return False
    if torch.cuda.is_current_stream_capturing():
        return False
    info, flag, probe, err, val = res[[_lib.RES_INFO, _lib.RES_FLAG, _lib.RES_PROBE,
                                       _lib.RES_MLL_ERR, _lib.RES_MLL]].tolist()
    if info != 0.0:
        # float32 storage moves Sigma (or C = Kzz + P / s) by ~1e-7 relative; a
        # pivot that then turns non-positive is a precision failure (float64 may
        # still factor it)
        msg = (f"{what}: the mixed-precision Cholesky broke down at pivot {int(info) - 1} "
               f"(n = {n})")
    elif flag == 1.0 and elbo:
        msg = (f"{what}: the mixed-precision a-posteriori check flagged this problem "
               f"(estimated ELBO error {err:.3g} of |ELBO| = {abs(val):.6g}, n = {n}); its "
               "ELBO / gradient may be outside 1e-6 / 1e-4 of the float64 answer")
    elif flag == 1.0:
        msg = (f"{what}: the mixed-precision a-posteriori check flagged this problem "
               f"(||I - Sigma X^T X||_F ~ {probe * math.sqrt(n):.3g}, n = {n}); its MLL / "
               "gradient may be outside 1e-6 / 1e-4 of the float64 answer")
    else:
        return False
    if config.mixed_check == "raise":
        raise MixedPrecisionError(msg)
    if config.mixed_check == "warn":
        warnings.warn(msg, RuntimeWarning, stacklevel=3)
        return False
    warnings.warn(msg + "; recomputing in float64", RuntimeWarning, stacklevel=3)
    return True


def _precision_id(precision: str | None) -> int:
    """``prec`` code of "fp64", "mixed" or "mixed:P,T,L" (the products of the
    Cholesky updates, the triangular inverse and the fused Sigma^-1 pass)."""
    p = precision or config.precision
    if isinstance(p, str) and p.startswith("mixed:"):
        try:
            stages = [int(v) for v in p[len("mixed:"):].split(",")]
        except ValueError:
            stages = []
        if len(stages) != 3:
            raise ValueError(f"mixed arithmetic must read 'mixed:P,T,L', got {p!r}")
        return _lib.mixed_arith(*stages)
    if p not in _lib.PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)} or "
                         f"'mixed:P,T,L', got {p!r}")
    return _lib.PRECISIONS[p]


# Exact-GP objectives sharing one factorisation and one adjoint pass:
# (workspace size, forward, backward) entry points of include/gpk.h.
_EXACT = {"mll": ("gpk_mll_workspace_bytes", "gpk_mll_forward", "gpk_mll_backward"),
          "loocv": ("gpk_loocv_workspace_bytes", "gpk_loocv_forward", "gpk_loocv_backward")}


class _ConjugateMLL(torch.autograd.Function):
    """Forward: gpk_mll_forward (or gpk_loocv_forward).  Backward: the matching
    backward entry point (fused adjoint)."""

    @staticmethod
    def forward(ctx, ls, var, sd, c, X, y, kind_id, prec_id, jitter, refine, nls, info_out,
                which="mll"):
        n, d = X.shape
        dev = X.device
        ws_fn, fwd_fn, bwd_fn = _EXACT[which]
        theta = torch.cat([ls.reshape(-1)[:nls], var.reshape(1), sd.reshape(1),
                           c.reshape(1)]).to(dev, torch.float64).contiguous()
        it = -1 if refine is None else int(refine)

        def run(prec):
            nbytes = getattr(_lib.lib(), ws_fn)(n, d, prec)
            ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            res = torch.empty(_lib.RESULT_LEN, dtype=torch.float64, device=dev)
            _lib.call(fwd_fn, kind_id, prec, _lib.ptr(X), _lib.ptr(y), n, d,
                      _lib.ptr(theta), nls, float(jitter), it, _lib.ptr(ws), nbytes,
                      _lib.ptr(res), _lib.stream_ptr(dev))
            return ws, nbytes, res

        ws, nbytes, res = run(prec_id)
        if mixed_check_says_recompute(res, prec_id, n, f"conjugate_{which}"):
            prec_id = _lib.PRECISIONS["fp64"]
            ws, nbytes, res = run(prec_id)
            res[_lib.RES_FLAG] = 2.0
        ctx.ws, ctx.nbytes, ctx.theta, ctx.bwd_fn = ws, nbytes, theta, bwd_fn
        ctx.X, ctx.kind_id, ctx.prec_id, ctx.nls = X, kind_id, prec_id, nls
        ctx.shapes = (ls.shape, var.shape, sd.shape, c.shape)
        if info_out is not None:
            info_out["result"] = res
        return res[_lib.RES_MLL].clone()

    @staticmethod
    def backward(ctx, gbar):
        X = ctx.X
        n, d = X.shape
        nls = ctx.nls
        grad = torch.empty(nls + 3, dtype=torch.float64, device=X.device)
        _lib.call(ctx.bwd_fn, ctx.kind_id, ctx.prec_id, _lib.ptr(X), n, d,
                  _lib.ptr(ctx.theta), nls, 1.0, _lib.ptr(ctx.ws), ctx.nbytes, _lib.ptr(grad),
                  _lib.stream_ptr(X.device))
        g = grad * gbar
        ls_s, var_s, sd_s, c_s = ctx.shapes
        ls_g = g[:nls].reshape(ls_s) if math.prod(ls_s) == nls else g[:nls].sum().expand(ls_s)
        ctx.ws = None
        return (ls_g, g[nls].reshape(var_s), g[nls + 1].reshape(sd_s), g[nls + 2].reshape(c_s),
                None, None, None, None, None, None, None, None, None)


def _mean_constant(mean_function, device) -> torch.Tensor:
    if not isinstance(mean_function, Constant):
        raise TypeError("the GPU path supports Constant / Zero mean functions")
    return mean_function.value_tensor(device)


def conjugate_mll_with_info(posterior, data, precision: str | None = None,
                            refine_iters: int | None = None, _which: str = "mll"):
    """(mll, result) where result is the device buffer of ``gpk_mll_forward``
    (slots: mll, logdet, quad, relative refinement residual, info)."""
    kernel = posterior.prior.kernel
    X = kernel.slice_input(data.X)
    if not X.is_cuda:
        raise ValueError(f"conjugate_{_which}: data must live on a GPU device")
    y = data.y.reshape(-1).to(torch.float64).contiguous()
    n, d = X.shape
    if y.shape[0] != n:
        raise ValueError("X and y disagree in length")
    nls = kernel.nls(d)
    c = _mean_constant(posterior.prior.mean_function, X.device)
    info: dict = {}
    mll = _ConjugateMLL.apply(kernel.lengthscale.value, kernel.variance.value,
                              posterior.likelihood.obs_stddev.value, c, X, y, kernel.kind_id,
                              _precision_id(precision), float(posterior.prior.jitter),
                              config.refine_iters if refine_iters is None else refine_iters,
                              nls, info, _which)
    return mll, info["result"]


def conjugate_mll(posterior, data, precision: str | None = None):
    """Log marginal likelihood (``gpjax/objectives.py:36-107``), differentiable."""
    return conjugate_mll_with_info(posterior, data, precision)[0]


def conjugate_loocv_with_info(posterior, data, precision: str | None = None,
                              refine_iters: int | None = None):
    """(loocv, result) where result is the device buffer of ``gpk_loocv_forward``
    (slots: loocv, sum log d_i, sum alpha_i^2 / d_i, refinement residual, info)."""
    return conjugate_mll_with_info(posterior, data, precision, refine_iters, _which="loocv")


def conjugate_loocv(posterior, data, precision: str | None = None):
    """Leave-one-out log predictive probability (``gpjax/objectives.py:110-178``),
    Rasmussen & Williams sec. 5.4.2, differentiable in the kernel hyperparameters,
    obs_stddev and a trainable constant mean.  Sigma = K + (obs_sd^2 + jitter) I
    with the prior's jitter (``objectives.py:163-168``)."""
    return conjugate_loocv_with_info(posterior, data, precision)[0]


class _CollapsedELBO(torch.autograd.Function):
    """Forward: gpk_elbo_forward.  Backward: gpk_elbo_backward (Kzx tiles recomputed)."""

    @staticmethod
    def forward(ctx, ls, var, sd, c, Z, X, y, kind_id, prec_id, jitter, nls, info_out):
        n, d = X.shape
        m = Z.shape[0]
        dev = X.device
        theta = torch.cat([ls.reshape(-1)[:nls], var.reshape(1), sd.reshape(1),
                           c.reshape(1)]).to(dev, torch.float64).contiguous()
        Zc = Z.detach().to(dev, torch.float64).contiguous()

        def run(prec):
            nbytes = _lib.lib().gpk_elbo_workspace_bytes(n, m, d, prec)
            ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            res = torch.empty(_lib.RESULT_LEN, dtype=torch.float64, device=dev)
            _lib.call("gpk_elbo_forward", kind_id, prec, _lib.ptr(X), _lib.ptr(y), n,
                      _lib.ptr(Zc), m, d, _lib.ptr(theta), nls, float(jitter), _lib.ptr(ws),
                      nbytes, _lib.ptr(res), _lib.stream_ptr(dev))
            return ws, nbytes, res

        ws, nbytes, res = run(prec_id)
        if mixed_check_says_recompute(res, prec_id, n, "collapsed_elbo", elbo=True):
            prec_id = _lib.PRECISIONS["fp64"]
            ws, nbytes, res = run(prec_id)
            res[_lib.RES_FLAG] = 2.0
        ctx.ws, ctx.nbytes, ctx.theta, ctx.Z = ws, nbytes, theta, Zc
        ctx.X, ctx.y, ctx.kind_id, ctx.prec_id, ctx.nls = X, y, kind_id, prec_id, nls
        ctx.shapes = (ls.shape, var.shape, sd.shape, c.shape, Z.shape)
        if info_out is not None:
            info_out["result"] = res
        return res[_lib.RES_MLL].clone()

    @staticmethod
    def backward(ctx, gbar):
        X, y, Zc = ctx.X, ctx.y, ctx.Z
        n, d = X.shape
        m = Zc.shape[0]
        nls = ctx.nls
        grad = torch.empty(nls + 3, dtype=torch.float64, device=X.device)
        gz = torch.empty((m, d), dtype=torch.float64, device=X.device)
        _lib.call("gpk_elbo_backward", ctx.kind_id, ctx.prec_id, _lib.ptr(X), _lib.ptr(y), n,
                  _lib.ptr(Zc), m, d, _lib.ptr(ctx.theta), nls, 1.0, _lib.ptr(ctx.ws), ctx.nbytes,
                  _lib.ptr(grad), _lib.ptr(gz), _lib.stream_ptr(X.device))
        g = grad * gbar
        ls_s, var_s, sd_s, c_s, z_s = ctx.shapes
        ls_g = g[:nls].reshape(ls_s) if math.prod(ls_s) == nls else g[:nls].sum().expand(ls_s)
        ctx.ws = None
        return (ls_g, g[nls].reshape(var_s), g[nls + 1].reshape(sd_s), g[nls + 2].reshape(c_s),
                (gz * gbar).reshape(z_s), None, None, None, None, None, None, None)


def collapsed_elbo_with_info(variational_family, data, precision: str | None = None):
    """(elbo, result) -- result is the device buffer of ``gpk_elbo_forward``.
    ``precision``: "fp64", "mixed" (float32 Kzx chunks, P = Kzx Kxz as
    six-product bf16 splits, the gradient pass's Kxz G_P as the fp16 split) or
    "mixed:P,T,L" (P: the product Kzx Kxz, T: the gradient pass's; L unused);
    default ``config.precision``."""
    post = variational_family.posterior
    kernel = post.prior.kernel
    X = kernel.slice_input(data.X)
    if not X.is_cuda:
        raise ValueError("collapsed_elbo: data must live on a GPU device")
    y = data.y.reshape(-1).to(torch.float64).contiguous()
    n, d = X.shape
    Z = variational_family.inducing_inputs.value
    if Z.ndim != 2 or Z.shape[1] != d:
        raise ValueError("inducing inputs must be [M, D] with the data's D")
    nls = kernel.nls(d)
    c = _mean_constant(post.prior.mean_function, X.device)
    info: dict = {}
    elbo = _CollapsedELBO.apply(kernel.lengthscale.value, kernel.variance.value,
                                post.likelihood.obs_stddev.value, c, Z, X, y, kernel.kind_id,
                                _precision_id(precision), float(variational_family.jitter), nls,
                                info)
    return elbo, info["result"]


def collapsed_elbo(variational_family, data, precision: str | None = None):
    """Collapsed ELBO (``gpjax/objectives.py:321-416``), differentiable in the kernel
    hyperparameters, obs_stddev, a trainable constant mean and the inducing inputs."""
    return collapsed_elbo_with_info(variational_family, data, precision)[0]
