"""Many independent exact GPs evaluated together (the many-small-GPs workload).

GPJax evaluates ``conjugate_mll`` (``gpjax/objectives.py:36-107``) for one
model at a time; a user with thousands of small regression problems maps it
over the problems (``jax.vmap`` of the objective and its ``jax.grad``).  Here
one ``gpk_batched_mll_grad`` call returns every problem's MLL and gradient:
problems are grouped by padded size and each group runs the exact-GP kernels
once, one problem per ``blockIdx.z``.

``BatchedConjugateMLL(X_list, y_list, kernel="rbf", ...)`` owns the
hyperparameters as ``PositiveReal`` parameters of shape ``[B, D]`` (ARD) or
``[B]`` (isotropic), ``[B]`` and ``[B]``; ``model.mll()`` is a differentiable
float64 tensor ``[B]``.  Zero mean.  The batch's problem map is written to the
device once, when the model is built, so ``mll()`` only launches kernels and
can be captured in a HIP graph (``gpjax_amd.fit(..., capture_graph=True)``).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, config
from .parameters import PositiveReal


class _BatchedMLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ls, var, sd, model):
        B, nls = model.count, model.nls
        dev = model.X.device
        theta = torch.cat([ls.reshape(B, nls), var.reshape(B, 1), sd.reshape(B, 1),
                           model._zero_col], dim=1).to(dev, torch.float64).contiguous()
        result = torch.empty((B, _lib.RESULT_LEN), dtype=torch.float64, device=dev)
        grad = torch.empty((B, nls + 3), dtype=torch.float64, device=dev)
        _lib.call("gpk_batched_mll_grad", model.kind_id, _lib.ptr(model.X), _lib.ptr(model.y),
                  model._offsets, B, model.d, _lib.ptr(theta), nls,
                  float(model.jitter), -1, _lib.ptr(model._ws), model._ws_bytes,
                  _lib.ptr(result), _lib.ptr(grad), _lib.stream_ptr(dev))
        ctx.save_for_backward(grad)
        ctx.shapes = (ls.shape, var.shape, sd.shape)
        ctx.nls = nls
        model.last_result = result
        return result[:, _lib.RES_MLL].clone()

    @staticmethod
    def backward(ctx, gbar):
        (grad,) = ctx.saved_tensors
        nls = ctx.nls
        g = grad * gbar.reshape(-1, 1)
        ls_s, var_s, sd_s = ctx.shapes
        return (g[:, :nls].reshape(ls_s), g[:, nls].reshape(var_s), g[:, nls + 1].reshape(sd_s),
                None)


class BatchedConjugateMLL:
    """B independent zero-mean exact GPs with Gaussian likelihoods."""

    def __init__(self, X_list, y_list, kernel: str = "rbf", lengthscale=1.0, variance=1.0,
                 obs_stddev=0.1, jitter: float = 1e-6, ard: bool = True, device=None):
        if kernel not in _lib.KERNEL_IDS:
            raise ValueError(f"kernel must be one of {sorted(_lib.KERNEL_IDS)}")
        if len(X_list) != len(y_list) or not X_list:
            raise ValueError("need as many target vectors as input arrays (at least one)")
        dev = torch.device(device or config.device)
        Xs = [np.asarray(x, dtype=np.float64).reshape(len(x), -1) for x in X_list]
        d = Xs[0].shape[1]
        if any(x.shape[1] != d for x in Xs):
            raise ValueError("all problems need the same input dimension")
        ns = [x.shape[0] for x in Xs]
        if any(len(np.asarray(y).reshape(-1)) != n for y, n in zip(y_list, ns)):
            raise ValueError("X and y disagree in length")
        self.count, self.d, self.kind_id = len(Xs), d, _lib.KERNEL_IDS[kernel]
        self.nls = d if ard else 1
        self.jitter = jitter
        self.X = torch.tensor(np.concatenate(Xs, 0), dtype=torch.float64, device=dev)
        self.y = torch.tensor(np.concatenate([np.asarray(y, dtype=np.float64).reshape(-1)
                                              for y in y_list]), device=dev)
        off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
        self._offsets = (ctypes.c_int64 * len(off))(*off.tolist())
        B = self.count
        ls = torch.as_tensor(lengthscale, dtype=torch.float64)
        if ard:   # scalar, [D] or [B, D]
            ls = ls.expand(B, d) if ls.dim() < 2 else ls.reshape(B, d)
        else:     # scalar or [B]
            ls = ls.expand(B) if ls.dim() == 0 else ls.reshape(B)
        self.lengthscale = PositiveReal(ls.clone())
        self.variance = PositiveReal(torch.as_tensor(variance, dtype=torch.float64).expand(B)
                                     .clone())
        self.obs_stddev = PositiveReal(torch.as_tensor(obs_stddev, dtype=torch.float64)
                                       .expand(B).clone())
        self._zero_col = torch.zeros((B, 1), dtype=torch.float64, device=dev)
        lib = _lib.lib()
        self._ws_bytes = lib.gpk_batched_workspace_bytes(self._offsets, B, d)
        if self._ws_bytes == 0:
            raise _lib.GPKError("gpk_batched_workspace_bytes", 1, lib.gpk_last_error().decode())
        self._ws = torch.empty(self._ws_bytes, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize(dev)
        _lib.call("gpk_batched_plan", self._offsets, B, d, _lib.ptr(self._ws), self._ws_bytes)
        self.last_result = None

    @property
    def sizes(self) -> list[int]:
        return [self._offsets[i + 1] - self._offsets[i] for i in range(self.count)]

    def mll(self) -> torch.Tensor:
        """Log marginal likelihood of every problem, [B], differentiable."""
        return _BatchedMLL.apply(self.lengthscale.value, self.variance.value,
                                 self.obs_stddev.value, self)
