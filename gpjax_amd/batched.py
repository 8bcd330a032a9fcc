"""Many independent exact GPs evaluated together (the many-small-GPs workload).

GPJax evaluates ``conjugate_mll`` (``gpjax/objectives.py:36-107``) for one
model at a time; a user with thousands of small regression problems maps it
over the problems (``jax.vmap`` of the objective and its ``jax.grad``).  Here
one ``gpk_batched_mll_grad`` call returns every problem's MLL and gradient:
problems are grouped by padded size (multiples of 128) and each group runs the
exact-GP kernels once, one problem per ``blockIdx.z``.  At construction the
model autotunes the batch (``gpk_batched_tune``): each group's launch variants
are timed on its own problems and the fastest is kept by the library.

``BatchedConjugateMLL(X_list, y_list, kernel="rbf", ...)`` owns the
hyperparameters as ``PositiveReal`` parameters of shape ``[B, D]`` (ARD) or
``[B]`` (isotropic), ``[B]`` and ``[B]``; ``model.mll()`` is a differentiable
float64 tensor ``[B]``.  Zero mean.  The batch's problem map is written to the
device once, when the model is built, so ``mll()`` only launches kernels and
can be captured in a HIP graph (``gpjax_amd.fit(..., capture_graph=True)``).

``precision="mixed"`` (or ``"mixed:P,T,L"`` with bf16 stages) runs every
problem on the mixed path, each with its own a-posteriori probe check;
``config.mixed_check`` then applies to the batch as a whole: when any problem
is flagged (by its probes, or by a Cholesky that broke down in float32),
"fallback" recomputes the batch in float64, "raise" raises,
"warn" warns (outside graph capture: the check reads the flags on the host).
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
        theta = model._theta(ls, var, sd)
        result = torch.empty((B, _lib.RESULT_LEN), dtype=torch.float64, device=dev)
        grad = torch.empty((B, nls + 3), dtype=torch.float64, device=dev)
        model._run(model.prec_id, theta, result, grad)
        if model._batch_says_recompute(result):
            model._run(_lib.PRECISIONS["fp64"], theta, result, grad)
            result[:, _lib.RES_FLAG] = 2.0   # recomputed in float64
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
                 obs_stddev=0.1, jitter: float = 1e-6, ard: bool = True, device=None,
                 precision: str = "fp64", autotune: bool = True):
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
        from .objectives import _precision_id
        self.prec_id = _precision_id(precision)
        self._plans = {}
        self._plan(self.prec_id)
        self.last_result = None
        self.tune_report = self.autotune() if autotune else None

    def _theta(self, ls, var, sd) -> torch.Tensor:
        B = self.count
        return torch.cat([ls.reshape(B, self.nls), var.reshape(B, 1), sd.reshape(B, 1),
                          self._zero_col], dim=1).to(self.X.device, torch.float64).contiguous()

    def autotune(self) -> list[dict]:
        """Time the launch variants of every size group on this batch (at the
        current hyperparameters) and let the library keep the fastest per group
        (``gpk_batched_tune``); returns one record per group.  Host-synchronous:
        call it outside graph capture (the constructor does, unless
        ``autotune=False``)."""
        B, dev = self.count, self.X.device
        with torch.no_grad():
            theta = self._theta(self.lengthscale.value, self.variance.value,
                                self.obs_stddev.value)
        result = torch.empty((B, _lib.RESULT_LEN), dtype=torch.float64, device=dev)
        grad = torch.empty((B, self.nls + 3), dtype=torch.float64, device=dev)
        ws, nbytes = self._plan(self.prec_id)
        ngroups = len({-(-n // 128) for n in self.sizes})
        report = (ctypes.c_float * (7 * ngroups))()
        _lib.call("gpk_batched_tune", self.kind_id, self.prec_id, _lib.ptr(self.X),
                  _lib.ptr(self.y), self._offsets, B, self.d, _lib.ptr(theta), self.nls,
                  float(self.jitter), -1, _lib.ptr(ws), nbytes, _lib.ptr(result), _lib.ptr(grad),
                  report, 7 * ngroups, _lib.stream_ptr(dev))
        out = []
        for g in range(ngroups):
            q = report[7 * g: 7 * g + 7]
            out.append({"padded_size": int(q[0]), "members": int(q[1]),
                        "stored_sigma": bool(int(q[2]) & 1),
                        "transposed_pass": bool(int(q[2]) & 2),
                        "ms": {"recomputed/columns": q[3], "stored/columns": q[4],
                               "recomputed/transposed": q[5], "stored/transposed": q[6]}})
        return out

    def _plan(self, prec_id: int):
        """Workspace of this batch at `prec_id`, planned once (outside capture)."""
        if prec_id not in self._plans:
            lib = _lib.lib()
            nbytes = lib.gpk_batched_workspace_bytes(self._offsets, self.count, self.d, prec_id)
            if nbytes == 0:
                raise _lib.GPKError("gpk_batched_workspace_bytes", 1,
                                    lib.gpk_last_error().decode())
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.X.device)
            torch.cuda.synchronize(self.X.device)
            _lib.call("gpk_batched_plan", self._offsets, self.count, self.d, prec_id,
                      _lib.ptr(ws), nbytes)
            self._plans[prec_id] = (ws, nbytes)
        return self._plans[prec_id]

    def _run(self, prec_id: int, theta, result, grad):
        ws, nbytes = self._plan(prec_id)
        _lib.call("gpk_batched_mll_grad", self.kind_id, prec_id, _lib.ptr(self.X),
                  _lib.ptr(self.y), self._offsets, self.count, self.d, _lib.ptr(theta), self.nls,
                  float(self.jitter), -1, _lib.ptr(ws), nbytes, _lib.ptr(result), _lib.ptr(grad),
                  _lib.stream_ptr(self.X.device))

    def _batch_says_recompute(self, result) -> bool:
        """config.mixed_check over the batch: True to recompute it in float64."""
        if self.prec_id == _lib.PRECISIONS["fp64"] or config.mixed_check == "off":
            return False
        if torch.cuda.is_current_stream_capturing():
            return False
        # flagged by the probe check, or a Cholesky that broke down in float32
        bad = (result[:, _lib.RES_FLAG] == 1.0) | (result[:, _lib.RES_INFO] != 0.0)
        nflag = int(bad.sum())
        if nflag == 0:
            return False
        import warnings
        from .objectives import MixedPrecisionError
        msg = (f"batched MLL: the mixed-precision check flagged {nflag} of {self.count} "
               "problems (probe estimate or Cholesky breakdown)")
        if config.mixed_check == "raise":
            raise MixedPrecisionError(msg)
        if config.mixed_check == "warn":
            warnings.warn(msg, RuntimeWarning, stacklevel=4)
            return False
        warnings.warn(msg + "; recomputing the batch in float64", RuntimeWarning, stacklevel=4)
        return True

    @property
    def sizes(self) -> list[int]:
        return [self._offsets[i + 1] - self._offsets[i] for i in range(self.count)]

    def mll(self) -> torch.Tensor:
        """Log marginal likelihood of every problem, [B], differentiable."""
        return _BatchedMLL.apply(self.lengthscale.value, self.variance.value,
                                 self.obs_stddev.value, self)
