"""Stationary kernels on the libgpk Gram builder.

Mirrors ``gpjax/kernels``: ``StationaryKernel`` (``stationary/base.py:42-106``:
``lengthscale`` PositiveReal of shape [] or [D], ``variance`` NonNegativeReal,
``active_dims``), ``RBF`` (``stationary/rbf.py:28-44``), ``Matern12``
(``matern12.py:31-48``), ``Matern32`` (``matern32.py:28-54``), ``Matern52``
(``matern52.py:28-53``), and the compute-engine seam of
``kernels/computations/base.py:36-109`` / ``dense.py:27-36``: ``gram``,
``cross_covariance`` and ``diagonal`` delegate to ``compute_engine``, whose
default here is :class:`GPKKernelComputation` (HIP kernels through the C ABI).

Matrices returned by ``gram`` / ``cross_covariance`` are values (no autograd
graph); the objectives differentiate through their own fused adjoint.
"""

from __future__ import annotations

import torch

from . import _lib
from .parameters import NonNegativeReal, Parameter, PositiveReal


class GPKKernelComputation:
    """Compute engine backed by ``gpk_gram`` / ``gpk_cross_covariance``.

    Contract of ``AbstractKernelComputation`` (``computations/base.py:56-109``):
    ``gram(kernel, x)`` [N,N], ``cross_covariance(kernel, x, y)`` [N,M],
    ``diagonal(kernel, x)`` [N].
    """

    def gram(self, kernel: "StationaryKernel", x: torch.Tensor) -> torch.Tensor:
        xs = kernel.slice_input(x)
        n, d = xs.shape
        theta = kernel.theta_tensor(d)
        out = torch.empty((n, n), dtype=torch.float64, device=xs.device)
        _lib.call("gpk_gram", kernel.kind_id, _lib.ptr(xs), n, d, _lib.ptr(theta),
                  kernel.nls(d), 0.0, 0, _lib.ptr(out), n, _lib.stream_ptr(xs.device))
        return out

    def cross_covariance(self, kernel, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        xs, ys = kernel.slice_input(x), kernel.slice_input(y)
        n, d = xs.shape
        m = ys.shape[0]
        if ys.shape[1] != d:
            raise ValueError("cross_covariance: inputs differ in dimension")
        theta = kernel.theta_tensor(d)
        out = torch.empty((n, m), dtype=torch.float64, device=xs.device)
        _lib.call("gpk_cross_covariance", kernel.kind_id, _lib.ptr(xs), n, _lib.ptr(ys), m, d,
                  _lib.ptr(theta), kernel.nls(d), _lib.ptr(out), m, _lib.stream_ptr(xs.device))
        return out

    def diagonal(self, kernel, x: torch.Tensor) -> torch.Tensor:
        # k(x, x) = variance * k(0): stationary kernels (computations/base.py:95-96).
        var = kernel.variance.value.detach()
        return torch.ones(x.shape[0], dtype=torch.float64, device=x.device) * var


class StationaryKernel:
    name = "StationaryKernel"
    kind = None

    def __init__(self, active_dims=None, lengthscale=1.0, variance=1.0, n_dims=None,
                 compute_engine=None):
        self.active_dims = active_dims
        ls = lengthscale if isinstance(lengthscale, Parameter) else PositiveReal(lengthscale)
        if ls.value.ndim > 1:
            raise ValueError("lengthscale must be a scalar or a 1-D array")
        if ls.value.ndim == 1 and n_dims is not None and ls.value.shape[0] != n_dims:
            raise ValueError(
                "Expected `lengthscale` to be compatible with the number of input dimensions. "
                f"Got `lengthscale` with shape {tuple(ls.value.shape)}, but the number of input "
                f"dimensions is {n_dims}.")
        self.lengthscale = ls
        self.variance = variance if isinstance(variance, Parameter) else NonNegativeReal(variance)
        self.n_dims = ls.value.shape[0] if ls.value.ndim == 1 else n_dims
        self.compute_engine = compute_engine or GPKKernelComputation()

    @property
    def kind_id(self) -> int:
        return _lib.KERNEL_IDS[self.kind]

    def slice_input(self, x: torch.Tensor) -> torch.Tensor:
        """``x[..., active_dims]`` (``kernels/base.py:137-149``), contiguous float64."""
        if self.active_dims is not None:
            x = x[..., self.active_dims]
        return x.to(torch.float64).contiguous()

    def nls(self, d: int) -> int:
        ls = self.lengthscale.value
        if ls.ndim == 0 or ls.numel() == 1:
            return 1
        if ls.numel() != d:
            raise ValueError(f"lengthscale has {ls.numel()} entries for {d} input dimensions")
        return d

    def theta_tensor(self, d: int, obs_stddev=None, mean=None) -> torch.Tensor:
        """Packed [lengthscale(s), variance, obs_stddev, mean] (layout of include/gpk.h)."""
        dev = self.lengthscale.value.device
        z = torch.zeros(1, dtype=torch.float64, device=dev)
        parts = [self.lengthscale.value.reshape(-1)[: self.nls(d)],
                 self.variance.value.reshape(1),
                 z if obs_stddev is None else obs_stddev.reshape(1),
                 z if mean is None else mean.reshape(1)]
        return torch.cat([p.to(dev, torch.float64) for p in parts]).contiguous()

    def gram(self, x: torch.Tensor) -> torch.Tensor:
        with torch.no_grad():
            return self.compute_engine.gram(self, x)

    def cross_covariance(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        with torch.no_grad():
            return self.compute_engine.cross_covariance(self, x, y)

    def diagonal(self, x: torch.Tensor) -> torch.Tensor:
        with torch.no_grad():
            return self.compute_engine.diagonal(self, x)

    def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        return self.cross_covariance(x.reshape(1, -1), y.reshape(1, -1)).reshape(())

    def parameters(self) -> dict:
        return {"lengthscale": self.lengthscale, "variance": self.variance}


class RBF(StationaryKernel):
    name = "RBF"
    kind = "rbf"


class Matern12(StationaryKernel):
    name = "Matérn12"
    kind = "matern12"


class Matern32(StationaryKernel):
    name = "Matérn32"
    kind = "matern32"


class Matern52(StationaryKernel):
    name = "Matérn52"
    kind = "matern52"
