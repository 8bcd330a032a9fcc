"""Collapsed variational family, mirroring ``gpjax/variational_families.py``.

``CollapsedVariationalGaussian(posterior, inducing_inputs, jitter=1e-6)``
(``variational_families.py:107-131,767-784``): the inducing inputs are a ``Real``
parameter (trainable, gradients from ``objectives.collapsed_elbo``), the
likelihood must be Gaussian.  ``predict`` (``variational_families.py:786-870``) runs
on the GPU through ``gpk_elbo_predict``.
"""

from __future__ import annotations

import torch

from .distributions import GaussianDistribution
from .likelihoods import Gaussian
from .parameters import Real


class CollapsedVariationalGaussian:
    def __init__(self, posterior, inducing_inputs, jitter: float = 1e-6):
        if not isinstance(posterior.likelihood, Gaussian):
            raise TypeError("Likelihood must be Gaussian.")
        if not isinstance(inducing_inputs, Real):
            inducing_inputs = Real(inducing_inputs)
        self.posterior = posterior
        self.inducing_inputs = inducing_inputs
        self.jitter = jitter

    @property
    def num_inducing(self) -> int:
        return self.inducing_inputs.value.shape[0]

    def __call__(self, test_inputs, train_data):
        return self.predict(test_inputs, train_data)

    def predict(self, test_inputs: torch.Tensor, train_data,
                full_covariance: bool = True) -> GaussianDistribution:
        """Latent predictive of the collapsed posterior (``variational_families.py:786-870``).

        mean = c + Kzt^T C^-1 Kzx (y - c) / s and
        cov = Ktt - Ktz Kzz^-1 Kzt + Ktz C^-1 Kzt + jitter I with
        C = Kzz + jitter I + Kzx Kxz / s, s = obs_stddev^2.  The training data enter
        through the same N-summed statistics as the ELBO.  ``full_covariance=False``
        forms only the diagonal, returned as a [T] variance vector (O(T) memory).
        Values only: no autograd graph is recorded.
        """
        from . import _lib
        from .objectives import _mean_constant

        post = self.posterior
        kernel = post.prior.kernel
        X = kernel.slice_input(train_data.X)
        if not X.is_cuda:
            raise ValueError("predict: data must live on a GPU device")
        T = kernel.slice_input(test_inputs)
        y = train_data.y.reshape(-1).to(torch.float64).contiguous()
        n, d = X.shape
        t = T.shape[0]
        if T.shape[1] != d:
            raise ValueError("test inputs and training inputs differ in dimension")
        m = self.num_inducing
        nls = kernel.nls(d)
        dev = X.device
        with torch.no_grad():
            Z = self.inducing_inputs.value.detach().to(dev, torch.float64).contiguous()
            if Z.ndim != 2 or Z.shape[1] != d:
                raise ValueError("inducing inputs must be [M, D] with the data's D")
            c = _mean_constant(post.prior.mean_function, dev)
            theta = kernel.theta_tensor(d, post.likelihood.obs_stddev.value, c)
            full = 1 if full_covariance else 0
            nbytes = _lib.lib().gpk_elbo_predict_workspace_bytes(n, m, d, t, full)
            ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            mean = torch.empty(t, dtype=torch.float64, device=dev)
            cov = torch.empty((t, t) if full else (t,), dtype=torch.float64, device=dev)
            _lib.call("gpk_elbo_predict", kernel.kind_id, 0, _lib.ptr(X), _lib.ptr(y), n,
                      _lib.ptr(Z), m, d, _lib.ptr(T), t, _lib.ptr(theta), nls,
                      float(self.jitter), full, _lib.ptr(ws), nbytes, _lib.ptr(mean),
                      _lib.ptr(cov), _lib.stream_ptr(dev))
        return GaussianDistribution(mean, cov)
