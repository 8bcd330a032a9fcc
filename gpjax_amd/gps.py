"""GP prior / posterior containers, mirroring ``gpjax/gps.py``.

``Prior(kernel, mean_function, jitter=1e-6)`` (``gps.py:61-116,122-250``),
``prior * likelihood`` -> ``construct_posterior`` (``gps.py:160-189,724-743``),
``ConjugatePosterior(prior, likelihood, jitter=1e-6)`` (``gps.py:333-526``).
"""

from __future__ import annotations

import torch

from .distributions import GaussianDistribution
from .likelihoods import Gaussian


class Prior:
    def __init__(self, kernel, mean_function, jitter: float = 1e-6):
        self.kernel = kernel
        self.mean_function = mean_function
        self.jitter = jitter

    def __mul__(self, other):
        return construct_posterior(prior=self, likelihood=other)

    def __rmul__(self, other):
        return self.__mul__(other)

    def predict(self, test_inputs: torch.Tensor) -> GaussianDistribution:
        """m(t) and K(t, t) + jitter I (``gps.py:244-250``)."""
        mean = self.mean_function(test_inputs).reshape(-1)
        K = self.kernel.gram(test_inputs)
        K.diagonal().add_(self.jitter)
        return GaussianDistribution(mean, K)

    def __call__(self, test_inputs):
        return self.predict(test_inputs)


class ConjugatePosterior:
    def __init__(self, prior: Prior, likelihood: Gaussian, jitter: float = 1e-6):
        self.prior = prior
        self.likelihood = likelihood
        self.jitter = jitter

    def __call__(self, test_inputs, train_data):
        return self.predict(test_inputs, train_data)

    def predict(self, test_inputs: torch.Tensor, train_data, full_covariance: bool = True,
                precision: str | None = None) -> GaussianDistribution:
        """Latent predictive N(mean, cov) at ``test_inputs`` (``gps.py:443-526``).

        Sigma uses this posterior's jitter (``gps.py:506``) and the predictive
        covariance the prior's (``gps.py:523``).  With ``full_covariance=False``
        only the diagonal is formed and returned as a [T] variance vector (the
        test points pass through the device in passes), so large T costs O(T)
        memory rather than a T x T product.  Values only: no autograd graph is
        recorded through predict.
        """
        from . import _lib, config
        from .objectives import _mean_constant, _precision_id, mixed_check_says_recompute

        kernel = self.prior.kernel
        X = kernel.slice_input(train_data.X)
        T = kernel.slice_input(test_inputs)
        y = train_data.y.reshape(-1).to(torch.float64).contiguous()
        n, d = X.shape
        t = T.shape[0]
        if T.shape[1] != d:
            raise ValueError("test inputs and training inputs differ in dimension")
        nls = kernel.nls(d)
        prec = _precision_id(precision)
        with torch.no_grad():
            c = _mean_constant(self.prior.mean_function, X.device)
            theta = kernel.theta_tensor(d, self.likelihood.obs_stddev.value, c)
            full = 1 if full_covariance else 0
            iters = -1 if config.refine_iters is None else int(config.refine_iters)

            def run(prec):
                nbytes = _lib.lib().gpk_predict_workspace_bytes(n, d, t, prec, full)
                ws = torch.empty(nbytes, dtype=torch.uint8, device=X.device)
                mean = torch.empty(t, dtype=torch.float64, device=X.device)
                cov = torch.empty((t, t) if full else (t,), dtype=torch.float64,
                                  device=X.device)
                res = torch.empty(_lib.RESULT_LEN, dtype=torch.float64, device=X.device)
                _lib.call("gpk_predict", kernel.kind_id, prec, _lib.ptr(X), _lib.ptr(y), n, d,
                          _lib.ptr(T), t, _lib.ptr(theta), nls, float(self.jitter),
                          float(self.prior.jitter), full, iters, _lib.ptr(ws), nbytes,
                          _lib.ptr(mean), _lib.ptr(cov), _lib.ptr(res),
                          _lib.stream_ptr(X.device))
                return mean, cov, res

            mean, cov, res = run(prec)
            if mixed_check_says_recompute(res, prec, n, "predict"):
                mean, cov, res = run(_lib.PRECISIONS["fp64"])
                res[_lib.RES_FLAG] = 2.0
        dist = GaussianDistribution(mean, cov)
        dist.factor_result = res
        return dist


def construct_posterior(prior: Prior, likelihood):
    if isinstance(likelihood, Gaussian):
        return ConjugatePosterior(prior=prior, likelihood=likelihood)
    raise TypeError("only the Gaussian likelihood (conjugate posterior) is on the GPU path")
