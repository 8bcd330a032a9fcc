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


def construct_posterior(prior: Prior, likelihood):
    if isinstance(likelihood, Gaussian):
        return ConjugatePosterior(prior=prior, likelihood=likelihood)
    raise TypeError("only the Gaussian likelihood (conjugate posterior) is on the GPU path")
