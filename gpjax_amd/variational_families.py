"""Collapsed variational family, mirroring ``gpjax/variational_families.py``.

``CollapsedVariationalGaussian(posterior, inducing_inputs, jitter=1e-6)``
(``variational_families.py:107-131,767-784``): the inducing inputs are a ``Real``
parameter (trainable, gradients from ``objectives.collapsed_elbo``), the
likelihood must be Gaussian.
"""

from __future__ import annotations

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
