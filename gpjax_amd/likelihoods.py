"""Gaussian likelihood, mirroring ``gpjax/likelihoods.py:129-186``."""

from __future__ import annotations

import torch

from .distributions import GaussianDistribution
from .parameters import NonNegativeReal


class Gaussian:
    def __init__(self, num_datapoints: int, obs_stddev=1.0):
        self.num_datapoints = num_datapoints
        if not isinstance(obs_stddev, NonNegativeReal):
            obs_stddev = NonNegativeReal(obs_stddev)
        self.obs_stddev = obs_stddev

    def predict(self, dist: GaussianDistribution) -> GaussianDistribution:
        """Add obs_stddev^2 to the covariance diagonal (``likelihoods.py:182-186``)."""
        noise = self.obs_stddev.value.detach() ** 2
        if dist.is_diagonal:
            return GaussianDistribution(dist.mean, dist.variance + noise)
        cov = dist.covariance().clone()
        cov.diagonal().add_(noise)
        return GaussianDistribution(dist.mean, cov)

    def __call__(self, dist: GaussianDistribution) -> GaussianDistribution:
        return self.predict(dist)
