"""Multivariate normal returned by ``predict`` (``gpjax/distributions.py:42-108``).

Holds a mean vector and a dense covariance (torch float64); ``variance`` and
``stddev`` read the diagonal.  ``log_prob`` of a dense covariance goes through
the same C-ABI factorisation as ``conjugate_mll``.
"""

from __future__ import annotations

import torch


class GaussianDistribution:
    def __init__(self, loc: torch.Tensor, scale: torch.Tensor):
        self.loc = loc
        self.scale = scale

    @property
    def mean(self) -> torch.Tensor:
        return self.loc

    @property
    def variance(self) -> torch.Tensor:
        return torch.diagonal(self.scale)

    def stddev(self) -> torch.Tensor:
        return torch.sqrt(self.variance)

    def covariance(self) -> torch.Tensor:
        return self.scale

    @property
    def covariance_matrix(self) -> torch.Tensor:
        return self.scale

    @property
    def event_shape(self):
        return tuple(self.loc.shape)
