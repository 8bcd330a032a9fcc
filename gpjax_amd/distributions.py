"""Multivariate normal returned by ``predict`` (``gpjax/distributions.py:42-108``).

Holds a mean vector and a covariance (torch float64): either the dense [T, T]
matrix or, for a diagonal covariance (``predict(..., full_covariance=False)``),
the [T] vector of variances, so that a predictive at large T needs O(T) memory.
``variance`` and ``stddev`` read the diagonal; ``covariance()`` forms the dense
matrix (for a diagonal one, only when asked); ``log_prob`` and ``entropy`` run
on the device Cholesky (``gpjax_amd.linalg``).
"""

from __future__ import annotations

import torch


class GaussianDistribution:
    def __init__(self, loc: torch.Tensor, scale: torch.Tensor):
        if scale.ndim not in (1, 2):
            raise ValueError("scale must be a [T, T] covariance or a [T] diagonal")
        self.loc = loc
        self.scale = scale

    @property
    def is_diagonal(self) -> bool:
        return self.scale.ndim == 1

    @property
    def mean(self) -> torch.Tensor:
        return self.loc

    @property
    def variance(self) -> torch.Tensor:
        return self.scale if self.is_diagonal else torch.diagonal(self.scale)

    def stddev(self) -> torch.Tensor:
        return torch.sqrt(self.variance)

    def covariance(self) -> torch.Tensor:
        return torch.diag(self.scale) if self.is_diagonal else self.scale

    @property
    def covariance_matrix(self) -> torch.Tensor:
        return self.covariance()

    @property
    def event_shape(self):
        return tuple(self.loc.shape)

    def log_prob(self, y: torch.Tensor) -> torch.Tensor:
        """log N(y; loc, scale) = -1/2 [n log 2 pi + log|scale| + (y - loc)^T scale^-1
        (y - loc)] (``gpjax/distributions.py:115-134``), through the device Cholesky
        (``gpk_mvn_log_prob``); a diagonal scale is evaluated directly."""
        from .linalg import mvn_log_prob
        return mvn_log_prob(self.loc, self.scale, y)[0]

    def entropy(self) -> torch.Tensor:
        """1/2 [n (1 + log 2 pi) + log|scale|] (``gpjax/distributions.py:82-86``)."""
        from .linalg import mvn_log_prob
        return mvn_log_prob(self.loc, self.scale, self.loc)[3]
