"""Host-side pieces of the Python mirror that need no GPU."""

import pytest
import torch

import gpjax_amd as gpx
from gpjax_amd.distributions import GaussianDistribution


def test_diagonal_distribution_stays_a_vector():
    """predict(..., full_covariance=False) returns the variances as a [T] vector:
    variance/stddev read it directly, covariance() forms the dense matrix only
    when asked, and the Gaussian likelihood adds obs_stddev^2 without forming
    it (gpjax/likelihoods.py:182-186)."""
    mu = torch.tensor([0.5, -1.0, 2.0], dtype=torch.float64)
    var = torch.tensor([0.2, 0.3, 0.4], dtype=torch.float64)
    dist = GaussianDistribution(mu, var)
    assert dist.is_diagonal
    assert dist.variance is var
    torch.testing.assert_close(dist.stddev(), var.sqrt())
    torch.testing.assert_close(dist.covariance(), torch.diag(var))
    lik = gpx.likelihoods.Gaussian(num_datapoints=3, obs_stddev=0.5)
    noisy = lik.predict(dist)
    assert noisy.is_diagonal
    torch.testing.assert_close(noisy.variance, var.to(noisy.variance.device) + 0.25)
    dense = lik.predict(GaussianDistribution(mu, torch.diag(var)))
    assert not dense.is_diagonal
    torch.testing.assert_close(dense.variance.cpu(), noisy.variance.cpu())


def test_distribution_rejects_bad_scale():
    with pytest.raises(ValueError, match="scale must be"):
        GaussianDistribution(torch.zeros(2), torch.zeros(2, 2, 2))
