"""GPU parity of ConjugatePosterior.predict (gpk_predict) against the float64 oracle.

Cases: the N=2,000, D=1 RBF workload with 500 test points on [-3.5, 3.5]
(``workloads.exact_small``), and smaller Matern ARD problems.  FP64 path:
mean within 1e-9 relative (norm-wise), covariance within 1e-9 of the largest
entry; mixed path: mean within 1e-6, covariance within 1e-5.
"""

import numpy as np
import pytest
import torch

import gpjax_amd as gpx
from gpjax_amd import workloads
from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu

KCLS = {"rbf": gpx.kernels.RBF, "matern12": gpx.kernels.Matern12,
        "matern32": gpx.kernels.Matern32, "matern52": gpx.kernels.Matern52}


def _post(kind, ls, var, sd, n, mean=0.0):
    k = KCLS[kind](lengthscale=ls, variance=var)
    mf = gpx.mean_functions.Constant(mean)
    return gpx.gps.Prior(mean_function=mf, kernel=k) * \
        gpx.likelihoods.Gaussian(num_datapoints=n, obs_stddev=sd)


def _check(cuda, kind, X, y, T, ls, var, sd, mean, precision, tol_mu, tol_cov):
    post = _post(kind, ls, var, sd, X.shape[0], mean)
    D = gpx.Dataset(X=torch.tensor(X, device=cuda), y=torch.tensor(y[:, None], device=cuda))
    Tt = torch.tensor(T, device=cuda)
    dist = post.predict(Tt, D, precision=precision)
    mu_ref, cov_ref = O.predict(kind, X, y, T, ls, var, sd, 1e-6, 1e-6, mean)
    mu = dist.mean.cpu().numpy()
    cov = dist.covariance().cpu().numpy()
    assert np.linalg.norm(mu - mu_ref) <= tol_mu * np.linalg.norm(mu_ref)
    assert np.max(np.abs(cov - cov_ref)) <= tol_cov * np.max(np.abs(cov_ref))
    diag = post.predict(Tt, D, full_covariance=False, precision=precision).variance.cpu().numpy()
    np.testing.assert_allclose(diag, np.diag(cov), rtol=0, atol=tol_cov * np.max(np.abs(cov_ref)))
    noisy = post.likelihood.predict(dist)
    np.testing.assert_allclose(torch.diagonal(noisy.covariance()).cpu().numpy(),
                               np.diag(cov) + sd ** 2, rtol=1e-14)


def test_predict_exact_small_workload(cuda):
    w = workloads.exact_small()
    _check(cuda, w["kind"], w["X"], w["y"], w["T"], w["lengthscale"], w["variance"],
           w["obs_stddev"], 0.0, "fp64", 1e-9, 1e-9)


@pytest.mark.parametrize("d", [24, 48])
def test_predict_wide_inputs(cuda, d):
    """The cross-covariance Kxt goes through the streaming builder for d <= 32
    and the general one beyond (both float64, same entries): FP64 parity at
    D = 24 and 48."""
    rng = np.random.default_rng(d)
    X = rng.uniform(0, 1, (400, d))
    y = np.sin(X[:, :4].sum(1)) + 0.05 * rng.standard_normal(400)
    T = rng.uniform(0, 1, (90, d))
    ls = rng.uniform(0.8, 2.0, d)
    _check(cuda, "matern52", X, y, T, ls, 1.1, 0.2, 0.1, "fp64", 1e-9, 1e-9)


@pytest.mark.parametrize("kind", ["matern12", "matern32", "matern52", "rbf"])
def test_predict_ard(cuda, kind):
    rng = np.random.default_rng(5)
    X = rng.uniform(-2, 2, (333, 3))
    y = np.sin(X[:, 0]) * np.cos(X[:, 1]) + 0.05 * rng.standard_normal(333)
    T = rng.uniform(-2.5, 2.5, (77, 3))
    _check(cuda, kind, X, y, T, np.array([0.8, 1.3, 2.0]), 1.4, 0.2, 0.3, "fp64", 1e-9, 1e-9)


def test_predict_mixed(cuda):
    """Mixed factorisation on the benchmark recipe (D = 16): mean 1e-6, cov 1e-5."""
    w = workloads.exact_large(n=1500)
    T = np.random.default_rng(9).standard_normal((64, 16))
    _check(cuda, w["kind"], w["X"], w["y"], T, w["lengthscale"], w["variance"],
           w["obs_stddev"], 0.0, "mixed", 1e-6, 1e-5)


def test_predict_mixed_small_1d_in_gate_or_flagged(cuda, monkeypatch):
    """The 1-D, 2,000-point workload is nearly low-rank.  With the check's action
    off, the raw mixed predictive is within 1e-6 (mean) / 1e-5 (covariance) of the
    oracle or its factorisation is flagged; with the default action the
    predictive comes back within those bounds either way (here: recomputed in
    float64, within 1e-9)."""
    w = workloads.exact_small()
    X, y, T = w["X"], w["y"], w["T"]
    ls, var, sd = w["lengthscale"], w["variance"], w["obs_stddev"]
    mu_ref, cov_ref = O.predict("rbf", X, y, T, ls, var, sd, 1e-6, 1e-6, 0.0)
    D = gpx.Dataset(X=torch.tensor(X, device=cuda), y=torch.tensor(y[:, None], device=cuda))
    Tt = torch.tensor(T, device=cuda)

    def errs(dist):
        mu, cov = dist.mean.cpu().numpy(), dist.covariance().cpu().numpy()
        return (np.linalg.norm(mu - mu_ref) / np.linalg.norm(mu_ref),
                np.max(np.abs(cov - cov_ref)) / np.max(np.abs(cov_ref)))

    monkeypatch.setattr(gpx.config, "mixed_check", "off")
    dist = _post("rbf", ls, var, sd, X.shape[0]).predict(Tt, D, precision="mixed")
    e_mu, e_cov = errs(dist)
    flagged = float(dist.factor_result[7]) == 1.0
    print(f"mixed predict: mean {e_mu:.2e} cov {e_cov:.2e} flagged={flagged}")
    assert flagged or (e_mu <= 1e-6 and e_cov <= 1e-5)
    monkeypatch.setattr(gpx.config, "mixed_check", "fallback")
    if flagged:
        with pytest.warns(RuntimeWarning):
            dist = _post("rbf", ls, var, sd, X.shape[0]).predict(Tt, D, precision="mixed")
        assert float(dist.factor_result[7]) == 2.0
        tol_mu, tol_cov = 1e-9, 1e-9
    else:
        dist = _post("rbf", ls, var, sd, X.shape[0]).predict(Tt, D, precision="mixed")
        tol_mu, tol_cov = 1e-6, 1e-5
    e_mu, e_cov = errs(dist)
    assert e_mu <= tol_mu and e_cov <= tol_cov, (e_mu, e_cov)


def test_predict_diagonal_streams_large_t(cuda):
    """The diagonal predictive streams the test points in passes (512 MB per
    buffer: 4,096 points per pass at N = 16,384).  Every test point's mean and
    variance depend only on its own column, so a slice that straddles pass
    boundaries, predicted in a call of its own, must match exactly; the first
    points also match the full-covariance path's diagonal."""
    rng = np.random.default_rng(21)
    n, d, t = 16384, 3, 10000
    X = rng.uniform(-3, 3, (n, d))
    y = np.sin(X[:, 0]) * np.cos(X[:, 1]) + 0.1 * rng.standard_normal(n)
    T = rng.uniform(-3.5, 3.5, (t, d))
    post = _post("matern32", np.array([0.9, 1.4, 2.0]), 1.1, 0.1, n, 0.2)
    D = gpx.Dataset(X=torch.tensor(X, device=cuda), y=torch.tensor(y[:, None], device=cuda))
    big = post.predict(torch.tensor(T, device=cuda), D, full_covariance=False, precision="fp64")
    mu, var = big.mean.cpu().numpy(), big.variance.cpu().numpy()
    assert np.all(np.isfinite(mu)) and np.all(var > 0)
    part = post.predict(torch.tensor(T[3000:9000], device=cuda), D, full_covariance=False,
                        precision="fp64")
    np.testing.assert_array_equal(part.mean.cpu().numpy(), mu[3000:9000])
    np.testing.assert_array_equal(part.variance.cpu().numpy(), var[3000:9000])
    full = post.predict(torch.tensor(T[:200], device=cuda), D, precision="fp64")
    np.testing.assert_allclose(full.mean.cpu().numpy(), mu[:200], rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.diag(full.covariance().cpu().numpy()), var[:200], rtol=0,
                               atol=1e-12)
