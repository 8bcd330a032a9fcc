"""The dense linear-algebra entry points behind gpjax/linalg/operations.py
(gpk_potrf / gpk_trsm / gpk_potri / gpk_logdet_tri) and GaussianDistribution's
log_prob / entropy (gpk_mvn_log_prob), against SciPy's LAPACK-backed routines
(scipy.linalg.cholesky / solve_triangular / inv, scipy.stats.multivariate_normal)
on the same float64 inputs.  Sizes cover n = 1, ragged sizes around the
128-row tiles and several recursion levels; FP64 results within 1e-12
relative of SciPy on well-conditioned matrices, the mixed factor within
float32 accuracy."""

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.stats as st
import torch

import gpjax_amd as gpx
from gpjax_amd import linalg as LA
from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu

SIZES = [1, 5, 128, 257, 700, 1500]


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-2, 2, (n, 3))
    return O.gram("matern52", X, np.array([0.7, 1.1, 1.6]), 1.3) + 0.2 * np.eye(n)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n", SIZES)
def test_potrf_and_logdet_match_lapack(cuda, n):
    A = _spd(n, n)
    At = torch.tensor(A, device=cuda)
    L, ld, info = LA.cholesky_with_info(At)
    L_ref = sla.cholesky(A, lower=True)
    assert int(info) == 0
    assert _rel(L.cpu().numpy(), L_ref) <= 1e-13
    assert np.allclose(np.triu(L.cpu().numpy(), 1), 0.0)
    _, ld_ref = np.linalg.slogdet(A)
    assert abs(float(ld) - ld_ref) <= 1e-12 * max(1.0, abs(ld_ref))
    assert abs(float(LA.logdet_triangular(L)) - 0.5 * ld_ref) <= 1e-12 * max(1.0, abs(ld_ref))


@pytest.mark.parametrize("n", [257, 1500])
def test_potrf_mixed_factor(cuda, n):
    A = _spd(n, 3 * n)
    L = LA.lower_cholesky(torch.tensor(A, device=cuda), precision="mixed").cpu().numpy()
    assert _rel(L, sla.cholesky(A, lower=True)) <= 1e-5


def test_potrf_symmetrises_its_input(cuda):
    """jnp.linalg.cholesky factors (A + A^T) / 2."""
    A = _spd(300, 9)
    rng = np.random.default_rng(1)
    skew = rng.standard_normal(A.shape) * 1e-3
    skew = skew - skew.T
    L = LA.lower_cholesky(torch.tensor(A + skew, device=cuda)).cpu().numpy()
    assert _rel(L, sla.cholesky(A, lower=True)) <= 1e-13


def test_potrf_not_positive_definite(cuda):
    """A negative pivot at row 200: info = 201, NaN log-determinant."""
    n = 300
    A = np.eye(n)
    A[200, 200] = -1.0
    L, ld, info = LA.cholesky_with_info(torch.tensor(A, device=cuda))
    assert int(info) == 201
    assert np.isnan(float(ld))


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("m", [1, 7])
def test_trsm_matches_solve_triangular(cuda, n, m):
    L = sla.cholesky(_spd(n, n + 1), lower=True)
    B = np.random.default_rng(n + m).standard_normal((n, m))
    Lt = torch.tensor(L, device=cuda)
    x = LA.solve_triangular(Lt, torch.tensor(B, device=cuda)).cpu().numpy()
    assert _rel(x, sla.solve_triangular(L, B, lower=True)) <= 1e-12
    x1 = LA.solve_triangular(Lt, torch.tensor(B[:, 0], device=cuda)).cpu().numpy()
    assert x1.shape == (n,)
    assert _rel(x1, sla.solve_triangular(L, B[:, 0], lower=True)) <= 1e-12


@pytest.mark.parametrize("n", SIZES)
def test_potri_matches_inverse(cuda, n):
    A = _spd(n, 2 * n + 5)
    L = sla.cholesky(A, lower=True)
    inv = LA.inverse_from_cholesky(torch.tensor(L, device=cuda)).cpu().numpy()
    ref = sla.inv(A)
    assert _rel(inv, ref) <= 1e-12
    assert np.array_equal(inv, inv.T)


@pytest.mark.parametrize("n", [1, 64, 300, 1000])
def test_gaussian_log_prob_and_entropy(cuda, n):
    rng = np.random.default_rng(n)
    S = _spd(n, n + 11)
    mu = rng.standard_normal(n)
    y = mu + rng.standard_normal(n)
    d = gpx.distributions.GaussianDistribution(torch.tensor(mu, device=cuda),
                                               torch.tensor(S, device=cuda))
    mvn = st.multivariate_normal(mu, S)
    lp = float(d.log_prob(torch.tensor(y, device=cuda)))
    assert abs(lp - mvn.logpdf(y)) <= 1e-11 * abs(mvn.logpdf(y))
    assert abs(float(d.entropy()) - mvn.entropy()) <= 1e-11 * abs(mvn.entropy())
    var = np.diag(S).copy()
    dd = gpx.distributions.GaussianDistribution(torch.tensor(mu, device=cuda),
                                                torch.tensor(var, device=cuda))
    ref = st.norm(mu, np.sqrt(var)).logpdf(y).sum()
    assert abs(float(dd.log_prob(torch.tensor(y, device=cuda))) - ref) <= 1e-12 * abs(ref)


def test_predictive_log_prob_scores_held_out_data(cuda):
    """likelihood(posterior.predict(t, D)).log_prob(y_t), the way GPJax scores
    held-out data, against the oracle's predictive scored by SciPy."""
    rng = np.random.default_rng(4)
    X = rng.uniform(-2, 2, (400, 2))
    f = lambda Z: np.sin(2 * Z[:, 0]) * np.cos(Z[:, 1])
    y = f(X) + 0.1 * rng.standard_normal(400)
    T = rng.uniform(-2, 2, (90, 2))
    yt = f(T) + 0.1 * rng.standard_normal(90)
    ls, var, sd = np.array([0.8, 1.2]), 1.1, 0.1
    k = gpx.kernels.RBF(lengthscale=ls, variance=var)
    post = gpx.gps.Prior(mean_function=gpx.mean_functions.Zero(), kernel=k) * \
        gpx.likelihoods.Gaussian(num_datapoints=400, obs_stddev=sd)
    D = gpx.Dataset(X=torch.tensor(X, device=cuda), y=torch.tensor(y[:, None], device=cuda))
    pred = post.likelihood.predict(post.predict(torch.tensor(T, device=cuda), D,
                                                precision="fp64"))
    lp = float(pred.log_prob(torch.tensor(yt, device=cuda)))
    mu_ref, cov_ref = O.predict("rbf", X, y, T, ls, var, sd, 1e-6, 1e-6, 0.0)
    ref = st.multivariate_normal(mu_ref, cov_ref + sd ** 2 * np.eye(90)).logpdf(yt)
    assert abs(lp - ref) <= 1e-9 * abs(ref)
