"""Pins of the float64 oracle (oracle/gp_oracle.py) -- CPU only.

GPJax cannot be imported here, so each oracle function is checked against
something that does not share its code path:
  * kernel values: scalar closed forms via ``math``, K(x, x) = sigma_f^2,
    ARD == isotropic for equal lengthscales;
  * Gram: pure-Python double loop on tiny inputs, symmetry, PSD after +1e-6
    (``tests/test_kernels/test_stationary.py:186-204`` in the reference);
  * MLL: ``scipy.stats.multivariate_normal.logpdf``, the N = 1 closed form,
    LU vs Cholesky log-determinants;
  * gradient: torch float64 autograd of an independent torch implementation
    of the objective (the stand-in for jax.grad), and central differences of
    the oracle MLL;
  * collapsed ELBO: the reference's invariant Z = X => ELBO == MLL within 1e-6
    relative (``tests/test_objectives.py:193-199``) and a dense evaluation
    log N(y; m, s^2 I + Q) - tr(K - Q) / (2 s^2);
  * predict: dense solve-based formulas, exact interpolation as noise -> 0, and
    likelihood-weighted Monte Carlo over joint prior draws (no solve at all);
  * collapsed predict: Z = X gives the exact predictive, Titsias' dense form;
  * LOOCV: brute-force leave-one-out conditioning (N - 1 point solves per
    held-out point), the N = 1 closed form; its gradient against torch
    autograd of that brute force and central differences.
A planted-mistake test shows the gradient pins reject a dropped factor.  The
committed fixtures in tests/golden/ must still be what the oracle computes.
"""

import math

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.stats as st
import torch

from oracle import gp_oracle as O

KINDS = O.KINDS


def _scalar_kernel(kind, x, z, ls, var):
    r2 = sum(((a / l) - (b / l)) ** 2 for a, b, l in zip(x, z, ls))
    if kind == "rbf":
        return var * math.exp(-0.5 * r2)
    r = math.sqrt(max(r2, 1e-36))
    if kind == "matern12":
        return var * math.exp(-r)
    if kind == "matern32":
        return var * (1 + math.sqrt(3) * r) * math.exp(-math.sqrt(3) * r)
    return var * (1 + math.sqrt(5) * r + 5.0 / 3.0 * r * r) * math.exp(-math.sqrt(5) * r)


@pytest.mark.parametrize("kind", KINDS)
def test_kernel_closed_forms(kind):
    rng = np.random.default_rng(1)
    X = rng.normal(size=(6, 3))
    Z = rng.normal(size=(5, 3))
    ls = np.array([0.5, 1.0, 2.0])
    K = O.cross_covariance(kind, X, Z, ls, 1.7)
    for i in range(6):
        for j in range(5):
            assert K[i, j] == pytest.approx(_scalar_kernel(kind, X[i], Z[j], ls, 1.7),
                                            rel=1e-14, abs=1e-300)
    # k(x, x) = variance exactly
    np.testing.assert_array_equal(np.diag(O.gram(kind, X, ls, 1.7)), np.full(6, 1.7))
    np.testing.assert_array_equal(O.kernel_diagonal(kind, X, ls, 1.7), np.full(6, 1.7))
    # ARD with equal lengthscales == isotropic
    np.testing.assert_allclose(O.gram(kind, X, np.full(3, 0.8), 1.0), O.gram(kind, X, 0.8, 1.0),
                               rtol=0, atol=0)
    # far apart -> 0
    far = O.cross_covariance(kind, np.zeros((1, 1)), np.full((1, 1), 1e3), 1.0, 1.0)
    assert far[0, 0] < 1e-100


def test_rbf_1d_math_exp():
    x = np.linspace(-3, 3, 7)[:, None]
    K = O.gram("rbf", x, 0.5, 1.0)
    for i in range(7):
        for j in range(7):
            assert K[i, j] == pytest.approx(math.exp(-0.5 * ((x[i, 0] - x[j, 0]) / 0.5) ** 2),
                                            rel=1e-15)


@pytest.mark.parametrize("kind", KINDS)
def test_gram_brute_force_symmetric_psd(kind):
    rng = np.random.default_rng(2)
    X = rng.uniform(-2, 2, size=(9, 2))
    ls = np.array([0.7, 1.4])
    K = O.gram(kind, X, ls, 1.3)
    B = np.array([[_scalar_kernel(kind, X[i], X[j], ls, 1.3) for j in range(9)] for i in range(9)])
    np.testing.assert_allclose(K, B, rtol=1e-14, atol=0)
    np.testing.assert_array_equal(K, K.T)
    assert np.linalg.eigvalsh(O.add_jitter(K, 1e-6)).min() > 0


def test_add_jitter_errors():
    with pytest.raises(ValueError, match="Expected 2D matrix"):
        O.add_jitter(np.ones(3), 1e-6)
    with pytest.raises(ValueError, match="Expected square matrix"):
        O.add_jitter(np.ones((2, 3)), 1e-6)
    np.testing.assert_array_equal(O.add_jitter(np.zeros((2, 2)), 0.5), 0.5 * np.eye(2))


def _data(n, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-2, 2, (n, d))
    y = np.sin(X[:, 0]) + 0.1 * rng.standard_normal(n)
    return X, y


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n,d", [(1, 1), (2, 2), (10, 3), (64, 1)])
def test_mll_vs_scipy_logpdf(kind, n, d):
    X, y = _data(n, d, 10 * n + d)
    ls, var, sd, c = np.linspace(0.6, 1.6, d), 1.4, 0.3, 0.25
    S = O.sigma_matrix(kind, X, ls, var, sd, 1e-6)
    ref = st.multivariate_normal(mean=np.full(n, c), cov=S).logpdf(y)
    assert O.conjugate_mll(kind, X, y, ls, var, sd, 1e-6, c) == pytest.approx(ref, rel=1e-11)


def test_mll_n1_closed_form():
    X, y = np.array([[0.3]]), np.array([1.7])
    var, sd, jit, c = 1.2, 0.4, 1e-6, 0.5
    s = var + jit + sd ** 2
    ref = -0.5 * (math.log(2 * math.pi) + math.log(s) + (y[0] - c) ** 2 / s)
    assert O.conjugate_mll("rbf", X, y, 1.0, var, sd, jit, c) == pytest.approx(ref, rel=1e-15)


def test_lu_logdet_matches_cholesky():
    X, y = _data(120, 3, 5)
    S = O.sigma_matrix("matern52", X, 0.9, 1.0, 0.2)
    parts = O.conjugate_mll_parts("matern52", X, y, 0.9, 1.0, 0.2)
    L = np.linalg.cholesky(S)
    assert parts["logdet"] == pytest.approx(2 * np.sum(np.log(np.diag(L))), rel=1e-12)
    np.testing.assert_allclose(parts["alpha"], sla.cho_solve((L, True), y), rtol=1e-9)


def _torch_mll(kind, X, y, ls, var, sd, c, jitter=1e-6):
    """Independent torch float64 MLL (slogdet + solve), differentiated by autograd."""
    xs = X / ls
    diff = xs[:, None, :] - xs[None, :, :]
    d2 = (diff * diff).sum(-1)
    if kind == "rbf":
        K = var * torch.exp(-0.5 * d2)
    else:
        r = torch.sqrt(torch.clamp(d2, min=1e-36))
        if kind == "matern12":
            K = var * torch.exp(-r)
        elif kind == "matern32":
            K = var * (1 + math.sqrt(3) * r) * torch.exp(-math.sqrt(3) * r)
        else:
            K = var * (1 + math.sqrt(5) * r + 5.0 / 3.0 * r * r) * torch.exp(-math.sqrt(5) * r)
    n = X.shape[0]
    S = K + torch.eye(n, dtype=torch.float64) * jitter + torch.eye(n, dtype=torch.float64) * sd ** 2
    r_ = y - c
    _, logdet = torch.linalg.slogdet(S)
    return -0.5 * (n * math.log(2 * math.pi) + logdet + r_ @ torch.linalg.solve(S, r_))


def _grad_vec(g, ard):
    return np.concatenate([np.atleast_1d(g["lengthscale"]),
                           [g["variance"], g["obs_stddev"], g["mean"]]])


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("ard", [False, True])
def test_grad_vs_torch_autograd(kind, ard):
    X, y = _data(40, 3, 3)
    ls = np.array([0.6, 1.1, 1.9]) if ard else 0.9
    var, sd, c = 1.3, 0.25, 0.1
    g = O.conjugate_mll_grad(kind, X, y, ls, var, sd, 1e-6, c)
    tl = torch.tensor(ls, dtype=torch.float64, requires_grad=True)
    tv = torch.tensor(var, dtype=torch.float64, requires_grad=True)
    ts = torch.tensor(sd, dtype=torch.float64, requires_grad=True)
    tc = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    m = _torch_mll(kind, torch.tensor(X), torch.tensor(y), tl, tv, ts, tc)
    m.backward()
    assert g["mll"] == pytest.approx(m.item(), rel=1e-12)
    ref = np.concatenate([np.atleast_1d(tl.grad.numpy()),
                          [tv.grad.item(), ts.grad.item(), tc.grad.item()]])
    np.testing.assert_allclose(_grad_vec(g, ard), ref, rtol=1e-9, atol=1e-11)


def _fd_grad(kind, X, y, ls, var, sd, c, h=1e-6):
    f = lambda a, b, s, m: O.conjugate_mll(kind, X, y, a, b, s, 1e-6, m)
    ls = np.atleast_1d(np.asarray(ls, float))
    out = []
    for i in range(ls.size):
        e = np.zeros_like(ls)
        e[i] = h
        out.append((f(ls + e, var, sd, c) - f(ls - e, var, sd, c)) / (2 * h))
    out.append((f(ls, var + h, sd, c) - f(ls, var - h, sd, c)) / (2 * h))
    out.append((f(ls, var, sd + h, c) - f(ls, var, sd - h, c)) / (2 * h))
    out.append((f(ls, var, sd, c + h) - f(ls, var, sd, c - h)) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("kind", KINDS)
def test_grad_vs_central_differences(kind):
    X, y = _data(30, 2, 4)
    ls, var, sd, c = np.array([0.8, 1.4]), 1.1, 0.3, -0.2
    g = O.conjugate_mll_grad(kind, X, y, ls, var, sd, 1e-6, c)
    np.testing.assert_allclose(_grad_vec(g, True), _fd_grad(kind, X, y, ls, var, sd, c),
                               rtol=2e-6, atol=1e-7)


def test_planted_mistake_is_caught(monkeypatch):
    """Drop the 5/3 factor of the Matern-5/2 derivative: the FD pin must fail."""
    good = O.unit_kernel_dlog_factor

    def wrong(kind, d2):
        f = good(kind, d2)
        return f * 0.6 if kind == "matern52" else f

    X, y = _data(30, 2, 4)
    ls, var, sd, c = np.array([0.8, 1.4]), 1.1, 0.3, -0.2
    fd = _fd_grad("matern52", X, y, ls, var, sd, c)
    monkeypatch.setattr(O, "unit_kernel_dlog_factor", wrong)
    g = O.conjugate_mll_grad("matern52", X, y, ls, var, sd, 1e-6, c)
    assert not np.allclose(_grad_vec(g, True), fd, rtol=2e-6, atol=1e-7)


@pytest.mark.parametrize("kind", KINDS)
def test_collapsed_elbo_equals_mll_when_z_is_x(kind):
    # The reference's invariant (tests/test_objectives.py:193-199).
    X, y = _data(20, 2, 6)
    mll = O.conjugate_mll(kind, X, y, 1.0, 1.0, 1.0, 1e-6, 0.0)
    elbo = O.collapsed_elbo(kind, X, y, X, 1.0, 1.0, 1.0, 1e-6, 0.0)
    assert abs(elbo - mll) / abs(mll) < 1e-6


@pytest.mark.parametrize("kind", KINDS)
def test_collapsed_elbo_dense_form(kind):
    X, y = _data(60, 2, 8)
    rng = np.random.default_rng(9)
    Z = rng.uniform(-2, 2, (15, 2))
    ls, var, sd, c, jit = np.array([0.7, 1.3]), 1.2, 0.35, 0.1, 1e-6
    Kzz = O.gram(kind, Z, ls, var) + jit * np.eye(15)
    Kzx = O.cross_covariance(kind, Z, X, ls, var)
    Q = Kzx.T @ np.linalg.solve(Kzz, Kzx)
    Kxx = O.gram(kind, X, ls, var)
    dense = st.multivariate_normal(mean=np.full(60, c), cov=Q + sd ** 2 * np.eye(60)).logpdf(y) \
        - np.trace(Kxx - Q) / (2 * sd ** 2)
    assert O.collapsed_elbo(kind, X, y, Z, ls, var, sd, jit, c) == pytest.approx(dense, rel=1e-8)


@pytest.mark.parametrize("kind", KINDS)
def test_collapsed_elbo_grad_value_and_fd(kind):
    rng = np.random.default_rng(21)
    X = rng.uniform(0, 1, (120, 3))
    y = np.sin(6 * X[:, 0]) + 0.1 * rng.standard_normal(120)
    Z = X[rng.choice(120, 25, replace=False)] + 0.01
    ls, var, sd, c = np.array([0.3, 0.5, 0.4]), 1.3, 0.2, 0.1
    g = O.collapsed_elbo_grad(kind, X, y, Z, ls, var, sd, 1e-6, c)
    f = lambda Z_=Z, l=ls, v=var, s=sd, m=c: O.collapsed_elbo(kind, X, y, Z_, l, v, s, 1e-6, m)
    assert g["elbo"] == pytest.approx(f(), rel=1e-8)
    h = 1e-6
    fd_l = np.array([(f(l=ls + h * e) - f(l=ls - h * e)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(g["lengthscale"], fd_l, rtol=1e-5)
    assert g["variance"] == pytest.approx((f(v=var + h) - f(v=var - h)) / (2 * h), rel=1e-5)
    assert g["obs_stddev"] == pytest.approx((f(s=sd + h) - f(s=sd - h)) / (2 * h), rel=1e-5)
    assert g["mean"] == pytest.approx((f(m=c + h) - f(m=c - h)) / (2 * h), rel=1e-5)
    for (i, d) in [(0, 0), (7, 2), (24, 1)]:
        E = np.zeros_like(Z)
        E[i, d] = h
        assert g["Z"][i, d] == pytest.approx((f(Z_=Z + E) - f(Z_=Z - E)) / (2 * h), rel=1e-4,
                                             abs=1e-6)


def test_collapsed_elbo_grad_vs_torch_autograd():
    """Autograd of an independent torch transcription of objectives.py:342-416."""
    rng = np.random.default_rng(22)
    X = rng.uniform(0, 1, (80, 2))
    y = np.cos(4 * X[:, 1]) + 0.1 * rng.standard_normal(80)
    Z = X[rng.choice(80, 15, replace=False)] + 0.02
    ls, var, sd, c, jit = np.array([0.35, 0.45]), 0.9, 0.25, -0.3, 1e-6
    g = O.collapsed_elbo_grad("rbf", X, y, Z, ls, var, sd, jit, c)
    T = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)
    tl, tv, ts, tc, tZ = T(ls), T(var), T(sd), T(c), T(Z)
    Xt, yt = torch.tensor(X), torch.tensor(y)

    def k(A, B):
        d = A[:, None, :] / tl - B[None, :, :] / tl
        return tv * torch.exp(-0.5 * (d * d).sum(-1))

    noise = ts ** 2
    Kzz = k(tZ, tZ) + jit * torch.eye(15, dtype=torch.float64)
    Kzx = k(tZ, Xt)
    Lz = torch.linalg.cholesky(Kzz)
    A = torch.linalg.solve_triangular(Lz, Kzx, upper=False) / torch.sqrt(noise)
    AAT = A @ A.T
    L = torch.linalg.cholesky(torch.eye(15, dtype=torch.float64) + AAT)
    diff = yt - tc
    LAd = torch.linalg.solve_triangular(L, (A @ diff)[:, None], upper=False)[:, 0]
    quad = ((diff ** 2).sum() - (LAd ** 2).sum()) / noise
    two_lp = -80 * torch.log(2 * math.pi * noise) - 2 * torch.log(torch.diagonal(L)).sum() - quad
    elbo = (two_lp - (80 * tv / noise - torch.trace(AAT))) / 2
    elbo.backward()
    # The lemma-based form and the reference's triangular-solve form differ in
    # rounding by ~1e-9 relative on this ill-conditioned RBF Kzz.
    assert g["elbo"] == pytest.approx(elbo.item(), rel=1e-8)
    np.testing.assert_allclose(g["lengthscale"], tl.grad.numpy(), rtol=1e-6)
    assert g["variance"] == pytest.approx(tv.grad.item(), rel=1e-6)
    assert g["obs_stddev"] == pytest.approx(ts.grad.item(), rel=1e-6)
    assert g["mean"] == pytest.approx(tc.grad.item(), rel=1e-6)
    np.testing.assert_allclose(g["Z"], tZ.grad.numpy(), rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("kind", KINDS)
def test_predict_dense_solve(kind):
    X, y = _data(50, 2, 11)
    T = np.random.default_rng(12).uniform(-2.5, 2.5, (16, 2))
    ls, var, sd, c = np.array([0.9, 1.2]), 1.0, 0.2, 0.3
    mu, cov = O.predict(kind, X, y, T, ls, var, sd, 1e-6, 1e-6, c)
    S = O.sigma_matrix(kind, X, ls, var, sd, 1e-6)
    Kxt = O.cross_covariance(kind, X, T, ls, var)
    mu_ref = c + Kxt.T @ np.linalg.solve(S, y - c)
    cov_ref = O.gram(kind, T, ls, var) - Kxt.T @ np.linalg.solve(S, Kxt) + 1e-6 * np.eye(16)
    np.testing.assert_allclose(mu, mu_ref, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(cov, cov_ref, rtol=1e-8, atol=1e-11)
    noisy = O.gaussian_likelihood_predict(cov, sd)
    np.testing.assert_allclose(np.diag(noisy), np.diag(cov) + sd ** 2, rtol=1e-15)


def test_predict_interpolates_without_noise():
    X, y = _data(25, 1, 13)
    y = np.sin(X[:, 0])
    mu, cov = O.predict("rbf", X, y, X, 0.7, 1.0, 1e-5, 1e-10, 1e-10)
    np.testing.assert_allclose(mu, y, atol=1e-6)
    assert np.max(np.abs(np.diag(cov))) < 1e-6


@pytest.mark.parametrize("kind", ["rbf", "matern32"])
def test_predict_monte_carlo_importance_sampling(kind):
    """Posterior moments without any solve: draw (f, f*) jointly from the prior
    (latent covariance with the posterior's jitter on the training block and the
    prior's jitter on the test block), weight each draw by the likelihood
    N(y; f, obs_sd^2 I), and compare the weighted mean and variance of f* with
    predict (the statistical check of the reference's tests/test_gps.py:283-342,
    restated without pathwise sampling).  Tolerance: 6 Monte-Carlo standard
    errors from the effective sample size (about 1.1e5 of the 1e6 draws);
    dropping the noise from Sigma or the -Ktx Sigma^-1 Kxt term moves some test
    point by 50 or more of them."""
    rng = np.random.default_rng(11)
    X = np.array([[-1.0], [0.2], [1.1]])
    T = np.array([[-0.4], [0.7], [2.0]])
    y = np.array([0.3, -0.5, 0.8])
    ls, var, sd, jp, jq, mean = 0.9, 1.3, 0.5, 1e-6, 1e-6, 0.1
    mu, cov = O.predict(kind, X, y, T, ls, var, sd, jp, jq, mean)
    A = np.vstack([X, T])
    K = O.gram(kind, A, ls, var)
    K[:3, :3] += jp * np.eye(3)
    K[3:, 3:] += jq * np.eye(3)
    Lk = np.linalg.cholesky(K)
    f = mean + rng.standard_normal((1_000_000, 6)) @ Lk.T
    logw = -0.5 * np.sum((y - f[:, :3]) ** 2, axis=1) / sd ** 2
    w = np.exp(logw - logw.max())
    w /= w.sum()
    ess = 1.0 / np.sum(w ** 2)
    m_mc = w @ f[:, 3:]
    v_mc = w @ (f[:, 3:] - m_mc) ** 2
    v = np.diag(cov)
    assert ess > 1e5
    np.testing.assert_array_less(np.abs(m_mc - mu), 6 * np.sqrt(v / ess))
    np.testing.assert_array_less(np.abs(v_mc - v), 6 * v * np.sqrt(2.0 / ess))


@pytest.mark.parametrize("kind", KINDS)
def test_collapsed_predict_equals_exact_when_z_is_x(kind):
    """With Z = X the collapsed posterior is the exact one: Q = K, so mean and
    covariance equal ConjugatePosterior.predict's (up to the jitter on Kzz)."""
    X, y = _data(40, 2, 21)
    T = np.random.default_rng(22).uniform(-2.5, 2.5, (12, 2))
    ls, var, sd, c = np.array([0.9, 1.4]), 1.1, 0.3, -0.2
    mu, cov = O.collapsed_predict(kind, X, y, X, T, ls, var, sd, 1e-10, c)
    mu_ex, cov_ex = O.predict(kind, X, y, T, ls, var, sd, 0.0, 1e-10, c)
    np.testing.assert_allclose(mu, mu_ex, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(cov, cov_ex, rtol=1e-6, atol=1e-8)


def test_collapsed_predict_dense_form():
    """Titsias' predictive in its textbook dense form:
    mean = c + Ktz (Kzz + Kzx Kxz / s)^-1 Kzx (y - c) / s,
    cov = Ktt - Ktz Kzz^-1 Kzt + Ktz (Kzz + Kzx Kxz / s)^-1 Kzt + jitter I."""
    X, y = _data(60, 3, 23)
    Z = X[:17] + 0.05
    T = np.random.default_rng(24).uniform(-2, 2, (9, 3))
    ls, var, sd, c, jit = np.array([0.8, 1.1, 1.5]), 0.9, 0.25, 0.1, 1e-6
    mu, cov = O.collapsed_predict("matern32", X, y, Z, T, ls, var, sd, jit, c)
    s = sd ** 2
    Kzz = O.gram("matern32", Z, ls, var) + jit * np.eye(17)
    Kzx = O.cross_covariance("matern32", Z, X, ls, var)
    Kzt = O.cross_covariance("matern32", Z, T, ls, var)
    Sig = Kzz + Kzx @ Kzx.T / s
    mu_ref = c + Kzt.T @ np.linalg.solve(Sig, Kzx @ (y - c)) / s
    cov_ref = (O.gram("matern32", T, ls, var) - Kzt.T @ np.linalg.solve(Kzz, Kzt)
               + Kzt.T @ np.linalg.solve(Sig, Kzt) + jit * np.eye(9))
    np.testing.assert_allclose(mu, mu_ref, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(cov, cov_ref, rtol=1e-7, atol=1e-10)


def _brute_loocv(kind, X, y, ls, var, sd, c, jitter=1e-6):
    """Leave each point out, condition the remaining N-1 on the noisy Gaussian,
    and score the held-out y_i under its noisy predictive (Rasmussen & Williams
    2006, eqs. 5.10-5.12 before their closed form)."""
    n = X.shape[0]
    S = O.gram(kind, X, ls, var) + np.eye(n) * (sd ** 2 + jitter)
    tot = 0.0
    for i in range(n):
        o = np.r_[0:i, i + 1:n]
        if o.size:
            sol = np.linalg.solve(S[np.ix_(o, o)], np.c_[y[o] - c, S[o, i]])
            mu = c + S[i, o] @ sol[:, 0]
            v = S[i, i] - S[i, o] @ sol[:, 1]
        else:
            mu, v = c, S[i, i]
        tot += st.norm(mu, math.sqrt(v)).logpdf(y[i])
    return tot


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n,d", [(1, 1), (2, 2), (10, 3), (33, 2)])
def test_loocv_vs_brute_force(kind, n, d):
    X, y = _data(n, d, 7 * n + d)
    ls, var, sd, c = np.linspace(0.7, 1.5, d), 1.2, 0.3, 0.2
    ref = _brute_loocv(kind, X, y, ls, var, sd, c)
    assert O.conjugate_loocv(kind, X, y, ls, var, sd, 1e-6, c) == pytest.approx(ref, rel=1e-10)


def test_loocv_n1_closed_form():
    """One point: the held-out predictive is the prior, N(c, var + jitter + sd^2)."""
    X, y = np.array([[0.3]]), np.array([1.7])
    var, sd, jit, c = 1.2, 0.4, 1e-6, 0.5
    s = var + jit + sd ** 2
    ref = -0.5 * (math.log(2 * math.pi) + math.log(s) + (y[0] - c) ** 2 / s)
    assert O.conjugate_loocv("rbf", X, y, 1.0, var, sd, jit, c) == pytest.approx(ref, rel=1e-14)


def _torch_brute_loocv(kind, X, y, ls, var, sd, c, jitter=1e-6):
    xs = X / ls
    diff = xs[:, None, :] - xs[None, :, :]
    d2 = (diff * diff).sum(-1)
    if kind == "rbf":
        K = var * torch.exp(-0.5 * d2)
    else:
        r = torch.sqrt(torch.clamp(d2, min=1e-36))
        if kind == "matern12":
            K = var * torch.exp(-r)
        elif kind == "matern32":
            K = var * (1 + math.sqrt(3) * r) * torch.exp(-math.sqrt(3) * r)
        else:
            K = var * (1 + math.sqrt(5) * r + 5.0 / 3.0 * r * r) * torch.exp(-math.sqrt(5) * r)
    n = X.shape[0]
    S = K + torch.eye(n, dtype=torch.float64) * (sd ** 2 + jitter)
    tot = 0.0
    for i in range(n):
        o = [k for k in range(n) if k != i]
        Soo = S[o][:, o]
        mu = c + S[i, o] @ torch.linalg.solve(Soo, y[o] - c)
        v = S[i, i] - S[i, o] @ torch.linalg.solve(Soo, S[o, i])
        tot = tot - 0.5 * (math.log(2 * math.pi) + torch.log(v) + (y[i] - mu) ** 2 / v)
    return tot


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("ard", [False, True])
def test_loocv_grad_vs_torch_autograd_of_brute_force(kind, ard):
    X, y = _data(16, 3, 8)
    ls = np.array([0.6, 1.1, 1.9]) if ard else 0.9
    var, sd, c = 1.3, 0.25, 0.1
    g = O.conjugate_loocv_grad(kind, X, y, ls, var, sd, 1e-6, c)
    tl = torch.tensor(ls, dtype=torch.float64, requires_grad=True)
    tv = torch.tensor(var, dtype=torch.float64, requires_grad=True)
    ts = torch.tensor(sd, dtype=torch.float64, requires_grad=True)
    tc = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    v = _torch_brute_loocv(kind, torch.tensor(X), torch.tensor(y), tl, tv, ts, tc)
    v.backward()
    assert g["loocv"] == pytest.approx(v.item(), rel=1e-10)
    ref = np.concatenate([np.atleast_1d(tl.grad.numpy()),
                          [tv.grad.item(), ts.grad.item(), tc.grad.item()]])
    np.testing.assert_allclose(_grad_vec(g, ard), ref, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("kind", KINDS)
def test_loocv_grad_vs_central_differences(kind):
    X, y = _data(30, 2, 9)
    ls, var, sd, c = np.array([0.8, 1.4]), 1.1, 0.3, -0.2
    g = O.conjugate_loocv_grad(kind, X, y, ls, var, sd, 1e-6, c)
    f = lambda a, b, s, m: O.conjugate_loocv(kind, X, y, a, b, s, 1e-6, m)
    h = 1e-6
    fd = []
    for i in range(2):
        e = np.zeros(2)
        e[i] = h
        fd.append((f(ls + e, var, sd, c) - f(ls - e, var, sd, c)) / (2 * h))
    fd.append((f(ls, var + h, sd, c) - f(ls, var - h, sd, c)) / (2 * h))
    fd.append((f(ls, var, sd + h, c) - f(ls, var, sd - h, c)) / (2 * h))
    fd.append((f(ls, var, sd, c + h) - f(ls, var, sd, c - h)) / (2 * h))
    np.testing.assert_allclose(_grad_vec(g, True), np.array(fd), rtol=2e-6, atol=1e-7)
    assert g["loocv"] == pytest.approx(f(ls, var, sd, c), rel=1e-12)


def test_oracle_reproduces_golden_fixtures():
    """tests/golden/exact_small.npz (written by tests/golden/make_fixtures.py
    from the oracle) is what the device path is checked against without the
    oracle; this keeps the two in step: any change to the oracle's answers on
    those 120 problems must come with regenerated fixtures."""
    import importlib.util
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location(
        "make_fixtures", os.path.join(here, "golden", "make_fixtures.py"))
    F = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(F)
    fresh = F.compute(with_spread=False)
    z = np.load(os.path.join(here, "golden", "exact_small.npz"))
    assert sorted(k for k in z.files if "_elbospread_" not in k) == sorted(fresh)
    for key in fresh:
        if key == "meta":
            assert str(z[key]) == str(fresh[key])
        else:
            np.testing.assert_allclose(z[key], fresh[key], rtol=1e-11, atol=1e-13, err_msg=key)
