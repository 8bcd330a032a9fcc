"""Float64 CPU restatement of GPJax's exact-GP and collapsed-sparse-GP path.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP library in
``gpjax_amd``: it may be imported by ``tests/``, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- never by the product path.
The product path fails loudly when its HIP library is missing; it never falls
back to anything in here.

Every function restates the reference algorithm in NumPy/SciPy float64 and
cites the GPJax 0.13.2 file:line it follows (paths relative to the reference
repository root).  The reference itself cannot be imported (it needs jax,
flax, numpyro and Python >= 3.11), so each function is pinned in
``tests/test_oracle.py`` by closed forms, SciPy/LAPACK routines, brute force
on tiny inputs, torch float64 autograd of the objective and central finite
differences, and by the reference's own invariant (Z = X gives ELBO == MLL,
``tests/test_objectives.py:193-199``).

Conventions
-----------
* ``kind`` is one of ``"rbf" | "matern12" | "matern32" | "matern52"``.
* ``ls`` is a float (isotropic) or a ``[D]`` array (ARD); ``var`` is the
  kernel variance sigma_f^2; ``obs_sd`` is the likelihood's ``obs_stddev``
  (a standard deviation, ``gpjax/likelihoods.py:135,148-150``).
* Gradients are in constrained coordinates, i.e. w.r.t. ``ls``, ``var``,
  ``obs_sd`` and the constant mean, as ``jax.grad`` over the constrained
  ``nnx.State`` returns them (``tests/test_objectives.py:68-85``).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

KINDS = ("rbf", "matern12", "matern32", "matern52")
LOG2PI = math.log(2.0 * math.pi)
_SQRT3 = math.sqrt(3.0)
_SQRT5 = math.sqrt(5.0)
# Row-chunk for pairwise differences so that [chunk, M, D] stays ~64 MB.
_PAIR_BUDGET = 8_000_000


def _as_ls(ls, D: int) -> np.ndarray:
    a = np.asarray(ls, dtype=np.float64)
    if a.ndim == 0:
        return np.full(D, float(a))
    if a.shape != (D,):
        raise ValueError(f"lengthscale shape {a.shape} incompatible with D={D}")
    return a


def squared_distance(xs: np.ndarray, zs: np.ndarray) -> np.ndarray:
    """sum_d (x_d - z_d)^2 for every row pair, by direct differences.

    Follows ``gpjax/kernels/stationary/utils.py:42-53`` (``jnp.sum((x - y)**2)``)
    applied through the double ``vmap`` of
    ``gpjax/kernels/computations/dense.py:32-36``.  Inputs are already divided
    by the lengthscale (``rbf.py:41-42``).
    """
    n, m = xs.shape[0], zs.shape[0]
    out = np.empty((n, m), dtype=np.float64)
    step = max(1, _PAIR_BUDGET // max(1, m * xs.shape[1]))
    for i0 in range(0, n, step):
        diff = xs[i0:i0 + step, None, :] - zs[None, :, :]
        out[i0:i0 + step] = np.einsum("ijd,ijd->ij", diff, diff)
    return out


def unit_kernel_from_d2(kind: str, d2: np.ndarray) -> np.ndarray:
    """k(rho) with sigma_f^2 = 1, from rho^2 (already lengthscale-scaled).

    RBF ``stationary/rbf.py:43``; Matern-1/2 ``matern12.py:47``;
    Matern-3/2 ``matern32.py:48-53``; Matern-5/2 ``matern52.py:47-52``; the
    Matern distance is ``sqrt(max(d2, 1e-36))`` (``stationary/utils.py:67``).
    """
    if kind == "rbf":
        return np.exp(-0.5 * d2)
    tau = np.sqrt(np.maximum(d2, 1e-36))
    if kind == "matern12":
        return np.exp(-tau)
    if kind == "matern32":
        return (1.0 + _SQRT3 * tau) * np.exp(-_SQRT3 * tau)
    if kind == "matern52":
        return (1.0 + _SQRT5 * tau + 5.0 / 3.0 * tau * tau) * np.exp(-_SQRT5 * tau)
    raise ValueError(f"unknown kernel kind {kind!r}")


def unit_kernel_dlog_factor(kind: str, d2: np.ndarray) -> np.ndarray:
    """F(rho) such that dK/dl_d = sigma_f^2 * F(rho) * u_d / l_d, u_d = (x_d-z_d)^2/l_d^2.

    Derived from the formulas cited in :func:`unit_kernel_from_d2`:
    RBF F = exp(-rho^2/2); M12 F = exp(-rho)/rho (0 where the 1e-36 clamp is
    active, since u_d = 0 there); M32 F = 3 exp(-sqrt3 rho);
    M52 F = 5/3 (1 + sqrt5 rho) exp(-sqrt5 rho).
    """
    if kind == "rbf":
        return np.exp(-0.5 * d2)
    tau = np.sqrt(np.maximum(d2, 1e-36))
    if kind == "matern12":
        f = np.exp(-tau) / tau
        return np.where(d2 > 1e-36, f, 0.0)
    if kind == "matern32":
        return 3.0 * np.exp(-_SQRT3 * tau)
    if kind == "matern52":
        return 5.0 / 3.0 * (1.0 + _SQRT5 * tau) * np.exp(-_SQRT5 * tau)
    raise ValueError(f"unknown kernel kind {kind!r}")


def cross_covariance(kind: str, X: np.ndarray, Z: np.ndarray, ls, var: float) -> np.ndarray:
    """K[i, j] = var * k(x_i / l, z_j / l)  (``kernels/computations/dense.py:32-36``)."""
    X = np.asarray(X, np.float64)
    Z = np.asarray(Z, np.float64)
    lsv = _as_ls(ls, X.shape[1])
    d2 = squared_distance(X / lsv, Z / lsv)
    return float(var) * unit_kernel_from_d2(kind, d2)


def gram(kind: str, X: np.ndarray, ls, var: float) -> np.ndarray:
    """Dense Gram ``cross_covariance(x, x)`` (``kernels/computations/base.py:56-72``)."""
    return cross_covariance(kind, X, X, ls, var)


def kernel_diagonal(kind: str, X: np.ndarray, ls, var: float) -> np.ndarray:
    """``vmap(kernel)(x, x)`` (``objectives.py:356``; ``computations/base.py:95-96``)."""
    X = np.asarray(X, np.float64)
    lsv = _as_ls(ls, X.shape[1])
    xs = X / lsv
    d2 = np.einsum("id,id->i", xs - xs, xs - xs)
    return float(var) * unit_kernel_from_d2(kind, d2)


def add_jitter(A: np.ndarray, jitter: float) -> np.ndarray:
    """``A + eye * jitter`` with the reference's shape checks (``linalg/utils.py:39-65``)."""
    if A.ndim != 2:
        raise ValueError(f"Expected 2D matrix, got {A.ndim}D array")
    if A.shape[0] != A.shape[1]:
        raise ValueError(f"Expected square matrix, got shape {A.shape}")
    return A + np.eye(A.shape[0]) * jitter


def sigma_matrix(kind, X, ls, var, obs_sd, jitter=1e-6) -> np.ndarray:
    """Sigma = (K + jitter I) + obs_sd^2 I, in the reference's order (``objectives.py:96-103``)."""
    K = gram(kind, X, ls, var)
    K = add_jitter(K, jitter)
    return K + np.eye(K.shape[0]) * (float(obs_sd) ** 2)


def _lu_slogdet_solve(S: np.ndarray, b: np.ndarray):
    """LU-based ``slogdet`` and ``solve`` (``linalg/operations.py:109-111,163-165``).

    ``jnp.linalg.slogdet`` and ``jnp.linalg.solve`` on a Dense operator both
    go through an LU factorisation; the sign is discarded
    (``operations.py:164``).
    """
    lu, piv = sla.lu_factor(S, check_finite=False)
    d = np.diag(lu)
    logdet = float(np.sum(np.log(np.abs(d))))
    x = sla.lu_solve((lu, piv), b, check_finite=False)
    return logdet, x, (lu, piv)


def gaussian_log_prob(mu: np.ndarray, S: np.ndarray, y: np.ndarray) -> float:
    """-1/2 [n log 2pi + log|S| + (y-mu)^T S^-1 (y-mu)] (``distributions.py:115-134``)."""
    diff = np.asarray(y, np.float64).reshape(-1) - np.asarray(mu, np.float64).reshape(-1)
    logdet, sol, _ = _lu_slogdet_solve(S, diff)
    n = diff.shape[0]
    return -0.5 * (n * LOG2PI + logdet + float(diff @ sol))


def conjugate_mll(kind, X, y, ls, var, obs_sd, jitter=1e-6, mean=0.0) -> float:
    """``objectives.conjugate_mll`` (``objectives.py:93-107``) with a constant mean."""
    X = np.asarray(X, np.float64)
    S = sigma_matrix(kind, X, ls, var, obs_sd, jitter)
    mu = np.full(X.shape[0], float(mean))
    return gaussian_log_prob(mu, S, np.asarray(y, np.float64).reshape(-1))


def conjugate_mll_parts(kind, X, y, ls, var, obs_sd, jitter=1e-6, mean=0.0) -> dict:
    """MLL with its pieces: logdet, quad = r^T Sigma^-1 r, alpha = Sigma^-1 r (LU path)."""
    X = np.asarray(X, np.float64)
    S = sigma_matrix(kind, X, ls, var, obs_sd, jitter)
    r = np.asarray(y, np.float64).reshape(-1) - float(mean)
    logdet, alpha, _ = _lu_slogdet_solve(S, r)
    quad = float(r @ alpha)
    n = r.shape[0]
    return {"mll": -0.5 * (n * LOG2PI + logdet + quad), "logdet": logdet,
            "quad": quad, "alpha": alpha}


def conjugate_mll_grad(kind, X, y, ls, var, obs_sd, jitter=1e-6, mean=0.0) -> dict:
    """Gradient of :func:`conjugate_mll` in constrained coordinates.

    The reference has no hand-written gradient: ``jax.grad`` differentiates
    through the LU ``slogdet``/``solve`` and the vmapped kernel
    (``fit.py:160``; ``tests/test_objectives.py:83-85``).  That derivative is
    d MLL/d theta = 1/2 tr(W dSigma/dtheta) with W = alpha alpha^T - Sigma^-1
    (Rasmussen & Williams 2006, eq. 5.9), Sigma^-1 formed from the same LU.
    Returns ``{"mll", "lengthscale", "variance", "obs_stddev", "mean"}``;
    ``lengthscale`` has the shape of ``ls`` (scalar -> summed over d).
    """
    X = np.asarray(X, np.float64)
    n, D = X.shape
    lsv = _as_ls(ls, D)
    xs = X / lsv
    d2 = squared_distance(xs, xs)
    ku = unit_kernel_from_d2(kind, d2)
    K = float(var) * ku
    S = add_jitter(K, jitter) + np.eye(n) * (float(obs_sd) ** 2)
    r = np.asarray(y, np.float64).reshape(-1) - float(mean)
    logdet, alpha, lupiv = _lu_slogdet_solve(S, r)
    Sinv = sla.lu_solve(lupiv, np.eye(n), check_finite=False)
    W = np.outer(alpha, alpha) - Sinv
    mll = -0.5 * (n * LOG2PI + logdet + float(r @ alpha))
    F = float(var) * unit_kernel_dlog_factor(kind, d2)
    WF = W * F
    g_ls = np.empty(D)
    for d in range(D):
        dd = X[:, d][:, None] / lsv[d] - X[:, d][None, :] / lsv[d]
        g_ls[d] = 0.5 * float(np.sum(WF * dd * dd)) / lsv[d]
    g_var = 0.5 * float(np.sum(W * ku))
    g_sd = float(obs_sd) * float(np.trace(W))
    g_mean = float(np.sum(alpha))
    g_l = g_ls if np.ndim(ls) else float(np.sum(g_ls))
    return {"mll": mll, "lengthscale": g_l, "variance": g_var,
            "obs_stddev": g_sd, "mean": g_mean, "alpha": alpha}


def collapsed_elbo(kind, X, y, Z, ls, var, obs_sd, jitter=1e-6, mean=0.0) -> float:
    """Titsias collapsed bound, ``objectives.collapsed_elbo`` (``objectives.py:342-416``)."""
    X = np.asarray(X, np.float64)
    Z = np.asarray(Z, np.float64)
    n = X.shape[0]
    m = Z.shape[0]
    noise = float(obs_sd) ** 2                                     # :350
    Kzz = add_jitter(gram(kind, Z, ls, var), jitter)              # :352-354
    Kzx = cross_covariance(kind, Z, X, ls, var)                   # :355
    Kxx_diag = kernel_diagonal(kind, X, ls, var)                  # :356
    Lz = np.linalg.cholesky(Kzz)                                  # :359
    A = sla.solve_triangular(Lz, Kzx, lower=True) / math.sqrt(noise)  # :387
    AAT = A @ A.T                                                 # :390
    B = np.eye(m) + AAT                                           # :393
    L = np.linalg.cholesky(B)                                     # :396
    log_det_B = 2.0 * float(np.sum(np.log(np.diag(L))))           # :399
    diff = np.asarray(y, np.float64).reshape(-1) - float(mean)    # :401
    L_inv_A_diff = sla.solve_triangular(L, A @ diff, lower=True)  # :404
    quad = (float(np.sum(diff ** 2)) - float(np.sum(L_inv_A_diff ** 2))) / noise  # :407
    two_log_prob = -n * math.log(2.0 * math.pi * noise) - log_det_B - quad        # :410
    two_trace = float(np.sum(Kxx_diag)) / noise - float(np.trace(AAT))            # :413
    return (two_log_prob - two_trace) / 2.0                                       # :416


def collapsed_elbo_grad(kind, X, y, Z, ls, var, obs_sd, jitter=1e-6, mean=0.0) -> dict:
    """Collapsed ELBO and its gradient w.r.t. (l, var, obs_sd, c, Z), constrained coordinates.

    Same objective as :func:`collapsed_elbo` (``objectives.py:342-416``), written
    through the N-summed statistics P = Kzx Kxz, b = Kzx r, q = Kzx 1 and
    C = Kzz + P / s (s = obs_sd^2), using the determinant and inversion lemmas
    the reference's own comments state (``objectives.py:361-385``):
      log|B| = log|C| - log|Kzz|,  ||L^-1 A r||^2 = b^T C^-1 b / s,
      tr(A A^T) = tr(Kzz^-1 P) / s,  sum diag(Kxx) = n var (stationary).
    Differentiating F gives 2 dF = 2 sum(W * dKzx) + tr(G_K dKzz) + g_s ds - n/s dvar
    with c = C^-1 b,
      G_P = -C^-1/s - c c^T/s^3 + Kzz^-1/s,  W = G_P Kzx + (c/s^2) r^T,
      G_K = Kzz^-1 - C^-1 - c c^T/s^2 - Kzz^-1 P Kzz^-1/s,
      g_s = -n/s + tr(C^-1 P)/s^2 + r^T r/s^2 + c^T P c/s^4 - 2 b^T c/s^3
            + n var/s^2 - tr(Kzz^-1 P)/s^2,
    and dKzx / dtheta, dKzx / dZ from the stationary kernel derivative
    (dk/drho^2 = -var F / 2, F as in :func:`unit_kernel_dlog_factor`).
    Returns ``{"elbo", "lengthscale", "variance", "obs_stddev", "mean", "Z"}``.
    """
    X = np.asarray(X, np.float64)
    Z = np.asarray(Z, np.float64)
    y = np.asarray(y, np.float64).reshape(-1)
    n, D = X.shape
    M = Z.shape[0]
    s = float(obs_sd) ** 2
    lsv = _as_ls(ls, D)
    r = y - float(mean)
    xs, zs = X / lsv, Z / lsv
    d2 = squared_distance(zs, xs)
    ku, F = unit_kernel_from_d2(kind, d2), unit_kernel_dlog_factor(kind, d2)
    Kzx = float(var) * ku
    P, b, q = Kzx @ Kzx.T, Kzx @ r, Kzx.sum(1)
    d2z = squared_distance(zs, zs)
    kzu, Fz = unit_kernel_from_d2(kind, d2z), unit_kernel_dlog_factor(kind, d2z)
    Kzz = add_jitter(float(var) * kzu, jitter)
    Lk = np.linalg.cholesky(Kzz)
    C = Kzz + P / s
    Lc = np.linalg.cholesky(C)
    c = sla.cho_solve((Lc, True), b)
    Kinv = sla.cho_solve((Lk, True), np.eye(M))
    Cinv = sla.cho_solve((Lc, True), np.eye(M))
    rr = float(r @ r)
    trKP = float(np.sum(Kinv * P))
    elbo = 0.5 * (-n * math.log(2 * math.pi * s) - 2 * np.sum(np.log(np.diag(Lc)))
                  + 2 * np.sum(np.log(np.diag(Lk))) - rr / s + float(b @ c) / s ** 2
                  - n * float(var) / s + trKP / s)
    GP = -Cinv / s - np.outer(c, c) / s ** 3 + Kinv / s
    W = GP @ Kzx + np.outer(c / s ** 2, r)
    GK = Kinv - Cinv - np.outer(c, c) / s ** 2 - Kinv @ P @ Kinv / s
    gs = (-n / s + float(np.sum(Cinv * P)) / s ** 2 + rr / s ** 2 + float(c @ P @ c) / s ** 4
          - 2 * float(b @ c) / s ** 3 + n * float(var) / s ** 2 - trKP / s ** 2)
    WF, GF = W * F, GK * Fz
    g_ls = np.empty(D)
    g_Z = np.empty((M, D))
    for d in range(D):
        dzx = zs[:, d][:, None] - xs[:, d][None, :]
        dzz = zs[:, d][:, None] - zs[:, d][None, :]
        g_ls[d] = float(var) * (float(np.sum(WF * dzx * dzx))
                                + 0.5 * float(np.sum(GF * dzz * dzz))) / lsv[d]
        g_Z[:, d] = -float(var) * (np.sum(WF * dzx, 1) + np.sum(GF * dzz, 1)) / lsv[d]
    g_var = float(np.sum(W * ku)) + 0.5 * float(np.sum(GK * kzu)) - 0.5 * n / s
    return {"elbo": elbo, "lengthscale": g_ls if np.ndim(ls) else float(np.sum(g_ls)),
            "variance": g_var, "obs_stddev": gs * float(obs_sd),
            "mean": float(np.sum(r)) / s - float(c @ q) / s ** 2, "Z": g_Z}


def predict(kind, X, y, T, ls, var, obs_sd, jitter_post=1e-6, jitter_prior=1e-6, mean=0.0):
    """``ConjugatePosterior.predict`` (``gps.py:494-526``): latent mean [T] and covariance [T,T].

    Sigma uses the posterior's jitter (``gps.py:506``), the predictive
    covariance the prior's (``gps.py:523``).
    """
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    noise = float(obs_sd) ** 2
    Kxx = add_jitter(gram(kind, X, ls, var), jitter_post)         # :505-507
    S = Kxx + np.eye(X.shape[0]) * noise                          # :509-510
    Ls = np.linalg.cholesky(S)                                    # :511
    Ktt = gram(kind, T, ls, var)                                  # :514
    Kxt = cross_covariance(kind, X, T, ls, var)                   # :515
    L_inv_Kxt = sla.solve_triangular(Ls, Kxt, lower=True)         # :517
    r = np.asarray(y, np.float64).reshape(-1) - float(mean)
    L_inv_y = sla.solve_triangular(Ls, r, lower=True)             # :518
    mu = float(mean) + L_inv_Kxt.T @ L_inv_y                      # :520
    cov = Ktt - L_inv_Kxt.T @ L_inv_Kxt                           # :522
    cov = add_jitter(cov, jitter_prior)                           # :523
    return mu, cov


def collapsed_predict(kind, X, y, Z, T, ls, var, obs_sd, jitter=1e-6, mean=0.0):
    """``CollapsedVariationalGaussian.predict`` (``variational_families.py:786-870``):
    latent mean [T] and covariance [T, T] of the collapsed (Titsias) posterior."""
    X = np.asarray(X, np.float64)
    Z = np.asarray(Z, np.float64)
    T = np.asarray(T, np.float64)
    m = Z.shape[0]
    noise_var = float(obs_sd) ** 2                                # :807
    Kzx = cross_covariance(kind, Z, X, ls, var)                   # :815
    Kzz = add_jitter(gram(kind, Z, ls, var), jitter)              # :816-818
    Lz = np.linalg.cholesky(Kzz)                                  # :822
    Lz_inv_Kzx = sla.solve_triangular(Lz, Kzx, lower=True)        # :825
    A = Lz_inv_Kzx / float(obs_sd)                                # :828
    L = np.linalg.cholesky(np.eye(m) + A @ A.T)                   # :831-834
    diff = np.asarray(y, np.float64).reshape(-1) - float(mean)    # :836-837
    w = sla.cho_solve((L, True), Lz_inv_Kzx @ diff)               # :840
    Kzz_inv_Kzx_diff = sla.solve_triangular(Lz.T, w, lower=False)  # :843
    Ktt = gram(kind, T, ls, var)                                  # :845
    Kzt = cross_covariance(kind, Z, T, ls, var)                   # :846
    Lz_inv_Kzt = sla.solve_triangular(Lz, Kzt, lower=True)        # :850
    L_inv_Lz_inv_Kzt = sla.solve_triangular(L, Lz_inv_Kzt, lower=True)  # :853
    mu = float(mean) + (Kzt.T / noise_var) @ Kzz_inv_Kzx_diff     # :856
    cov = Ktt - Lz_inv_Kzt.T @ Lz_inv_Kzt + L_inv_Lz_inv_Kzt.T @ L_inv_Lz_inv_Kzt  # :859-863
    return mu, add_jitter(cov, jitter)                            # :864-866


def conjugate_loocv(kind, X, y, ls, var, obs_sd, jitter=1e-6, mean=0.0) -> float:
    """``objectives.conjugate_loocv`` (``objectives.py:156-178``): sum over i of
    log N(y_i; mu_i, 1 / [Sigma^-1]_ii), mu_i = y_i - [Sigma^-1 r]_i / [Sigma^-1]_ii,
    Sigma = K + (obs_sd^2 + jitter) I, r = y - c; Sigma^-1 by ``inv`` (``:171``)."""
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    y = np.asarray(y, np.float64).reshape(-1)
    S = gram(kind, X, ls, var) + np.eye(n) * (float(obs_sd) ** 2 + jitter)  # :163-168
    r = y - float(mean)
    Sinv_r = np.linalg.solve(S, r)                                # :170
    dg = np.diag(np.linalg.inv(S))                                # :171-172
    mu = float(mean) + r - Sinv_r / dg                            # :174
    sd = np.sqrt(1.0 / dg)                                        # :175
    z = (y - mu) / sd
    return float(np.sum(-0.5 * z * z - np.log(sd) - 0.5 * LOG2PI))  # :177-178


def conjugate_loocv_grad(kind, X, y, ls, var, obs_sd, jitter=1e-6, mean=0.0) -> dict:
    """Gradient of :func:`conjugate_loocv` in constrained coordinates (``jax.grad``
    in the reference, ``tests/test_objectives.py:123-126``).

    With alpha = Sigma^-1 r and d_i = [Sigma^-1]_ii the objective is
    sum_i [-1/2 log 2pi + 1/2 log d_i - 1/2 alpha_i^2 / d_i], so
    d LOOCV = 1/2 tr(W dSigma) + sum(w) dc with v = alpha / d, w = Sigma^-1 v,
    b = 1/(2d) + alpha^2/(2 d^2) and W = alpha w^T + w alpha^T - 2 Sigma^-1 diag(b) Sigma^-1
    (Rasmussen & Williams 2006, sec. 5.4.2 and eq. 5.13).
    """
    X = np.asarray(X, np.float64)
    n, D = X.shape
    lsv = _as_ls(ls, D)
    xs = X / lsv
    d2 = squared_distance(xs, xs)
    ku = unit_kernel_from_d2(kind, d2)
    S = float(var) * ku + np.eye(n) * (float(obs_sd) ** 2 + jitter)
    r = np.asarray(y, np.float64).reshape(-1) - float(mean)
    Sinv = np.linalg.inv(S)
    alpha = Sinv @ r
    dg = np.diag(Sinv).copy()
    val = float(np.sum(-0.5 * LOG2PI + 0.5 * np.log(dg) - 0.5 * alpha ** 2 / dg))
    w = Sinv @ (alpha / dg)
    b = 0.5 / dg + 0.5 * alpha ** 2 / dg ** 2
    W = np.outer(alpha, w) + np.outer(w, alpha) - 2.0 * (Sinv * b[None, :]) @ Sinv
    F = float(var) * unit_kernel_dlog_factor(kind, d2)
    WF = W * F
    g_ls = np.empty(D)
    for d in range(D):
        dd = X[:, d][:, None] / lsv[d] - X[:, d][None, :] / lsv[d]
        g_ls[d] = 0.5 * float(np.sum(WF * dd * dd)) / lsv[d]
    g_l = g_ls if np.ndim(ls) else float(np.sum(g_ls))
    return {"loocv": val, "lengthscale": g_l, "variance": 0.5 * float(np.sum(W * ku)),
            "obs_stddev": float(obs_sd) * float(np.trace(W)), "mean": float(np.sum(w))}


def gaussian_likelihood_predict(cov: np.ndarray, obs_sd: float) -> np.ndarray:
    """Noisy predictive covariance, diag += obs_sd^2 (``likelihoods.py:182-186``)."""
    return cov + np.eye(cov.shape[0]) * (float(obs_sd) ** 2)
