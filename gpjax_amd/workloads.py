"""Synthetic workloads of the benchmark configurations (numpy PCG64, float64).

Every generator draws from ``numpy.random.Generator(PCG64(seed))`` in the order
inputs -> hyperparameters -> weights -> noise, and returns a dict with
``X [N, D]``, ``y [N]``, ``kind``, ``lengthscale``, ``variance``, ``obs_stddev``,
``jitter`` (plus ``Z`` / ``T`` / batch lists where the configuration has them).
``n`` overrides the number of data points (same recipe, smaller sample).

  exact_small   N=2,000, D=1, X~U[-3,3], y = sin 3x + cos 2x + N(0, 0.1^2);
                RBF l=0.5, var=1, obs_sd=0.1; 500 test points on [-3.5, 3.5].
  exact_matern  N=16,384, D=8, X~N(0, I); Matern-5/2 ARD, l_d = exp U[ln .5, ln 2],
                var=1.5, obs_sd=0.05; y = sum_d sin(x_d w_d) + N(0, 0.05^2),
                w_d ~ U[0.5, 2].
  exact_large   N=65,536, D=16, X = G diag(s), G~N(0, I), s_d~U[0.5, 2]; RBF ARD,
                l_d = exp U[0, ln 4], var=1, obs_sd=0.1; y from a 2,048-feature
                random-Fourier-feature prior sample (omega~N(0, I), features
                sqrt(var/F)[cos(X (omega/l)^T), sin(...)], theta~N(0, I)) + noise,
                the construction of gpjax/gps.py:315-322,769-772 and
                gpjax/kernels/computations/basis_functions.py:58-62.
  sparse        N=1,000,000, D=8, X~U[0,1]^8, Z = 2,048 rows of X without
                replacement; Matern-3/2 ARD l_d~U[0.1, 0.5], var=1, obs_sd=0.2;
                y = sum_d sin(2 pi x_d) + N(0, 0.2^2).
  batched       4,096 GPs, N_i~U{32..1024}, D=4, X_i~U[0,1]^4; per GP RBF ARD
                l~logU[0.05, 1], var~logU[0.5, 2], obs_sd~logU[1e-3, 1e-1];
                y = sin(6 x_1) + x_2^2 + N(0, obs_sd^2).
"""

from __future__ import annotations

import numpy as np

SEEDS = {"exact_small": 0, "exact_matern": 1, "exact_large": 2, "sparse": 3, "batched": 4}


def _rng(name: str) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(SEEDS[name]))


def exact_small(n: int = 2000) -> dict:
    rng = _rng("exact_small")
    X = rng.uniform(-3.0, 3.0, (n, 1))
    y = np.sin(3 * X[:, 0]) + np.cos(2 * X[:, 0]) + 0.1 * rng.standard_normal(n)
    T = np.linspace(-3.5, 3.5, 500)[:, None]
    return dict(X=X, y=y, T=T, kind="rbf", lengthscale=0.5, variance=1.0, obs_stddev=0.1,
                jitter=1e-6)


def exact_matern(n: int = 16384) -> dict:
    rng = _rng("exact_matern")
    D = 8
    X = rng.standard_normal((n, D))
    ls = np.exp(rng.uniform(np.log(0.5), np.log(2.0), D))
    w = rng.uniform(0.5, 2.0, D)
    y = np.sin(X * w).sum(1) + 0.05 * rng.standard_normal(n)
    return dict(X=X, y=y, kind="matern52", lengthscale=ls, variance=1.5, obs_stddev=0.05,
                jitter=1e-6)


def exact_large(n: int = 65536, features: int = 2048, chunk: int = 8192) -> dict:
    rng = _rng("exact_large")
    D = 16
    G = rng.standard_normal((n, D))
    s = rng.uniform(0.5, 2.0, D)
    X = G * s
    ls = np.exp(rng.uniform(0.0, np.log(4.0), D))
    var, sd = 1.0, 0.1
    omega = rng.standard_normal((features, D))
    theta = rng.standard_normal(2 * features)
    noise = sd * rng.standard_normal(n)
    W = (omega / ls).T                              # [D, F]
    scale = np.sqrt(var / features)
    y = np.empty(n)
    for i0 in range(0, n, chunk):
        z = X[i0:i0 + chunk] @ W
        y[i0:i0 + chunk] = scale * (np.cos(z) @ theta[:features] + np.sin(z) @ theta[features:])
    y += noise
    return dict(X=X, y=y, kind="rbf", lengthscale=ls, variance=var, obs_stddev=sd, jitter=1e-6)


def sparse(n: int = 1_000_000, m: int = 2048) -> dict:
    rng = _rng("sparse")
    D = 8
    X = rng.uniform(0.0, 1.0, (n, D))
    idx = rng.choice(n, size=m, replace=False)
    ls = rng.uniform(0.1, 0.5, D)
    y = np.sin(2 * np.pi * X).sum(1) + 0.2 * rng.standard_normal(n)
    return dict(X=X, y=y, Z=X[idx].copy(), kind="matern32", lengthscale=ls, variance=1.0,
                obs_stddev=0.2, jitter=1e-6)


def batched(num: int = 4096, nmin: int = 32, nmax: int = 1024) -> dict:
    rng = _rng("batched")
    D = 4
    sizes = rng.integers(nmin, nmax + 1, num)
    Xs = [rng.uniform(0.0, 1.0, (int(k), D)) for k in sizes]
    ls = np.exp(rng.uniform(np.log(0.05), np.log(1.0), (num, D)))
    var = np.exp(rng.uniform(np.log(0.5), np.log(2.0), num))
    sd = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), num))
    ys = [np.sin(6 * X[:, 0]) + X[:, 1] ** 2 + s * rng.standard_normal(X.shape[0])
          for X, s in zip(Xs, sd)]
    return dict(Xs=Xs, ys=ys, sizes=sizes, kind="rbf", lengthscale=ls, variance=var,
                obs_stddev=sd, jitter=1e-6)
