"""Writes tests/golden/exact_small.npz: seeded small problems and the float64
oracle's answers for them.

GPJax itself cannot run in this environment (it needs jax and Python >= 3.11),
so these vectors come from the oracle (oracle/gp_oracle.py), which the CPU
tests pin against closed forms, SciPy, brute force, torch autograd and finite
differences.  The fixtures freeze those answers: tests/test_oracle.py checks
that the oracle still reproduces them, and tests/test_gpu_golden.py checks the
device path against them without running the oracle.

Grid: N in {1, 2, 10, 64, 257} x D in {1, 3, 8} x {rbf, matern12, matern32,
matern52} x {isotropic, ARD}.  Per (N, D): X ~ U[-2, 2]^D, y = sin(x_1) +
0.1 eps, 16 test points T ~ U[-2.5, 2.5]^D, from numpy's PCG64 with seed
1000 N + D.  Per case: variance 1.3, obs_stddev 0.3, Constant mean 0.2,
jitter 1e-6, lengthscale 0.9 (isotropic) or exp U[ln 0.5, ln 2]^D (ARD, seed
1000 N + D + 17 k, k the kernel's index).  Outputs: the MLL (objectives.py:
36-107), alpha = Sigma^-1 (y - m), the gradient in constrained coordinates,
the latent predictive mean and variance at T (gps.py:443-526) and the
collapsed ELBO and its gradient with Z = X[:M], M in {max(1, N // 2), N}
(objectives.py:321-416).

Run from the repository root:  python tests/golden/make_fixtures.py
"""

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import gp_oracle as O  # noqa: E402

NS = (1, 2, 10, 64, 257)
DS = (1, 3, 8)
KINDS = ("rbf", "matern12", "matern32", "matern52")
VAR, SD, MEAN, JIT = 1.3, 0.3, 0.2, 1e-6
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "exact_small.npz")


def problem(n, d):
    rng = np.random.default_rng(1000 * n + d)
    X = rng.uniform(-2, 2, (n, d))
    y = np.sin(X[:, 0]) + 0.1 * rng.standard_normal(n)
    T = rng.uniform(-2.5, 2.5, (16, d))
    return X, y, T


def lengthscale(n, d, k, ard):
    if not ard:
        return np.array([0.9])
    rng = np.random.default_rng(1000 * n + d + 17 * k)
    return np.exp(rng.uniform(np.log(0.5), np.log(2.0), d))


def elbo_sizes(n):
    return sorted({max(1, n // 2), n})


def elbo_spread(kind, X, y, m, ls, elbo, draws=3, eps=1e-15):
    """How many digits the problem determines its ELBO to: the largest relative
    change over the reference's form of the ELBO (objectives.py:381-416, against
    the streamed form's `elbo`) and over `draws` evaluations with X (and so
    Z = X[:m]) perturbed by relative 1e-15 (a few units in the last place)."""
    out = abs(O.collapsed_elbo(kind, X, y, X[:m], ls, VAR, SD, JIT, MEAN) - elbo) / abs(elbo)
    rng = np.random.default_rng(len(X) * 31 + m)
    for _ in range(draws):
        Xp = X * (1.0 + eps * rng.standard_normal(X.shape))
        e = O.collapsed_elbo_grad(kind, Xp, y, Xp[:m], ls, VAR, SD, JIT, MEAN)["elbo"]
        out = max(out, abs(e - elbo) / abs(elbo))
    return out


def group_key(n, d, ard):
    return f"N{n}_D{d}_{'ard' if ard else 'iso'}"


def flat_grad(g):
    return np.concatenate([np.atleast_1d(g["lengthscale"]),
                           [g["variance"], g["obs_stddev"], g["mean"]]])


def compute(with_spread=True):
    """Arrays per (N, D, isotropic/ARD) group, first axis over KINDS: _ls
    [4, nls], _mll [4], _alpha [4, N], _grad [4, nls + 3] (lengthscales,
    variance, obs_stddev, mean constant), _pmean and _pvar [4, 16], and per M
    _elbo_M{M} [4], _elbograd_M{M} [4, nls + 3] and _elbospread_M{M} [4]
    (elbo_spread: up to ~6e-7 where Z = X with D = 1 makes Kzz + 1e-6 I nearly
    singular, ~1e-15 where it is well conditioned)."""
    out = {}
    for n in NS:
        for d in DS:
            X, y, T = problem(n, d)
            out[f"N{n}_D{d}_X"], out[f"N{n}_D{d}_y"], out[f"N{n}_D{d}_T"] = X, y, T
            for ard in (False, True):
                rows = {}
                for k, kind in enumerate(KINDS):
                    ls = lengthscale(n, d, k, ard)
                    lsa = ls if ard else float(ls[0])
                    g = O.conjugate_mll_grad(kind, X, y, lsa, VAR, SD, JIT, MEAN)
                    S = O.sigma_matrix(kind, X, lsa, VAR, SD, JIT)
                    mu, cov = O.predict(kind, X, y, T, lsa, VAR, SD, JIT, JIT, MEAN)
                    r = {"ls": ls, "mll": O.conjugate_mll(kind, X, y, lsa, VAR, SD, JIT, MEAN),
                         "alpha": np.linalg.solve(S, y - MEAN), "grad": flat_grad(g),
                         "pmean": mu, "pvar": np.diag(cov).copy()}
                    for m in elbo_sizes(n):
                        e = O.collapsed_elbo_grad(kind, X, y, X[:m], lsa, VAR, SD, JIT, MEAN)
                        r[f"elbo_M{m}"], r[f"elbograd_M{m}"] = e["elbo"], flat_grad(e)
                        if with_spread:
                            r[f"elbospread_M{m}"] = elbo_spread(kind, X, y, m, lsa, e["elbo"])
                    for f, v in r.items():
                        rows.setdefault(f, []).append(v)
                gk = group_key(n, d, ard)
                for f, v in rows.items():
                    out[f"{gk}_{f}"] = np.array(v, dtype=np.float64)
    meta = {"ns": NS, "ds": DS, "kinds": KINDS, "variance": VAR, "obs_stddev": SD,
            "mean": MEAN, "jitter": JIT}
    out["meta"] = np.array(json.dumps(meta))
    return out


if __name__ == "__main__":
    np.savez_compressed(OUT, **compute())
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
