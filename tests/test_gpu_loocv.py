"""GPU parity of the leave-one-out objective (gpk_loocv_forward / _backward).

The oracle's ``conjugate_loocv`` restates ``objectives.py:156-178`` and is
pinned in tests/test_oracle.py against brute-force leave-one-out conditioning;
``conjugate_loocv_grad`` against torch autograd of that brute force and central
differences.  Tolerances: FP64 value within 1e-10 relative and gradient within
1e-8 relative norm-wise; mixed value within 1e-6 and gradient within 1e-4.
The reference's own test sizes (n in 1, 2, 10; d in 1, 2, 3; RBF with default
parameters and a Constant mean, ``tests/test_objectives.py:88-126``) run first.
"""

import numpy as np
import pytest
import torch

import gpjax_amd as gpx
from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu

KCLS = {"rbf": gpx.kernels.RBF, "matern12": gpx.kernels.Matern12,
        "matern32": gpx.kernels.Matern32, "matern52": gpx.kernels.Matern52}


def _model(cuda, X, y, kind, ls, var, sd, mean, jitter=1e-6):
    k = KCLS[kind](lengthscale=ls, variance=var)
    mf = gpx.mean_functions.Constant(gpx.parameters.Real(mean))
    prior = gpx.gps.Prior(mean_function=mf, kernel=k, jitter=jitter)
    post = prior * gpx.likelihoods.Gaussian(num_datapoints=X.shape[0], obs_stddev=sd)
    D = gpx.Dataset(X=torch.tensor(X, device=cuda), y=torch.tensor(y.reshape(-1, 1), device=cuda))
    return post, D, k


def _grads(post, k):
    return np.concatenate([np.atleast_1d(k.lengthscale.value.grad.cpu().numpy()),
                           [k.variance.value.grad.item(),
                            post.likelihood.obs_stddev.value.grad.item(),
                            post.prior.mean_function.constant.value.grad.item()]])


def _ref_grads(g):
    return np.concatenate([np.atleast_1d(g["lengthscale"]),
                           [g["variance"], g["obs_stddev"], g["mean"]]])


@pytest.mark.parametrize("n", [1, 2, 10])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_loocv_reference_sizes(cuda, n, d):
    rng = np.random.default_rng(100 * n + d)
    X = rng.uniform(-2, 2, (n, d))
    y = np.sin(X[:, 0]) + 0.1 * rng.standard_normal(n)
    post, D, k = _model(cuda, X, y, "rbf", 1.0, 1.0, 1.0, 0.0)
    v, res = gpx.objectives.conjugate_loocv_with_info(post, D, precision="fp64")
    (-v).backward()
    g = O.conjugate_loocv_grad("rbf", X, y, 1.0, 1.0, 1.0, 1e-6, 0.0)
    assert int(res[4].item()) == 0
    assert v.item() == pytest.approx(g["loocv"], rel=1e-10)
    gr = -_ref_grads(g)
    assert np.linalg.norm(_grads(post, k) - gr) <= 1e-8 * np.linalg.norm(gr)


@pytest.mark.parametrize("kind,n,d,ard", [("matern32", 300, 3, True), ("matern52", 1000, 4, False),
                                          ("matern12", 777, 2, True), ("rbf", 513, 5, True)])
def test_loocv_fp64_parity(cuda, kind, n, d, ard):
    rng = np.random.default_rng(n + d)
    X = rng.uniform(-2, 2, (n, d))
    y = np.sin(X[:, 0]) + 0.1 * rng.standard_normal(n)
    ls = np.exp(rng.uniform(np.log(0.5), np.log(2.0), d)) if ard else 0.9
    var, sd, c = 1.3, 0.3, 0.15
    post, D, k = _model(cuda, X, y, kind, ls, var, sd, c)
    v = gpx.objectives.conjugate_loocv(post, D, precision="fp64")
    v.backward()
    g = O.conjugate_loocv_grad(kind, X, y, ls, var, sd, 1e-6, c)
    assert v.item() == pytest.approx(g["loocv"], rel=1e-10)
    gr = _ref_grads(g)
    assert np.linalg.norm(_grads(post, k) - gr) <= 1e-8 * np.linalg.norm(gr)


def test_loocv_mixed_parity(cuda):
    rng = np.random.default_rng(11)
    n, d = 3000, 3
    X = rng.uniform(-2, 2, (n, d))
    y = np.sin(X[:, 0]) + 0.1 * rng.standard_normal(n)
    ls, var, sd, c = np.array([0.7, 1.1, 1.6]), 1.3, 0.3, 0.1
    post, D, k = _model(cuda, X, y, "matern52", ls, var, sd, c)
    v = gpx.objectives.conjugate_loocv(post, D, precision="mixed")
    v.backward()
    g = O.conjugate_loocv_grad("matern52", X, y, ls, var, sd, 1e-6, c)
    rel = abs(v.item() - g["loocv"]) / abs(g["loocv"])
    gr = _ref_grads(g)
    grel = np.linalg.norm(_grads(post, k) - gr) / np.linalg.norm(gr)
    print(f"loocv mixed n={n}: rel value {rel:.3e}, rel grad {grel:.3e}", flush=True)
    assert rel <= 1e-6
    assert grel <= 1e-4


def test_loocv_then_mll_on_fresh_workspaces(cuda):
    """The LOOCV backward overwrites the factor L in its own workspace; the MLL
    on the same model afterwards is unaffected (separate workspaces)."""
    rng = np.random.default_rng(5)
    X = rng.uniform(-2, 2, (400, 2))
    y = np.sin(X[:, 0]) + 0.1 * rng.standard_normal(400)
    post, D, k = _model(cuda, X, y, "matern32", np.array([0.8, 1.2]), 1.1, 0.25, 0.0)
    gpx.objectives.conjugate_loocv(post, D, precision="fp64").backward()
    m = gpx.objectives.conjugate_mll(post, D, precision="fp64").item()
    ref = O.conjugate_mll("matern32", X, y, np.array([0.8, 1.2]), 1.1, 0.25, 1e-6, 0.0)
    assert m == pytest.approx(ref, rel=1e-10)
