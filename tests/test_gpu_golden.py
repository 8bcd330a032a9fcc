"""The device path against the committed golden fixtures (tests/golden/exact_small.npz).

120 small problems: N in {1, 2, 10, 64, 257} x D in {1, 3, 8} x four kernels x
{isotropic, ARD}; the fixture file's generator documents the data.  The
oracle is not run here: the fixtures hold its answers (and
tests/test_oracle.py keeps them in step with it).  FP64 path, per problem:

* conjugate MLL within 1e-10 relative; its gradient (lengthscales, variance,
  obs_stddev, mean constant) within 1e-8 relative, norm-wise;
* latent predictive mean at 16 test points within 1e-9 relative (norm-wise),
  variance within 1e-9 of the largest;
* collapsed ELBO with Z = X[:M], M in {max(1, N // 2), N}, within
  max(1e-9, 10 s) relative, where s is the fixture's spread: the largest
  relative change of that ELBO under a second float64 formulation and under
  perturbations of X by a few units in the last place (up to 6e-7 when D = 1
  and Z = X makes Kzz + 1e-6 I nearly singular; a sample, hence the factor
  10); its gradient within 1e-7 relative where s <= 1e-10, and not compared
  where the problem does not determine it (31 of the 216 ELBOs).
"""

import json
import os

import numpy as np
import pytest
import torch

import gpjax_amd as gpx

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "exact_small.npz"))
META = json.loads(str(FIX["meta"]))
KCLS = {"rbf": gpx.kernels.RBF, "matern12": gpx.kernels.Matern12,
        "matern32": gpx.kernels.Matern32, "matern52": gpx.kernels.Matern52}
CASES = [(n, d, kind) for n in META["ns"] for d in META["ds"] for kind in META["kinds"]]


def _posterior(kind, ls, n):
    k = KCLS[kind](lengthscale=ls, variance=META["variance"])
    mf = gpx.mean_functions.Constant(gpx.parameters.Real(META["mean"]))
    prior = gpx.gps.Prior(mean_function=mf, kernel=k, jitter=META["jitter"])
    post = prior * gpx.likelihoods.Gaussian(num_datapoints=n, obs_stddev=META["obs_stddev"])
    return post, k


def _grads(post, k):
    return np.concatenate([np.atleast_1d(k.lengthscale.value.grad.cpu().numpy()),
                           [k.variance.value.grad.item(),
                            post.likelihood.obs_stddev.value.grad.item(),
                            post.prior.mean_function.constant.value.grad.item()]])


def _rel(a, b):
    return np.linalg.norm(np.atleast_1d(a - b)) / np.linalg.norm(np.atleast_1d(b))


@pytest.mark.parametrize("n,d,kind", CASES)
def test_golden_fp64(cuda, n, d, kind):
    ki = META["kinds"].index(kind)
    X, y, T = FIX[f"N{n}_D{d}_X"], FIX[f"N{n}_D{d}_y"], FIX[f"N{n}_D{d}_T"]
    Xt = torch.tensor(X, device=cuda)
    data = gpx.Dataset(X=Xt, y=torch.tensor(y[:, None], device=cuda))
    Tt = torch.tensor(T, device=cuda)
    for ard in (False, True):
        g = f"N{n}_D{d}_{'ard' if ard else 'iso'}"
        ls = FIX[g + "_ls"][ki]
        lsa = ls if ard else float(ls[0])
        case = f"{g} {kind}"

        post, k = _posterior(kind, lsa, n)
        mll = gpx.objectives.conjugate_mll(post, data, precision="fp64")
        mll.backward()
        ref = FIX[g + "_mll"][ki]
        assert abs(mll.item() - ref) <= 1e-10 * abs(ref), case
        assert _rel(_grads(post, k), FIX[g + "_grad"][ki]) <= 1e-8, case

        post, k = _posterior(kind, lsa, n)
        dist = post.predict(Tt, data, full_covariance=False, precision="fp64")
        mu, var = dist.mean.cpu().numpy(), dist.variance.cpu().numpy()
        pm, pv = FIX[g + "_pmean"][ki], FIX[g + "_pvar"][ki]
        assert _rel(mu, pm) <= 1e-9, case
        assert np.max(np.abs(var - pv)) <= 1e-9 * np.max(np.abs(pv)), case

        for m in sorted({max(1, n // 2), n}):
            post, k = _posterior(kind, lsa, n)
            q = gpx.variational_families.CollapsedVariationalGaussian(
                posterior=post, inducing_inputs=Xt[:m].clone())
            elbo, res = gpx.objectives.collapsed_elbo_with_info(q, data)
            elbo.backward()
            spread = FIX[g + f"_elbospread_M{m}"][ki]
            eref = FIX[g + f"_elbo_M{m}"][ki]
            assert int(res[4].item()) == 0, case
            assert abs(elbo.item() - eref) <= max(1e-9, 10 * spread) * abs(eref), (case, m)
            if spread <= 1e-10:
                assert _rel(_grads(post, k), FIX[g + f"_elbograd_M{m}"][ki]) <= 1e-7, (case, m)
