"""The device-resident training loop (``gpjax_amd.fit``, mirroring
``gpjax/fit.py:133-180``) driving the HIP objective, eager and replayed from a
captured HIP graph.

The graph holds one whole step -- softplus bijection, ``gpk_mll_forward``,
``gpk_mll_backward``, the chain rule and the Adam update -- and is replayed
with the program's own eager work in between.  Every replay must reproduce the
eager trajectory (same kernels, same deterministic reductions: agreement to
1e-12 relative), and the loss must go down.
"""

import pytest
import torch

import gpjax_amd as gpx
from gpjax_amd import workloads

pytestmark = pytest.mark.gpu


def _model(w, n):
    k = gpx.kernels.RBF(lengthscale=1.0, variance=0.5)
    return gpx.gps.Prior(mean_function=gpx.mean_functions.Zero(), kernel=k, jitter=1e-6) * \
        gpx.likelihoods.Gaussian(num_datapoints=n, obs_stddev=0.3)


def _objective(model, data):
    return gpx.objectives.conjugate_mll(model, data, precision="fp64")


@pytest.mark.parametrize("n", [256, 700])
def test_fit_graph_replay_matches_eager(cuda, n):
    w = workloads.exact_small(n=n)
    D = gpx.Dataset(X=torch.tensor(w["X"], device=cuda), y=torch.tensor(w["y"][:, None], device=cuda))
    steps = 12
    m1, h1 = gpx.fit(model=_model(w, n), objective=_objective, train_data=D, num_iters=steps,
                     learning_rate=0.05)
    torch.cuda.synchronize()
    m2, h2 = gpx.fit(model=_model(w, n), objective=_objective, train_data=D, num_iters=steps,
                     learning_rate=0.05, capture_graph=True, warmup=3)
    torch.cuda.synchronize()
    h1, h2 = h1.cpu(), h2.cpu()
    assert torch.isfinite(h1).all()
    assert torch.allclose(h1, h2, rtol=1e-12, atol=0.0), (h1, h2)
    assert float(h1[-1]) < float(h1[0])
    for a, b in [(m1.prior.kernel.lengthscale, m2.prior.kernel.lengthscale),
                 (m1.prior.kernel.variance, m2.prior.kernel.variance),
                 (m1.likelihood.obs_stddev, m2.likelihood.obs_stddev)]:
        assert torch.allclose(a.value.detach(), b.value.detach(), rtol=1e-12)
