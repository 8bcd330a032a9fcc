"""Host logic of ``gpjax_amd.fit`` (mirrors ``gpjax/fit.py:133-180``) on the CPU.

The Adam step is pinned against ``torch.optim.Adam`` -- an independent
implementation of Kingma & Ba (2015) with the same conventions as optax's
``adam`` defaults (b1 = 0.9, b2 = 0.999, eps = 1e-8 added outside the square
root, eps_root = 0) -- on the unconstrained coordinates, and the first step
against its closed form (bias-corrected m / sqrt(v) = sign(g) at t = 1).
"""

import math

import pytest
import torch

from gpjax_amd.fit import fit
from gpjax_amd.parameters import PositiveReal, Real, inverse_softplus, softplus


class _Toy:
    def __init__(self):
        self.a = Real(torch.tensor([0.0, -1.0]))
        self.b = PositiveReal(2.0)


def _objective(model, data):
    a, b = model.a.value, model.b.value
    return -((a - torch.tensor([3.0, 0.5], dtype=torch.float64)) ** 4).sum() - (b - 0.5) ** 2


def test_first_step_closed_form():
    m = _Toy()
    _, hist = fit(model=m, objective=_objective, train_data=None, num_iters=1, learning_rate=0.1)
    # d loss / d a = 4 (a - target)^3 < 0 for both entries -> a moves up by lr (|g| >> eps)
    assert torch.allclose(m.a.value.detach(), torch.tensor([0.1, -0.9], dtype=torch.float64),
                          atol=1e-9)
    # b: loss gradient 2 (b - 0.5) > 0 -> the raw (softplus^-1) coordinate moves down by lr
    raw0 = inverse_softplus(torch.tensor(2.0, dtype=torch.float64))
    assert abs(float(m.b.value.detach()) - float(softplus(raw0 - 0.1))) < 1e-9
    # the recorded loss is -objective at the starting point
    assert abs(float(hist[0]) - (3.0 ** 4 + 1.5 ** 4 + 1.5 ** 2)) < 1e-9


def test_trajectory_matches_torch_adam():
    steps, lr = 25, 0.05
    m = _Toy()
    _, hist = fit(model=m, objective=_objective, train_data=None, num_iters=steps,
                  learning_rate=lr)
    # reference: torch.optim.Adam on the unconstrained coordinates
    ua = torch.tensor([0.0, -1.0], dtype=torch.float64, requires_grad=True)
    ub = inverse_softplus(torch.tensor(2.0, dtype=torch.float64)).clone().requires_grad_(True)
    opt = torch.optim.Adam([ua, ub], lr=lr, betas=(0.9, 0.999), eps=1e-8)
    ref_hist = []
    for _ in range(steps):
        opt.zero_grad()
        b = softplus(ub)
        loss = ((ua - torch.tensor([3.0, 0.5], dtype=torch.float64)) ** 4).sum() + (b - 0.5) ** 2
        loss.backward()
        ref_hist.append(float(loss))
        opt.step()
    assert torch.allclose(hist, torch.tensor(ref_hist, dtype=torch.float64), rtol=1e-12)
    assert torch.allclose(m.a.value.detach(), ua.detach(), rtol=1e-12)
    assert abs(float(m.b.value.detach()) - float(softplus(ub.detach()))) < 1e-12


def test_converges_and_validates():
    m = _Toy()
    _, hist = fit(model=m, objective=_objective, train_data=None, num_iters=400,
                  learning_rate=0.05)
    assert float(hist[-1]) < 1e-3 * float(hist[0])
    assert abs(float(m.b.value.detach()) - 0.5) < 1e-2
    with pytest.raises(ValueError):
        fit(model=m, objective=_objective, train_data=None, num_iters=0)
    with pytest.raises(ValueError):
        fit(model=object(), objective=_objective, train_data=None, num_iters=1)
    assert math.isfinite(float(hist[0]))
