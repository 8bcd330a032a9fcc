"""Host logic of ``gpjax_amd.fit`` (mirrors ``gpjax/fit.py:44-180``) on the CPU.

The optimiser step itself is a libgpk kernel (``gpk_adam_step``); its
trajectories are pinned on the GPU (``tests/test_gpu_fit.py``).  Here: argument
validation, parameter discovery, the bijection helpers, and that ``fit``
refuses host tensors instead of falling back to a host optimiser.
"""

import math

import pytest
import torch

from gpjax_amd.fit import _collect_parameters, fit
from gpjax_amd.parameters import (NonNegativeReal, PositiveReal, Real, inverse_softplus,
                                  softplus)


class _Toy:
    def __init__(self):
        self.a = Real(torch.tensor([0.0, -1.0]))
        self.b = PositiveReal(2.0)
        self.nested = [NonNegativeReal(0.5), {"not": "walked"}]
        self.same = self.b


def _objective(model, data):
    a, b = model.a.value, model.b.value
    return ((a - torch.tensor([3.0, 0.5], dtype=torch.float64)) ** 4).sum() + (b - 0.5) ** 2


def test_collects_each_parameter_once():
    m = _Toy()
    ps = _collect_parameters(m)
    assert [type(p).__name__ for p in ps] == ["Real", "PositiveReal", "NonNegativeReal"]


def test_softplus_pair_is_inverse_and_overflow_free():
    x = torch.tensor([-40.0, -3.0, 0.0, 1e-3, 2.0, 30.0, 800.0], dtype=torch.float64)
    y = softplus(x)
    assert torch.isfinite(y).all()
    # log(1 + e^x) in closed form where it does not overflow
    small = x[:-1]
    assert torch.allclose(y[:-1], torch.log1p(torch.exp(small)), rtol=1e-15, atol=0.0)
    assert float(y[-1]) == 800.0
    z = torch.tensor([1e-6, 0.1, 1.0, 5.0, 50.0], dtype=torch.float64)
    assert torch.allclose(softplus(inverse_softplus(z)), z, rtol=1e-12)


def test_validates_arguments():
    m = _Toy()
    with pytest.raises(ValueError):
        fit(model=m, objective=_objective, train_data=None, num_iters=0)
    with pytest.raises(ValueError):
        fit(model=object(), objective=_objective, train_data=None, num_iters=1)
    with pytest.raises(ValueError, match="warmup"):
        fit(model=m, objective=_objective, train_data=None, num_iters=10, capture_graph=True,
            warmup=0)


def test_host_parameters_are_refused():
    """No host optimiser: parameters on the CPU are an error, not a fallback."""
    m = _Toy()
    for p in _collect_parameters(m):
        p.value = p.value.detach().cpu().requires_grad_(True)
    with pytest.raises(ValueError, match="GPU"):
        fit(model=m, objective=_objective, train_data=None, num_iters=1)
    assert math.isfinite(float(m.b.value))
