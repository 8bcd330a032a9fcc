"""Parameter containers and the softplus bijection.

Mirrors ``gpjax/parameters.py``: ``Parameter`` (69-80), ``NonNegativeReal``
(83-88), ``PositiveReal`` (91-96), ``Real`` (99-103) and ``DEFAULT_BIJECTION``
(140-146: softplus for "positive"/"non_negative", identity for "real").
Values are torch float64 leaf tensors with ``requires_grad``; gradients of the
objectives land on ``.value.grad`` in constrained coordinates, which is what
``jax.grad`` over the constrained ``nnx.State`` returns
(``tests/test_objectives.py:68-85``).
"""

from __future__ import annotations

import torch

from . import config


def _as_tensor(value) -> torch.Tensor:
    if isinstance(value, Parameter):
        value = value.value
    t = torch.as_tensor(value, dtype=torch.float64)
    dev = config.device if torch.cuda.is_available() else "cpu"
    return t.detach().clone().to(dev).requires_grad_(True)


class Parameter:
    tag = "parameter"

    def __init__(self, value, tag: str | None = None):
        self.value = _as_tensor(value)
        if tag is not None:
            self.tag = tag

    def __repr__(self) -> str:
        return f"{type(self).__name__}({self.value.detach().cpu().numpy()!r})"

    def zero_grad(self) -> None:
        self.value.grad = None


class NonNegativeReal(Parameter):
    tag = "non_negative"

    def __init__(self, value, tag: str = "non_negative"):
        super().__init__(value, tag)
        if bool((self.value.detach() < 0).any()):
            raise ValueError("NonNegativeReal value must be >= 0")


class PositiveReal(Parameter):
    tag = "positive"

    def __init__(self, value, tag: str = "positive"):
        super().__init__(value, tag)
        if bool((self.value.detach() <= 0).any()):
            raise ValueError("PositiveReal value must be > 0")


class Real(Parameter):
    tag = "real"

    def __init__(self, value, tag: str = "real"):
        super().__init__(value, tag)


def softplus(x: torch.Tensor) -> torch.Tensor:
    """log(1 + e^x) as max(x, 0) + log1p(e^-|x|): no overflow, and the same
    formula as the device Adam step (``gpk_adam_step``)."""
    return torch.clamp(x, min=0.0) + torch.log1p(torch.exp(-x.abs()))


def inverse_softplus(y: torch.Tensor) -> torch.Tensor:
    return y + torch.log(-torch.expm1(-y))


DEFAULT_BIJECTION = {
    "positive": (softplus, inverse_softplus),
    "non_negative": (softplus, inverse_softplus),
    "real": (lambda x: x, lambda y: y),
}
