"""Mean functions, mirroring ``gpjax/mean_functions.py:123-164``.

``Constant`` holds a plain value or a trainable ``Parameter``; ``Zero`` is the
fixed constant 0.  Only constant means are on the GPU path: the constant enters
the C ABI as the last slot of ``theta`` (r = y - c).
"""

from __future__ import annotations

import torch

from .parameters import Parameter


class Constant:
    def __init__(self, constant=0.0):
        if isinstance(constant, Parameter):
            self.constant = constant
        else:
            self.constant = torch.as_tensor(constant, dtype=torch.float64)
        self._on_device: dict = {}

    def value_tensor(self, device) -> torch.Tensor:
        if isinstance(self.constant, Parameter):
            return self.constant.value.reshape(()).to(device, torch.float64)
        # A fixed constant is copied to each device once, so that later calls
        # (e.g. inside a HIP-graph capture) issue no host-to-device copy.
        key = str(torch.device(device))
        if key not in self._on_device:
            self._on_device[key] = self.constant.reshape(()).to(device, torch.float64)
        return self._on_device[key]

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return torch.ones((x.shape[0], 1), dtype=torch.float64, device=x.device) * \
            self.value_tensor(x.device)


class Zero(Constant):
    def __init__(self):
        super().__init__(constant=0.0)
