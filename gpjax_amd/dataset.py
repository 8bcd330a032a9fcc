"""``Dataset`` container, mirroring ``gpjax/dataset.py:26-126``."""

from __future__ import annotations

import warnings
from dataclasses import dataclass

import torch


@dataclass
class Dataset:
    X: torch.Tensor | None = None
    y: torch.Tensor | None = None

    def __post_init__(self) -> None:
        if self.X is not None and self.y is not None and self.X.shape[0] != self.y.shape[0]:
            raise ValueError(
                "Inputs, X, and outputs, y, must have the same number of rows."
                f" Got X.shape={tuple(self.X.shape)} and y.shape={tuple(self.y.shape)}.")
        if self.X is not None and self.X.ndim != 2:
            raise ValueError(f"Inputs, X, must be a 2-dimensional array. Got X.ndim={self.X.ndim}.")
        if self.y is not None and self.y.ndim != 2:
            raise ValueError(f"Outputs, y, must be a 2-dimensional array. Got y.ndim={self.y.ndim}.")
        for name, a in (("X", self.X), ("y", self.y)):
            if a is not None and a.dtype != torch.float64:
                warnings.warn(f"{name} is not of type float64. Got {name}.dtype={a.dtype}. "
                              "This may lead to numerical instability.", stacklevel=2)

    def __repr__(self) -> str:
        return f"Dataset(Number of observations: {self.n} - Input dimension: {self.in_dim})"

    def __add__(self, other: "Dataset") -> "Dataset":
        X = torch.cat((self.X, other.X)) if self.X is not None and other.X is not None else None
        y = torch.cat((self.y, other.y)) if self.y is not None and other.y is not None else None
        return Dataset(X=X, y=y)

    @property
    def n(self) -> int:
        return self.X.shape[0]

    @property
    def in_dim(self) -> int:
        return self.X.shape[1]
