"""Process-wide defaults: device and factorisation precision.

GPJax runs float64 throughout (``jax_enable_x64``, ``tests/conftest.py:4``);
here ``precision`` picks between the float64 path ("fp64") and the
mixed-precision path ("mixed": float32 storage, BF16 split products on the
matrix cores, float64 diagonal blocks, log-determinant and refinement).
"""

from __future__ import annotations

device = "cuda"
precision = "fp64"
# Float64 iterative-refinement steps for alpha = Sigma^-1 r; None = library
# default (1 for fp64, 4 for mixed).
refine_iters = None


def set_precision(p: str) -> None:
    """'fp64', 'mixed', or 'mixed:P,T,L' (products per stage, see objectives)."""
    global precision
    if p not in ("fp64", "mixed") and not (isinstance(p, str) and p.startswith("mixed:")):
        raise ValueError(f"precision must be 'fp64', 'mixed' or 'mixed:P,T,L', got {p!r}")
    precision = p
