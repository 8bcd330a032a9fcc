"""Process-wide defaults: device and factorisation precision.

GPJax runs float64 throughout (``jax_enable_x64``, ``tests/conftest.py:4``);
here ``precision`` picks between the float64 path ("fp64") and the
mixed-precision path ("mixed": float32 storage, BF16 split products on the
matrix cores, float64 diagonal blocks, log-determinant and refinement) for
every objective and predictive that takes a ``precision`` argument and is
called without one (conjugate_mll, conjugate_loocv, collapsed_elbo,
ConjugatePosterior.predict).
"""

from __future__ import annotations

device = "cuda"
precision = "fp64"
# Most float64 iterative-refinement steps for alpha = Sigma^-1 r (they stop
# early once the residual is below 1e-12); None = library default (1 for fp64,
# 6 for mixed).
refine_iters = None
# What a mixed-precision evaluation does when its a-posteriori check (result
# slot GPK_RES_FLAG: for the exact objectives the Rademacher probes of
# I - Sigma X^T X, for the collapsed ELBO the error model of its split-product
# statistics) says the result may be outside 1e-6 (objective) / 1e-4
# (gradient) of the float64 answer:
#   "fallback" recompute in float64 and return that (result slot FLAG = 2),
#   "raise"    raise objectives.MixedPrecisionError,
#   "warn"     return the mixed result with a warning,
#   "off"      no check (and no host sync for it).
# Under HIP-graph capture the check is skipped (it needs a host read).
mixed_check = "fallback"


def set_precision(p: str) -> None:
    """'fp64', 'mixed', or 'mixed:P,T,L' (products per stage, see objectives)."""
    global precision
    if p not in ("fp64", "mixed") and not (isinstance(p, str) and p.startswith("mixed:")):
        raise ValueError(f"precision must be 'fp64', 'mixed' or 'mixed:P,T,L', got {p!r}")
    precision = p
