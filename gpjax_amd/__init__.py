"""gpjax_amd: GPJax's exact-GP and collapsed-sparse-GP hot path on MI355X (gfx950).

User code shaped like GPJax's examples works against these modules:
``Dataset``, ``kernels.RBF/Matern12/Matern32/Matern52``,
``mean_functions.Zero/Constant``, ``likelihoods.Gaussian``,
``gps.Prior`` / ``ConjugatePosterior``, ``objectives.conjugate_mll``.
All arithmetic runs in the HIP library ``libgpk.so`` (C ABI in include/gpk.h).
"""

from . import (config, gps, kernels, likelihoods, mean_functions, objectives, parameters,
               variational_families)
from .dataset import Dataset

__version__ = "0.1.0"
__all__ = ["Dataset", "config", "gps", "kernels", "likelihoods", "mean_functions",
           "objectives", "parameters", "variational_families"]
