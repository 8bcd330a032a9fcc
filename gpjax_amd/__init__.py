"""gpjax_amd: GPJax's exact-GP and collapsed-sparse-GP hot path on MI355X (gfx950).

User code shaped like GPJax's examples works against these modules:
``Dataset``, ``kernels.RBF/Matern12/Matern32/Matern52``,
``mean_functions.Zero/Constant``, ``likelihoods.Gaussian``,
``gps.Prior`` / ``ConjugatePosterior``, ``objectives.conjugate_mll``,
``fit`` (device-resident Adam, optionally replayed from a HIP graph).
All arithmetic runs in the HIP library ``libgpk.so`` (C ABI in include/gpk.h).
"""

from . import (batched, config, gps, kernels, likelihoods, linalg, mean_functions, objectives,
               parameters, variational_families)
from .dataset import Dataset
from .fit import fit

__version__ = "0.1.0"
__all__ = ["Dataset", "batched", "config", "fit", "gps", "kernels", "likelihoods", "linalg",
           "mean_functions", "objectives", "parameters", "variational_families"]
