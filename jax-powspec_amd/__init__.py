"""MI355X-native power-spectrum estimator with the call surface of jax-powspec.

Modules
  mas           cic_mas_vec, cic_mas (reference src/mas.py), tsc_mas_vec, assign
  correlations  powspec_vec, powspec_vec_fundamental (reference src/correlations.py),
                estimate (particles -> P_l in one call), rfftn
  binning       exact integer k-shell thresholds (host side)

The compute path is libpowspec_amd.so (hand-written HIP for gfx950, C ABI in
include/pk_api.h); importing this package does not load it, the first GPU call does.
"""
from . import binning  # noqa: F401  (host-only; importable without a GPU)

__version__ = "0.1.0"


def __getattr__(name):
    # lazy: mas / correlations import torch
    if name in ("mas", "correlations"):
        import importlib
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
