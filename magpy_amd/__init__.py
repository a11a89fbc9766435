"""magpy_amd: MI355X-native stochastic Landau-Lifshitz-Gilbert ensemble integrator
with owlas/magpy's Model / EnsembleModel / Results interface.

The compute path is libmagpy_amd.so (HIP, gfx950) reached through the Cython
binding magpy_amd.core; importing this package fails if either is missing.
"""
from . import core
from .model import DOModel, EnsembleModel, Model
from .results import EnsembleResults, Results

__all__ = ["core", "Model", "EnsembleModel", "DOModel", "Results", "EnsembleResults"]
