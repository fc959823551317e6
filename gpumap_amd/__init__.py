"""gpumap_amd -- progressive UMAP on AMD Instinct MI355X (gfx950).

Drop-in for the reference package ``gpumap`` (python/gpumap/__init__.py:1-29):
same exported names.  ``GPUMAP`` and ``GPUMAPParams`` come from the compiled
gfx950 engine and are loaded on first use; importing them without the built
extension raises ImportError (there is no CPU fallback).
"""

from __future__ import annotations

from .metric import Metric
from .schedule import (
    BaseStrategy,
    DefaultLinearEstimator,
    Estimator,
    MeanRateEstimator,
    PersistentEstimator,
    SequentialPriorityStrategy,
    WeightedStrategy,
)

__all__ = [
    "GPUMAP",
    "GPUMAPParams",
    "Metric",
    "BaseStrategy",
    "WeightedStrategy",
    "SequentialPriorityStrategy",
    "Estimator",
    "MeanRateEstimator",
    "DefaultLinearEstimator",
    "PersistentEstimator",
]


def __getattr__(name):
    if name == "GPUMAP":
        from .gpumap import GPUMAP

        return GPUMAP
    if name == "GPUMAPParams":
        from ._native import _core

        return _core.GPUMAPParams
    if name == "find_ab_params":
        from .gpumap import find_ab_params

        return find_ab_params
    raise AttributeError(f"module 'gpumap_amd' has no attribute {name!r}")
