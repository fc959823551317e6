"""The reference's module path (python/gpumap/__init__.py:1-29): ``import
gpumap`` gives the gfx950 engine under the names and submodule paths the
reference's users import --

    gpumap.GPUMAP, gpumap.GPUMAPParams, gpumap.Metric, the strategies and
    estimators; gpumap.gpumap, gpumap._core, gpumap.data_loader,
    gpumap.metric, gpumap.schedule (.strategy, .estimator, .sequential,
    .weighted), gpumap.validators (.validate_data_chunk) and gpumap.utils
    (.as_f32_c_contig_2d, whose function is _as_f32_c_contig_2d).

Each of them is the gpumap_amd object itself, not a copy.  GPUMAP and
GPUMAPParams need the built extension (ImportError without it: there is no
CPU fallback); everything else is plain Python.
"""

from __future__ import annotations

import sys
import types

import gpumap_amd as _impl
from gpumap_amd import data_loader, metric, schedule, validators
from gpumap_amd.metric import Metric
from gpumap_amd.schedule import (
    BaseStrategy,
    DefaultLinearEstimator,
    Estimator,
    MeanRateEstimator,
    PersistentEstimator,
    SequentialPriorityStrategy,
    WeightedStrategy,
)

__all__ = list(_impl.__all__)


def _alias(name: str, module: types.ModuleType) -> None:
    sys.modules[f"{__name__}.{name}"] = module


_alias("data_loader", data_loader)
_alias("metric", metric)
_alias("schedule", schedule)
for _sub in ("strategy", "estimator", "sequential", "weighted"):
    _alias(f"schedule.{_sub}", getattr(schedule, _sub))

# the reference's validators/ and utils/ packages, one function module each
_vpkg = types.ModuleType(f"{__name__}.validators")
_vpkg.validate_data_chunk = validators.validate_data_chunk
_vpkg.__all__ = ["validate_data_chunk"]
_vmod = types.ModuleType(f"{__name__}.validators.validate_data_chunk")
_vmod.validate_data_chunk = validators.validate_data_chunk
_upkg = types.ModuleType(f"{__name__}.utils")
_upkg._as_f32_c_contig_2d = validators.as_f32_c_contig_2d
_upkg.__all__ = ["_as_f32_c_contig_2d"]
_umod = types.ModuleType(f"{__name__}.utils.as_f32_c_contig_2d")
_umod._as_f32_c_contig_2d = validators.as_f32_c_contig_2d
for _m in (_vpkg, _upkg):
    _m.__path__ = []  # packages: their one submodule is registered below
_alias("validators", _vpkg)
_alias("validators.validate_data_chunk", _vmod)
_alias("utils", _upkg)
_alias("utils.as_f32_c_contig_2d", _umod)
validators = _vpkg
utils = _upkg


def __getattr__(name):
    if name in ("GPUMAP", "GPUMAPParams", "find_ab_params"):
        return getattr(_impl, name)
    raise AttributeError(f"module 'gpumap' has no attribute {name!r}")
