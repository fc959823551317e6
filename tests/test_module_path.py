"""The reference's module path `gpumap` (python/gpumap/__init__.py:1-29):
every name and submodule the reference's users import resolves to the
gpumap_amd object itself.  CPU only (the engine's names are checked on the
GPU in test_gpu_facade.py)."""

import numpy as np
import pytest


def test_reference_names_and_submodules_resolve_to_the_engine_package():
    import gpumap
    import gpumap_amd
    from gpumap.data_loader import DataLoader, NumpyDataLoader
    from gpumap.metric import Metric
    from gpumap.schedule import (DefaultLinearEstimator, MeanRateEstimator,
                                 PersistentEstimator, WeightedStrategy)
    from gpumap.schedule.estimator import Estimator
    from gpumap.schedule.sequential import SequentialPriorityStrategy
    from gpumap.schedule.strategy import BaseStrategy
    from gpumap.schedule.weighted import WeightedStrategy as W2
    from gpumap.utils import _as_f32_c_contig_2d
    from gpumap.utils.as_f32_c_contig_2d import _as_f32_c_contig_2d as u2
    from gpumap.validators import validate_data_chunk
    from gpumap.validators.validate_data_chunk import validate_data_chunk as v2

    assert gpumap.__all__ == gpumap_amd.__all__
    assert W2 is WeightedStrategy is gpumap_amd.WeightedStrategy
    assert BaseStrategy is gpumap_amd.BaseStrategy
    assert SequentialPriorityStrategy is gpumap_amd.SequentialPriorityStrategy
    assert Estimator is gpumap_amd.Estimator
    assert (MeanRateEstimator, DefaultLinearEstimator, PersistentEstimator) == (
        gpumap_amd.MeanRateEstimator, gpumap_amd.DefaultLinearEstimator,
        gpumap_amd.PersistentEstimator)
    assert Metric is gpumap_amd.Metric and gpumap.Metric is Metric
    assert DataLoader is gpumap_amd.data_loader.DataLoader
    assert NumpyDataLoader is gpumap_amd.data_loader.NumpyDataLoader
    assert v2 is validate_data_chunk is gpumap_amd.validators.validate_data_chunk
    assert u2 is _as_f32_c_contig_2d is gpumap_amd.validators.as_f32_c_contig_2d
    # the functions behave as the reference's
    assert validate_data_chunk(np.zeros((3, 4)), 4) == 4
    with pytest.raises(ValueError):
        validate_data_chunk(np.zeros((3, 4)), 5)
    out = _as_f32_c_contig_2d(np.zeros((2, 3), np.float64).T)
    assert out.dtype == np.float32 and out.flags["C_CONTIGUOUS"]


def test_facade_module_is_the_engine_facade():
    import gpumap.gpumap as facade

    import gpumap_amd.gpumap

    assert facade is gpumap_amd.gpumap
    assert facade.find_ab_params(1.0, 0.1) == gpumap_amd.gpumap.find_ab_params(
        1.0, 0.1)
