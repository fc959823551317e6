"""The Python drop-in surface on the GPU: GPUMAP facade, DLPack getters,
error behaviour, run(target_latency) with the default strategy."""

import numpy as np
import pytest

import oracle
from gpumap_amd import evaluation as ev

pytestmark = pytest.mark.gpu


def _params(n, d, k=15, metric="euclidean"):
    from gpumap_amd import GPUMAPParams

    p = GPUMAPParams()
    p.n_instances, p.n_features, p.n_neighbors = n, d, k
    p.metric = metric
    return p


def test_facade_getters_are_device_views_equal_to_oracle():
    import torch

    from gpumap_amd import GPUMAP
    from gpumap_amd.data_loader import ArrayDataLoader

    x = ev.gaussian10(n=2000, d=24)
    m = GPUMAP(ArrayDataLoader(x), _params(2000, 24))
    o = oracle.Oracle(2000, 24, a=m.params.a, b=m.params.b, n_threads=8,
                      knn_mode=1)
    for s in range(0, 2000, 400):
        res = m.run_with_n_ops(400, 400, 0)
        o.append(x[s:s + 400])
        o.run(400, 400, 0)
    while not res.update_completed:
        res = m.run_with_n_ops(0, 2000, 0)
        o.run(0, 2000, 0)
    idx = m.get_neighbor_indices()
    dist = m.get_neighbor_dists()
    emb = m.get_embedding()
    assert idx.is_cuda and dist.is_cuda and emb.is_cuda
    assert idx.dtype == torch.uint32 and dist.dtype == torch.float32
    assert tuple(idx.shape) == (2000, 15) and tuple(emb.shape) == (2000, 2)
    i2, d2 = o.knn()
    assert np.array_equal(idx.cpu().numpy(), i2)
    assert np.array_equal(dist.cpu().numpy(), d2)
    # the exported tensor keeps the session's memory alive
    del m
    import gc

    gc.collect()
    assert np.array_equal(idx.cpu().numpy(), i2)


def test_capacity_and_feature_errors():
    from gpumap_amd._native import _core

    core = _core.GPUMAP(_params(100, 8))
    core.append_data(np.zeros((100, 8), np.float32))
    with pytest.raises(RuntimeError, match="Exceeding maximum number of rows"):
        core.append_data(np.zeros((1, 8), np.float32))
    with pytest.raises(RuntimeError, match="n_features mismatch"):
        core.append_data(np.zeros((1, 9), np.float32))
    with pytest.raises(RuntimeError, match="float32"):
        core.append_data(np.zeros((1, 8), np.float64))
    core.append_data(np.zeros((0, 8), np.float32))  # no-op
    core.close()


def test_run_target_latency_completes():
    from gpumap_amd import GPUMAP
    from gpumap_amd.data_loader import ArrayDataLoader

    x = ev.gaussian10(n=3000, d=16)
    m = GPUMAP(ArrayDataLoader(x), _params(3000, 16))
    done, cycles = False, 0
    while not done and cycles < 300:
        r = m.run(target_latency=0.2)
        done = r.insertion_completed and r.update_completed and r.embedding_completed
        cycles += 1
    assert done, cycles
    assert m.n_inserted_instances == 3000
    emb = m.get_embedding().cpu().numpy()
    assert np.isfinite(emb).all()
    assert ev.trustworthiness(x, emb, sample=3000) > 0.85
    m.close()


def test_reference_module_path_runs_the_engine():
    # `import gpumap` as the reference's users write it: the same session as
    # through gpumap_amd, and the same kNN graph as the oracle
    import gpumap
    from gpumap._core import GPUMAPParams
    from gpumap.data_loader import ArrayDataLoader

    x = ev.gaussian10(n=1200, d=16, index=44)
    p = GPUMAPParams()
    p.n_instances, p.n_features, p.n_neighbors = 1200, 16, 15
    m = gpumap.GPUMAP(ArrayDataLoader(x), p)
    assert isinstance(m.strategy, gpumap.WeightedStrategy)
    o = oracle.Oracle(1200, 16, a=m.params.a, b=m.params.b, n_threads=8,
                      knn_mode=1)
    for s in range(0, 1200, 400):
        res = m.run_with_n_ops(400, 400, 0)
        o.append(x[s:s + 400])
        o.run(400, 400, 0)
    i2, d2 = o.knn()
    assert np.array_equal(m.get_neighbor_indices().cpu().numpy(), i2)
    assert np.array_equal(m.get_neighbor_dists().cpu().numpy(), d2)
    m.close()
