"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every function include/gpumap_c.h declares, the pybind11 _core mirrors the
reference's _core surface, and the product fails loudly without a GPU (no CPU
fallback)."""

import ctypes

import pytest

from gpumap_amd import capi

REFERENCE_ALL = [
    "GPUMAP", "GPUMAPParams", "Metric", "BaseStrategy", "WeightedStrategy",
    "SequentialPriorityStrategy", "Estimator", "MeanRateEstimator",
    "DefaultLinearEstimator", "PersistentEstimator",
]


def test_c_abi_exports_every_declared_symbol():
    names = capi.header_functions()
    assert "gm_create" in names and "gm_run" in names
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    """Field offsets of the ctypes mirrors equal the C compiler's."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>',
             f'#include "{capi.HEADER}"', "int main(void) {"]
    for cname, cls in (("gm_params", capi.GmParams),
                       ("gm_result", capi.GmResult), ("gm_view", capi.GmView)):
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run(
        [str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    for cname, cls in (("gm_params", capi.GmParams),
                       ("gm_result", capi.GmResult), ("gm_view", capi.GmView)):
        assert int(out[cname]) == ctypes.sizeof(cls)
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


def test_package_surface_matches_reference():
    import gpumap_amd

    assert sorted(gpumap_amd.__all__) == sorted(REFERENCE_ALL)
    for name in REFERENCE_ALL:
        assert getattr(gpumap_amd, name) is not None


def test_params_defaults_and_metric_property():
    from gpumap_amd import GPUMAPParams, Metric

    p = GPUMAPParams()
    assert (p.n_neighbors, p.n_components, p.negative_sample_rate,
            p.max_epoch) == (15, 2, 5, 200)
    assert p.min_dist == pytest.approx(0.1)
    assert (p.a, p.b, p.alpha, p.verbose) == (1.0, 1.0, 1.0, False)
    assert p.metric == "euclidean"
    assert p.knn_mode == "exact_bf16"
    p.knn_mode = "nndescent"
    assert p.knn_mode == "nndescent"
    p.knn_mode = "exact_bf16"
    assert p.knn_mode == "exact_bf16"
    with pytest.raises(ValueError):
        p.knn_mode = "approximate"
    p.metric = "CoSiNe"
    assert p.metric == "cosine"
    p.metric = Metric.Euclidean
    assert p.metric == "euclidean"
    p.metric = 1
    assert p.metric == "cosine"
    with pytest.raises(ValueError):
        p.metric = "manhattan"


def test_result_fields():
    from gpumap_amd._native import _core

    fields = ["n_insertion_ops", "n_update_ops", "n_embedding_ops",
              "insertion_time", "update_time", "embedding_time",
              "insertion_completed", "update_completed", "embedding_completed",
              "n_inserted_instances", "n_update_changed_edges",
              "embedding_pending_edges", "insertion_queue_size",
              "update_queue_size", "embedding_queue_size"]
    for f in fields:
        assert hasattr(_core.GPUMAPResult, f)
    assert _core.GPUMAPIterationResult is _core.GPUMAPResult


def test_find_ab_params_fixture():
    from gpumap_amd.gpumap import find_ab_params

    a, b = find_ab_params(1.0, 0.1)
    assert a == pytest.approx(1.5769434602697652, rel=1e-6)
    assert b == pytest.approx(0.8950608778515733, rel=1e-6)


def test_invalid_params_rejected_before_any_device_call():
    from gpumap_amd import GPUMAPParams
    from gpumap_amd._native import _core

    p = GPUMAPParams()
    p.n_instances, p.n_features = 100, 8
    p.n_neighbors = 64
    with pytest.raises(ValueError):
        _core.GPUMAP(p)
    p.n_neighbors, p.n_components = 15, 7
    with pytest.raises(ValueError):
        _core.GPUMAP(p)
    # the embedding's byte size must stay below 2^31 (32-bit buffer offsets)
    p.n_components = 4
    p.n_instances = 1 << 27
    with pytest.raises(ValueError, match="n_components"):
        _core.GPUMAP(p)


def test_facade_min_dist_check():
    import numpy as np

    from gpumap_amd import GPUMAP, GPUMAPParams
    from gpumap_amd.data_loader import ArrayDataLoader

    p = GPUMAPParams()
    p.min_dist = 1.5
    with pytest.raises(ValueError, match="min_dist"):
        GPUMAP(ArrayDataLoader(np.zeros((10, 4), np.float32)), p)


def test_no_cpu_fallback_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present; this checks the no-GPU behaviour")
    from gpumap_amd import GPUMAPParams
    from gpumap_amd._native import _core

    p = GPUMAPParams()
    p.n_instances, p.n_features = 100, 8
    with pytest.raises(RuntimeError, match="device"):
        _core.GPUMAP(p)


def test_debug_sgd_rejects_unknown_mode_before_any_device_call():
    # gm_debug_sgd's mode: 0 the kernel that fits, 1 one launch per epoch,
    # 2 the streaming kernel, 3 the row-owner kernel; anything else is
    # refused by the argument check that runs before the engine touches the
    # device
    import ctypes

    import numpy as np

    from gpumap_amd import capi

    emb = np.zeros((4, 2), np.float32)
    head = np.array([0, 1], np.int32)
    tail = np.array([1, 0], np.int32)
    w = np.ones(2, np.float32)
    f32, i32 = ctypes.c_float, ctypes.c_int32
    for mode in (-1, 4):
        rc = capi.lib().gm_debug_sgd(
            capi._p(emb, f32), 4, 2, capi._p(head, i32), capi._p(tail, i32),
            capi._p(w, f32), 2, 2, 1, ctypes.c_float(1.577),
            ctypes.c_float(0.895), ctypes.c_float(1.0), ctypes.c_float(1.0),
            ctypes.c_uint64(42), mode)
        assert rc != 0
        assert "invalid SGD arguments" in capi.lib().gm_last_error().decode()
