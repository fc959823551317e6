"""The FP8 (E4M3, per-row power-of-two scales) screened insertion against
the FP32 tiles and the oracle, through the C ABI: the graph is the same bit
for bit on every workload shape, including rows whose magnitudes differ by
many orders (each row's own scale), rows in E4M3's subnormal range, zero
rows and rows too small to be scaled (left unscreened); the device's E4M3
rounding equals the host's."""

import numpy as np
import pytest

import oracle
from gpumap_amd import capi
from gpumap_amd import evaluation as ev
from tests.test_e4m3 import _reference

pytestmark = pytest.mark.gpu


def test_device_e4m3_rounding_equals_the_definition():
    g = np.random.default_rng(48)
    x = np.exp(g.uniform(np.log(2.0 ** -14), np.log(600.0), 50000))
    x = np.concatenate([x, -x, [0.0, 448.0, 464.0, 1e30]]).astype(np.float32)
    dev = capi.e4m3(x, on_device=True)
    assert np.array_equal(dev, capi.e4m3(x))
    assert np.array_equal(dev, _reference(x))


def _against_fp32_and_oracle(x, chunk, k=15, metric=0):
    n, d = x.shape
    e1 = capi.Engine(n, d, n_neighbors=k, metric=metric, knn_mode=1)
    e3 = capi.Engine(n, d, n_neighbors=k, metric=metric, knn_mode=3)
    o = oracle.Oracle(n, d, n_neighbors=k, metric=metric, n_threads=16,
                      knn_mode=1)
    for s in range(0, n, chunk):
        for m in (e1, e3, o):
            m.append(x[s:s + chunk])
            m.run(chunk, chunk, 0)
        (i1, d1), (i3, d3) = e1.knn(), e3.knn()
        assert np.array_equal(i1, i3) and np.array_equal(d1, d3), s
    i2, d2 = o.knn()
    assert np.array_equal(i3, i2) and np.array_equal(d3, d2)
    assert np.array_equal(np.sort(e3.pool(), 1), np.sort(o.pool(), 1))
    return e3


def test_fp8_screening_on_image_like_rows():
    x = ev.image_like(n=3000, d=784)
    e = _against_fp32_and_oracle(x, 500)
    e.set_profiling(True)


def test_fp8_screening_vmf_cosine_k30():
    x = ev.vmf50(n=3000, d=300)
    _against_fp32_and_oracle(x, 500, k=30, metric=1)


def test_fp8_per_row_scales_over_many_magnitudes():
    # each row scaled by 10^U(-15, 15): the scales are per row, so a row's
    # neighbours among rows of its own magnitude are screened as well as in
    # a uniform set; one zero row, rows below 2^-117 (unscreened), rows in
    # E4M3's subnormal range after scaling (a few large features, many tiny)
    g = np.random.default_rng(49)
    n, d = 1600, 72
    x = g.standard_normal((n, d)).astype(np.float64)
    mag = 10.0 ** g.uniform(-15, 15, size=(n, 1))
    x = x * mag
    x[5] = 0.0
    x[10:20] *= 1e-37 / np.abs(x[10:20]).max(1, keepdims=True)  # < 2^-117
    x[30:60, 1:] *= 1e-4                                # subnormal after scale
    x = x.astype(np.float32)
    assert np.isfinite(x).all()
    _against_fp32_and_oracle(x, 200)


def test_fp8_tiny_batches_duplicates_and_padding():
    g = np.random.default_rng(50)
    base = (g.standard_normal((120, 70)) * 5).astype(np.float32)  # D % 64 != 0
    x = np.concatenate([base, base[:40]], 0)
    x = x[g.permutation(len(x))]
    _against_fp32_and_oracle(x, 9)
