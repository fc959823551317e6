"""The allocation-free SGD entry point (gm_sgd_noalloc) against the
allocating one (gm_debug_sgd) through the C ABI: the same kernels on the same
inputs give the same layout, on two calls in a row (the second smaller, so
that what the first left in the shared scratch must not leak into it), in
all three kernel modes; sizes beyond the scratch and out-of-range indices
are refused; a call allocates no device memory."""

import numpy as np
import pytest

import oracle  # noqa: F401  (the oracle-side case builder below uses it)
from gpumap_amd import capi
from gpumap_amd import evaluation as ev
from tests.test_gpu_embed_parity import _conflict_free_case

pytestmark = pytest.mark.gpu

A, B = 1.577, 0.895


@pytest.mark.parametrize("mode", ["auto", "per_epoch", "stream"])
def test_noalloc_equals_the_allocating_entry_point_on_two_calls(mode):
    kw = dict(per_epoch=mode == "per_epoch", stream=mode == "stream")
    for n, seed0 in ((1 << 16, 100), (1 << 13, 400)):
        emb, head, tail, w, seed, _ = _conflict_free_case(n=n, seed0=seed0)
        print(f"case n={n} seed={seed} mode={mode}", flush=True)
        want = capi.sgd(emb, head, tail, w, 30, 5, A, B, seed=seed, **kw)
        got = capi.sgd_noalloc(emb, head, tail, w, 30, 5, A, B, seed=seed,
                               **kw)
        assert np.abs(want[:16] - emb[:16]).max() > 0.5
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


def test_noalloc_on_a_fuzzy_graph_matches_in_distribution():
    # many edges (Hogwild, so not bit for bit): the same layout quality; a
    # larger call followed by a smaller one
    import torch

    x = ev.gaussian10(n=4000, d=16, index=41)
    e = capi.Engine(4000, 16, knn_mode=1)
    e.append(x)
    e.run(4000, 4000, 0)
    for n in (4000, 1500):
        batch = np.arange(n, dtype=np.uint32)
        h, t, w, _, _ = e.fuzzy(batch)
        keep = (h < n) & (t < n)
        h, t, w = h[keep], t[keep], w[keep]
        g = np.random.default_rng(42)
        emb0 = g.uniform(-10, 10, size=(n, 2)).astype(np.float32)
        want = capi.sgd(emb0, h, t, w, 200, 5, A, B)
        got = capi.sgd_noalloc(emb0, h, t, w, 200, 5, A, B)
        tw = ev.trustworthiness(x[:n], want, sample=min(n, 2000))
        tg = ev.trustworthiness(x[:n], got, sample=min(n, 2000))
        print(f"fuzzy graph n={n}: trust {tg:.4f} vs {tw:.4f}", flush=True)
        assert abs(tg - tw) <= 0.01, (n, tg, tw)
    torch.cuda.synchronize()


def test_noalloc_refuses_sizes_beyond_its_scratch_and_bad_indices():
    import torch

    dev = torch.device("cuda")
    emb = torch.zeros((8, 2), dtype=torch.float32, device=dev)
    h = torch.tensor([0, 1], dtype=torch.int32, device=dev)
    t = torch.tensor([1, 9], dtype=torch.int32, device=dev)   # 9 >= n
    w = torch.ones(2, dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError, match="out of range"):
        capi.sgd_noalloc_device(emb, h, t, w, 10, 5, A, B)
    assert (emb == 0).all()
    big = torch.zeros(((1 << 20) + 1, 2), dtype=torch.float32, device=dev)
    t2 = torch.tensor([1, 0], dtype=torch.int32, device=dev)
    with pytest.raises(RuntimeError, match="at most"):
        capi.sgd_noalloc_device(big, h, t2, w, 10, 5, A, B)
    many = torch.zeros((1 << 20) + 1, dtype=torch.int32, device=dev)
    wm = torch.ones((1 << 20) + 1, dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError, match="at most"):
        capi.sgd_noalloc_device(emb, many, many, wm, 10, 5, A, B)


def test_noalloc_allocates_no_device_memory():
    import torch

    emb, head, tail, w, seed, _ = _conflict_free_case(n=1 << 13, seed0=400)
    dev = torch.device("cuda")
    e = torch.from_numpy(emb.copy()).to(dev)
    h = torch.from_numpy(head).to(dev)
    t = torch.from_numpy(tail).to(dev)
    ww = torch.from_numpy(w).to(dev)
    capi.sgd_noalloc_device(e, h, t, ww, 30, 5, A, B, seed=seed)  # warm up
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for mode in (dict(), dict(per_epoch=True), dict(stream=True)):
        capi.sgd_noalloc_device(e, h, t, ww, 30, 5, A, B, seed=seed, **mode)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
