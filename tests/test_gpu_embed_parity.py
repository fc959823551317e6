"""Device parity of the embedding-side kernels against the CPU oracle, through
the C ABI: the initial positions of inserted rows (k_init_embedding), the
embedding noise (k_minmax_col0, k_noise), the reposition of rows outside a
batch (k_reposition), the engine's result fields on a hand-computed script,
the SGD with negative samples and its rounding grid at head degrees above 1.

Tolerances: the device's Box-Muller and gradient use the hardware log, cos,
exp2 and reciprocal (a few ulp); everything else is the same arithmetic."""

import numpy as np
import pytest

import oracle
from gpumap_amd import capi
from gpumap_amd import evaluation as ev
from tests.test_oracle_embed_pins import EXPECTED, FIELDS, scripted_run
from tests.test_oracle_sgd_pins import W_REL

pytestmark = pytest.mark.gpu


def test_orchestration_follows_the_hand_computed_sequence():
    e = capi.Engine(12, 2, n_neighbors=3, knn_mode=1, max_epoch=0)
    got = scripted_run(e)
    for step, (g, w) in enumerate(zip(got, EXPECTED)):
        assert g == w, (step + 1, dict(zip(FIELDS, g)), dict(zip(FIELDS, w)))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_initial_positions_match_oracle(mode):
    # insertion-only cycles: each batch's rows start from the device's draws
    # or neighbour means; the graphs are equal, so the positions must be
    x = ev.gaussian10(n=2400, d=16, index=35)
    n, d = x.shape
    e = capi.Engine(n, d, knn_mode=mode)
    o = oracle.Oracle(n, d, n_threads=8, knn_mode=min(mode, 1))
    for s in range(0, n, 400):
        for m in (e, o):
            m.append(x[s:s + 400])
            m.run(400, 400, 0)
        i1, _ = e.knn()
        i2, _ = o.knn()
        assert np.array_equal(i1, i2), s
        np.testing.assert_allclose(e.embedding(), o.embedding(), rtol=1e-5,
                                   atol=1e-5, err_msg=f"after row {s + 400}")


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
@pytest.mark.parametrize("metric", [0, 1])
def test_noise_and_reposition_match_oracle(metric, graph, monkeypatch):
    # max_epoch = 0: every embedding op is the noise on its batch and the
    # reposition of the other rows, deterministic on both sides; partial ops
    # (cursor sweeps, reposition active) and full sweeps, new-row queues.
    # The op's kernel sequences run as HIP graphs (updated in place from op
    # to op, re-instantiated when the batch size changes) or directly.
    monkeypatch.setenv("GPUMAP_AMD_GRAPH", "1" if graph else "0")
    x = ev.gaussian10(n=3000, d=24, index=36)
    n, d = x.shape
    e = capi.Engine(n, d, knn_mode=1, max_epoch=0, metric=metric)
    o = oracle.Oracle(n, d, n_threads=8, knn_mode=1, max_epoch=0,
                      metric=metric)
    g = np.random.default_rng(37)
    plan = [(500, 300), (500, 700), (500, 200), (0, 900), (1500, 100),
            (0, 3000), (0, 450), (0, 450), (0, 3000), (0, 1000)]
    pos = 0
    for cyc, (ins, emb_ops) in enumerate(plan):
        for m in (e, o):
            if ins:
                m.append(x[pos:pos + ins])
        pos += ins
        if cyc == 3:
            # a spread-out start so that the noise and the means are large
            e0 = g.uniform(-8, 8, size=(pos, 2)).astype(np.float32)
            e.set_embedding(e0)
            o.set_embedding(e0)
        r1, r2 = e.run(ins, n, emb_ops), o.run(ins, n, emb_ops)
        assert r1.n_embedding_ops == r2.n_embedding_ops
        assert r1.embedding_completed == r2.embedding_completed
        np.testing.assert_allclose(e.embedding(), o.embedding(), rtol=1e-5,
                                   atol=2e-5, err_msg=f"cycle {cyc}")
    gs = e.kernel_stats(capi.KSTAT_GRAPH)
    if graph:
        # two sequences per op with a reposition, fewer instantiations
        assert gs.launches >= len(plan) and 0 < gs.work[0] < gs.launches
    else:
        assert gs.launches == 0


def _conflict_free_case(n=1 << 16, seed0=100, c=2):
    """Eight edges on rows 0..15 with relative weights in W_REL-like ratios,
    positions spread wide (c components); a seed whose negative samples never
    land on an edge's rows (checked with the oracle's trace), so that the
    device's Hogwild order cannot matter."""
    g = np.random.default_rng(38)
    emb = g.uniform(-10, 10, size=(n, c)).astype(np.float32)
    emb[:16] = g.uniform(-2, 2, size=(16, c))
    head = np.arange(0, 16, 2, dtype=np.int32)
    tail = head + 1
    w = np.array((W_REL * 2)[:8], np.float32)
    for seed in range(seed0, seed0 + 50):
        _, evs, _ = oracle.sgd_trace(emb, head, tail, w, 30, 5, 1.577, 0.895,
                                     seed=seed)
        negs = evs[evs["kind"] != oracle.SGD_POSITIVE]
        if not (negs["other"] < 16).any():
            return emb, head, tail, w, seed, evs
    raise AssertionError("no conflict-free seed")


def test_sgd_with_negative_samples_matches_oracle():
    # one launch per epoch (the reference's launch pattern): each epoch reads
    # the rows as the previous one left them, as the oracle does; 30 epochs,
    # 5 negative samples per positive sample
    emb, head, tail, w, seed, evs = _conflict_free_case()
    assert (evs["kind"] == oracle.SGD_NEGATIVE).sum() > 300
    want = oracle.sgd(emb, head, tail, w, 30, 5, 1.577, 0.895, seed=seed)
    got = capi.sgd(emb, head, tail, w, 30, 5, 1.577, 0.895, seed=seed,
                   per_epoch=True)
    moved = np.abs(want[:16] - emb[:16]).max()
    assert moved > 0.5
    np.testing.assert_allclose(got[:16], want[:16], rtol=1e-4, atol=1e-4)
    np.testing.assert_array_equal(got[16:], emb[16:])


@pytest.mark.parametrize("kernel,epochs", [
    ("pipe", 30), ("pipe", 29), ("pipe", 1), ("pipe", 5), ("pipe_coop", 30),
    ("waves", 30), ("waves_coop", 30), ("stream", 30), ("pull", 30),
    ("pull", 29)])
def test_sgd_cooperative_kernels_match_oracle(kernel, epochs, monkeypatch):
    # the cooperative kernels (pipelined k_sgd_pipe, per-wave k_sgd_waves,
    # streaming k_sgd_stream, row-owner k_sgd_pull): k_sgd_pipe and
    # k_sgd_waves read an epoch's positions before the last epochs' updates
    # of their own edges have landed and add their own edges' steps of those
    # epochs, k_sgd_pull reads every row as the previous epoch left it, so
    # on the conflict-free set all follow the oracle's trajectory (29
    # epochs: k_sgd_pull's sums end in its second buffer; k_sgd_pipe: 1 and
    # 5 epochs are shorter than its pipeline, 29 and 30 end before its
    # padding to a multiple of 4); the per-cycle kernels with a plain launch
    # (default) and the cooperative one (GPUMAP_AMD_SGD_COOP=1)
    monkeypatch.setenv("GPUMAP_AMD_SGD_PIPE",
                       "0" if kernel.startswith("waves") else "1")
    monkeypatch.setenv("GPUMAP_AMD_SGD_COOP",
                       "1" if kernel.endswith("_coop") else "0")
    emb, head, tail, w, seed, _ = _conflict_free_case()
    want = oracle.sgd(emb, head, tail, w, epochs, 5, 1.577, 0.895, seed=seed)
    got = capi.sgd(emb, head, tail, w, epochs, 5, 1.577, 0.895, seed=seed,
                   stream=kernel == "stream", pull=kernel == "pull")
    np.testing.assert_array_equal(got[16:], emb[16:])
    print(f"max |device - oracle| {np.abs(got[:16] - want[:16]).max():.3g}",
          flush=True)
    np.testing.assert_allclose(got[:16], want[:16], rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("c,fix", [(1, True), (3, True), (4, True),
                                   (2, False), (3, False)])
def test_sgd_pipelined_kernel_matches_oracle_in_every_width(c, fix,
                                                            monkeypatch):
    # k_sgd_pipe's other instantiations: 1, 3 and 4 components on the
    # fixed-point records (64-bit buffer atomics per coordinate pair, a 32-bit
    # one for an odd last coordinate) and the float embedding's buffer float
    # atomics (GPUMAP_AMD_SGD_FIX=0); the conflict-free set, 30 epochs (20 in
    # one dimension: past about epoch 20 this 1-D trajectory turns the
    # device's few-ulp pow and reciprocal into 1.5e-3 in every device path,
    # one launch per epoch included)
    monkeypatch.setenv("GPUMAP_AMD_SGD_PIPE", "1")
    monkeypatch.setenv("GPUMAP_AMD_SGD_FIX", "1" if fix else "0")
    epochs = 20 if c == 1 else 30
    emb, head, tail, w, seed, _ = _conflict_free_case(c=c)
    want = oracle.sgd(emb, head, tail, w, epochs, 5, 1.577, 0.895, seed=seed)
    got = capi.sgd(emb, head, tail, w, epochs, 5, 1.577, 0.895, seed=seed)
    np.testing.assert_array_equal(got[16:], emb[16:])
    assert np.abs(want[:16] - emb[:16]).max() > 0.5
    print(f"max |device - oracle| {np.abs(got[:16] - want[:16]).max():.3g}",
          flush=True)
    np.testing.assert_allclose(got[:16], want[:16], rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("maxdeg,step", [(1, 2.0 ** -21), (3, 2.0 ** -20),
                                         (6, 2.0 ** -19)])
def test_sgd_rounding_grid_follows_the_maximum_head_degree(maxdeg, step):
    # edge (0, 1) at the maximum weight is sampled once (2 epochs, nsr 1);
    # maxdeg - 1 edges of head 5 with weights too small to be sampled before
    # the end set the maximum degree.  The rounding factor is the smallest
    # power of two above maxdeg * 4 / (1 - 2 maxdeg eps): 8, 16, 32, and the
    # updates are multiples of factor * 2^-24 (the floats just below it);
    # starting from coordinates on a 2^-8 grid, the moved rows land on the
    # finer grid and not on the next coarser one
    emb = np.zeros((12, 2), np.float32)
    emb[0] = [0.0, 0.0]
    emb[1] = [1.3125, -0.6875]
    head = np.array([0] + [5] * (maxdeg - 1), np.int32)
    tail = np.array([1] + list(range(6, 5 + maxdeg)), np.int32)
    w = np.array([1.0] + [1e-3] * (maxdeg - 1), np.float32)
    rf = oracle.rounding_factor(head, 12, 1.0)
    assert rf * 2.0 ** -24 == step
    for out in (capi.sgd(emb, head, tail, w, 2, 1, 1.577, 0.895),
                capi.sgd(emb, head, tail, w, 2, 1, 1.577, 0.895,
                         per_epoch=True),
                oracle.sgd(emb, head, tail, w, 2, 1, 1.577, 0.895)):
        dlt = (out[:2] - emb[:2]).astype(np.float64)
        assert (dlt != 0).all()
        q = dlt / step
        assert np.array_equal(q, np.round(q)), q
        assert not np.array_equal(q / 2, np.round(q / 2)), q
