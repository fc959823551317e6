"""GPU parity of the gfx950 engine against the CPU oracle, through the C ABI.

The kNN pipeline (insertion, NN-Descent update, public rows) is deterministic
by construction -- canonical FP32 distances, counter-based RNG, order-
independent lock-free pool inserts -- so the device graph must equal the
oracle's bit for bit after every run() call.  Fuzzy weights are compared at
1e-6 relative (FP64 on both sides, exp may differ by an ulp); the Hogwild SGD
is compared exactly on conflict-free edge sets and statistically
(trustworthiness) otherwise."""

import numpy as np
import pytest

import oracle
from gpumap_amd import evaluation as ev
from gpumap_amd import capi

pytestmark = pytest.mark.gpu

RESULT_FIELDS = ["n_insertion_ops", "n_update_ops", "n_embedding_ops",
                 "insertion_completed", "update_completed",
                 "embedding_completed", "n_inserted_instances",
                 "n_update_changed_edges", "embedding_pending_edges",
                 "insertion_queue_size", "update_queue_size",
                 "embedding_queue_size"]


def _lockstep(x, chunk, k=15, metric=0, n_ops=None, embed=0, max_cycles=80,
              check_every=1, mode=0, c=2):
    n, d = x.shape
    n_ops = n_ops or chunk
    e = capi.Engine(n, d, n_neighbors=k, n_components=c, metric=metric,
                    knn_mode=mode)
    # the oracle's NN-Descent (0) for both NN-Descent joins, else exact (1)
    o = oracle.Oracle(n, d, n_neighbors=k, n_components=c, metric=metric,
                      n_threads=8, knn_mode=0 if mode in (0, 4) else 1)
    pos, cyc = 0, 0
    while True:
        ins = min(chunk, n - pos)
        if ins:
            e.append(x[pos:pos + ins])
            o.append(x[pos:pos + ins])
            pos += ins
        upd = n_ops if ins else n
        r1, r2 = e.run(ins, upd, embed), o.run(ins, upd, embed)
        for f in RESULT_FIELDS:
            assert getattr(r1, f) == getattr(r2, f), (cyc, f)
        if cyc % check_every == 0 or r1.update_completed:
            i1, d1 = e.knn()
            i2, d2 = o.knn()
            assert np.array_equal(i1, i2), f"indices differ at cycle {cyc}"
            assert np.array_equal(d1.view(np.uint32), d2.view(np.uint32)), cyc
        cyc += 1
        if (r1.update_completed and pos == n) or cyc >= max_cycles:
            break
    assert r1.update_completed
    # pools are unordered slot arrays: equal as sets of keys (flags included)
    assert np.array_equal(np.sort(e.pool(), 1), np.sort(o.pool(), 1))
    return e, o


def test_philox_device_matches_published_vectors():
    from tests.test_oracle_pins import PHILOX_KAT

    for ctr, key, want in PHILOX_KAT:
        assert tuple(int(v) for v in capi.philox(ctr, key)) == want


MODES = [pytest.param(0, id="nndescent"), pytest.param(1, id="exact_fp32"),
         pytest.param(2, id="exact_bf16"), pytest.param(3, id="exact_fp8"),
         pytest.param(4, id="nndescent_bf16")]
EXACT_MODES = MODES[1:4]


@pytest.mark.parametrize("mode", MODES)
def test_knn_bit_exact_euclidean(mode):
    x = ev.gaussian10(n=3000, d=32)
    _lockstep(x, chunk=500, mode=mode)


@pytest.mark.parametrize("mode", MODES)
def test_knn_bit_exact_cosine(mode):
    x = ev.gaussian10(n=2000, d=40)
    _lockstep(x, chunk=400, metric=1, mode=mode)


@pytest.mark.parametrize("mode", MODES)
def test_knn_bit_exact_k30_odd_dim(mode):
    # k = 30 fills the 256-candidate cap; D = 37 exercises the padding
    g = np.random.default_rng(3)
    x = (g.standard_normal((2500, 37)) * 3).astype(np.float32)
    _lockstep(x, chunk=500, k=30, mode=mode)


@pytest.mark.parametrize("mode", MODES)
def test_knn_bit_exact_tiny_batches_and_duplicates(mode):
    # first batches smaller than the pool, exact duplicate rows (distance-0
    # ties broken by id), partial update ops drawn from the queue
    g = np.random.default_rng(4)
    base = g.standard_normal((150, 8)).astype(np.float32)
    x = np.concatenate([base, base[:50]], 0)
    x = x[g.permutation(len(x))]
    _lockstep(x, chunk=7, n_ops=40, max_cycles=200, mode=mode)


@pytest.mark.parametrize("mode", EXACT_MODES)
def test_exact_mode_is_exact_knn(mode):
    # the exact pools give the FP64 brute-force neighbours (tie-aware)
    x = ev.gaussian10(n=3000, d=32)
    e, o = _lockstep(x, chunk=700, mode=mode)
    idx, _ = e.knn()
    ei, ed = ev.exact_knn(x, 15)
    assert ev.recall_at_k(idx, x, ei, ed) == 1.0


def test_nndescent_bf16_join_on_image_like_rows():
    # 784 features (25 BF16 chunks, padding to 800) of the 70,000 x 784
    # config's image-like rows: the BF16-screened local join keeps the
    # oracle's NN-Descent graph every cycle, and its screen re-ranks fewer
    # pairs than the join computes
    x = ev.image_like(n=2500, d=784)
    e, o = _lockstep(x, chunk=500, mode=4, check_every=2)
    st = e.kernel_stats(0)
    print(f"join: members {st.work[0]} pairs {st.work[1]} offers "
          f"{st.work[2]} re-ranked {st.work[3]}", flush=True)
    assert 0 < st.work[3] < st.work[1]


def test_nndescent_bf16_equals_fp32_join_at_scale():
    # 20,000 x 96 in 1,000-row chunks through the two joins on the device:
    # equal public rows after every cycle, equal pools at the end
    x = ev.gaussian10(n=20_000, d=96)
    n, d = x.shape
    e0 = capi.Engine(n, d, knn_mode=0)
    e4 = capi.Engine(n, d, knn_mode=4)
    for eng in (e0, e4):
        eng.set_profiling(True)
    pos, cyc = 0, 0
    while True:
        ins = min(1000, n - pos)
        if ins:
            for eng in (e0, e4):
                eng.append(x[pos:pos + ins])
            pos += ins
        upd = 1000 if ins else n
        r0, r4 = e0.run(ins, upd, 0), e4.run(ins, upd, 0)
        for f in RESULT_FIELDS:
            assert getattr(r0, f) == getattr(r4, f), (cyc, f)
        i0, d0 = e0.knn()
        i4, d4 = e4.knn()
        assert np.array_equal(i0, i4), cyc
        assert np.array_equal(d0.view(np.uint32), d4.view(np.uint32)), cyc
        cyc += 1
        if (r0.update_completed and pos == n) or cyc > 200:
            break
    assert r0.update_completed
    assert np.array_equal(np.sort(e0.pool(), 1), np.sort(e4.pool(), 1))
    s0, s4 = e0.kernel_stats(0), e4.kernel_stats(0)
    assert s0.work[0] == s4.work[0] and s0.work[1] == s4.work[1]
    print(f"join FP32 {s0.device_ms:.1f} ms, BF16 {s4.device_ms:.1f} ms; "
          f"pairs {s4.work[1]}, re-ranked {s4.work[3]}", flush=True)
    e0.close()
    e4.close()


def test_vmf_cosine_k30_shape_bf16():
    # the 200,000 x 300 cosine config's shape at 4,000 rows: von Mises-Fisher
    # clusters with Zipf sizes, k = 30, BF16-screened insertion; the graph
    # equals the oracle's every cycle and the exact cosine kNN
    x = ev.vmf50(n=4000, d=300)
    e, o = _lockstep(x, chunk=500, k=30, metric=1, mode=2, check_every=2)
    idx, _ = e.knn()
    ei, ed = ev.exact_knn(x, 30, metric="cosine")
    assert ev.recall_at_k(idx, x, ei, ed, metric="cosine") == 1.0


def test_three_components_stream_shape():
    # the 2,000,000 x 128 streaming config's shape at 20,000 rows: 100
    # clusters, 2,000-row chunks, three components; the graph equals the
    # oracle's, and the embedding after the op plan's embedding ops agrees
    # with the oracle's in trustworthiness
    x = ev.stream100(n=20_000, d=128)
    e, o = _lockstep(x, chunk=2000, mode=2, c=3, embed=2000, check_every=4)
    for eng in (e, o):
        eng.run(0, 0, len(x))
    ge, oe = e.embedding(), o.embedding()
    assert ge.shape == (len(x), 3) and np.isfinite(ge).all()
    tg = ev.trustworthiness(x, ge, sample=5000)
    to = ev.trustworthiness(x, oe, sample=5000)
    assert abs(tg - to) <= 0.01, (tg, to)


def _fp64_union(idx, dist, sig, rho, batch, k):
    """FP64 fuzzy union of the batch rows from given sigma/rho (scan form)."""
    n = idx.shape[0]
    ei = np.concatenate([np.arange(n)[:, None], idx[:, :k - 1]], 1)
    ed = np.concatenate([np.zeros((n, 1)), np.sqrt(dist[:, :k - 1]).astype(np.float64)], 1)

    def memb(i, c):
        d = ed[i, c]
        if d - rho[i] <= 0:
            return 1.0
        return max(np.exp(-(d - rho[i]) / sig[i]), np.finfo(np.float32).tiny)

    out = {}
    for i in batch:
        for c in range(1, k):
            j = ei[i, c]
            p = memb(i, c)
            hit = np.nonzero(ei[j] == i)[0]
            q = memb(j, hit[0]) if hit.size else 0.0
            out[(int(i), int(j))] = p + q - p * q
            if not hit.size:
                out[(int(j), int(i))] = p + q - p * q
    return out, ed


@pytest.mark.parametrize("k", [15, 16, 17])
def test_fuzzy_set_matches_oracle_and_fp64(k):
    # sigma's bisection sums over 16-lane rows by DPP at k <= 16 and over the
    # 32-lane half-wave above
    x = ev.gaussian10(n=3000, d=32)
    e, o = _lockstep(x, chunk=1000, k=k, check_every=10, mode=0)
    idx, dist = e.knn()
    for batch in (np.arange(3000, dtype=np.uint32),
                  np.arange(5, 3000, 7, dtype=np.uint32)):
        h1, t1, w1, s1, r1 = e.fuzzy(batch)
        h2, t2, w2, s2, r2 = oracle.fuzzy(idx, dist, k, True, batch, 200)
        # rho is a stored distance: exact; sigma is fixed by the bisection
        # only to its stopping tolerance |sum - log2(k-1)| < 1e-5
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_allclose(s1, s2, rtol=2e-4)
        ref, ed = _fp64_union(idx, dist, s1, r1, batch, k)
        xs = ed[:, 2:] - r1[:, None]
        psum = np.where(xs > 0, np.exp(-np.maximum(xs, 0) / s1[:, None]), 1.0).sum(1)
        floor = np.isclose(s1, 1e-3 * ed.mean(1), rtol=1e-9)
        assert (np.abs(psum - np.log2(k - 1)) < 2e-5)[~floor].all()
        # weights: within 1e-5 relative of FP64 given the same sigma / rho
        got = {(int(a), int(b)): float(v) for a, b, v in zip(h1, t1, w1)}
        for key, v in got.items():
            assert v == pytest.approx(ref[key], rel=1e-5)
        # same edge set as the oracle's (trim threshold max_w / 200)
        assert set(got) == set(zip(h2.tolist(), t2.tolist()))


def test_sgd_single_edge_matches_closed_form():
    emb = np.array([[0.0, 0.0], [1.5, -2.0]], np.float32)
    out_g = capi.sgd(emb, [0], [1], np.array([1.0], np.float32), 2, 1, 1.577, 0.895)
    out_o = oracle.sgd(emb, [0], [1], np.array([1.0], np.float32), 2, 1, 1.577, 0.895)
    # hardware exp2/log2 in the device pow (about 2 ulp) can move the
    # gradient by one step of the 2^-20 truncation grid
    np.testing.assert_allclose(out_g, out_o, rtol=1e-5, atol=2e-6)
    assert not np.array_equal(out_g, emb)


def test_sgd_disjoint_edges_match_oracle():
    # disjoint pairs, attraction only (epochs = 2, nsr = 1): no write
    # conflicts, so the Hogwild kernel is deterministic here
    g = np.random.default_rng(5)
    n = 256
    emb = g.standard_normal((n, 3)).astype(np.float32)
    head = np.arange(0, n, 2, dtype=np.int32)
    tail = head + 1
    w = g.uniform(0.5, 1.0, size=head.size).astype(np.float32)
    w[0] = 1.0
    out_g = capi.sgd(emb, head, tail, w, 2, 1, 1.577, 0.895)
    out_o = oracle.sgd(emb, head, tail, w, 2, 1, 1.577, 0.895)
    np.testing.assert_allclose(out_g, out_o, rtol=1e-5, atol=4e-6)


def _twin_paired_wave():
    """64 COO slots (one wave) in the engine's layout: each directed edge
    (i, j) in an even slot, its reverse (j, i) or an empty slot after it;
    runs of equal i over the even slots, one of them across a 16-lane row."""
    spec = [(0, [10, 11, 12, 13]), (1, [14, 15]), (2, [16]),
            (3, [17, 18, 19, 20, 21, 22, 23, 24]), (4, [25, 26, 27, 10, 11]),
            (5, [28, 29, 30]), (6, [31, 32, 33, 14, 2, 3, 4, 34, 35])]
    g = np.random.default_rng(11)
    head, tail = [], []
    for i, js in spec:
        for j in js:
            head.append(i)
            tail.append(j)
            rev = g.random() < 0.6
            head.append(j if rev else -1)
            tail.append(i if rev else -1)
    assert len(head) == 64
    return np.array(head, np.int32), np.array(tail, np.int32)


@pytest.mark.parametrize("per_epoch,stream",
                         [(False, False), (True, False), (False, True)])
def test_sgd_twin_paired_aggregation_equals_per_edge_aggregation(per_epoch,
                                                                 stream):
    # one wave, every edge at the maximum weight with 2 epochs and nsr 1:
    # each edge is sampled once (epoch 1) with no negative samples and the
    # wave reads every row before it adds.  The contributions are multiples
    # of the rounding grid (2^-17 here) and the coordinates multiples of
    # 2^-10, so every sum is exact and the result independent of how the
    # contributions are grouped.  The twin-paired slots take the paired
    # aggregation (both edges' contributions to a row in one atomic); the
    # same edges compacted and ordered by tail take the per-head / per-tail
    # runs.
    h, t = _twin_paired_wave()
    w = np.where(h >= 0, 1.0, 0.0).astype(np.float32)
    live = h >= 0
    order = np.lexsort((h[live], t[live]))
    h2, t2 = h[live][order], t[live][order]
    assert not (h2[0] == t2[1] and t2[0] == h2[1])  # not twins: unpaired
    g = np.random.default_rng(12)
    emb = (np.round(g.standard_normal((40, 2)) * 1024) / 1024).astype(np.float32)
    paired = capi.sgd(emb, h, t, w, 2, 1, 1.577, 0.895, per_epoch=per_epoch,
                      stream=stream)
    plain = capi.sgd(emb, h2, t2, np.ones(h2.size, np.float32), 2, 1, 1.577,
                     0.895, per_epoch=per_epoch, stream=stream)
    np.testing.assert_array_equal(paired, plain)
    moved = np.unique(np.concatenate([h2, t2]))
    assert (paired[moved] != emb[moved]).any(axis=1).all()
    rest = np.setdiff1d(np.arange(40), moved)
    np.testing.assert_array_equal(paired[rest], emb[rest])


def test_bf16_queue_overflow_redoes_insertion_with_fp32_tiles():
    # 2 distinct points repeated 3000 times: all pairs within a group are at
    # distance 0, so every duplicate survives the screening; once a row has
    # more than 2048 of them its survivor list overflows and the batch is
    # redone by the FP32 tiles
    g = np.random.default_rng(9)
    x = np.repeat(g.normal(size=(2, 24)).astype(np.float32), 3000, axis=0)
    x = np.ascontiguousarray(x[g.permutation(len(x))])
    n, d = x.shape
    e1, e2 = capi.Engine(n, d, knn_mode=1), capi.Engine(n, d, knn_mode=2)
    e2.set_profiling(True)
    o = oracle.Oracle(n, d, n_threads=16, knn_mode=1)
    for s in range(0, n, 500):
        for e in (e1, e2, o):
            e.append(x[s:s + 500])
            e.run(500, 500, 0)
        i1, d1 = e1.knn()
        i2, d2 = e2.knn()
        i3, d3 = o.knn()
        assert np.array_equal(i2, i1) and np.array_equal(d2, d1), s
        assert np.array_equal(i2, i3) and np.array_equal(d2, d3), s
        assert np.array_equal(np.sort(e2.pool(), 1), np.sort(o.pool(), 1)), s
    st = e2.kernel_stats(capi.KSTAT_INSERT)
    assert st.work[2] >= 1, "no insertion overflowed a survivor list"
    assert st.work[2] < 12, "every insertion overflowed"


def test_bf16_screening_on_image_like_rows():
    # quantised sparse image-like rows (many exact-zero features, values
    # k/255 not representable in BF16): the BF16-screened graph equals the
    # FP32 tiles' and the oracle's
    x = ev.image_like(n=3000, d=784)
    n, d = x.shape
    engines = [capi.Engine(n, d, knn_mode=m) for m in (1, 2)]
    o = oracle.Oracle(n, d, n_threads=16, knn_mode=1)
    for s in range(0, n, 500):
        for e in engines:
            e.append(x[s:s + 500])
            e.run(500, 500, 0)
        o.append(x[s:s + 500])
        o.run(500, 500, 0)
    (i1, d1), (i2, d2) = engines[0].knn(), engines[1].knn()
    i3, d3 = o.knn()
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2)
    assert np.array_equal(i1, i3) and np.array_equal(d1, d3)
    assert np.array_equal(np.sort(engines[1].pool(), 1), np.sort(o.pool(), 1))


@pytest.mark.parametrize("lb_store", [True, False])
def test_bf16_screening_survivor_paths(lb_store, monkeypatch):
    # the survivor lists come from the stored first-pass lower bounds, or,
    # when the store is not allocated, from a second pass over the tiles;
    # both give the FP32 tiles' graph
    if not lb_store:
        monkeypatch.setenv("GPUMAP_AMD_LBM_MAX_BYTES", "0")
    x = ev.image_like(n=2000, d=200)
    n, d = x.shape
    e1, e2 = capi.Engine(n, d, knn_mode=1), capi.Engine(n, d, knn_mode=2)
    for s in range(0, n, 400):
        for e in (e1, e2):
            e.append(x[s:s + 400])
            e.run(400, 400, 0)
    (i1, d1), (i2, d2) = e1.knn(), e2.knn()
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2)
    assert np.array_equal(np.sort(e1.pool(), 1), np.sort(e2.pool(), 1))


def test_screen_shape_tuner_times_candidates_and_keeps_the_graph():
    # the BF16 screen's launch shape (column splits, survivor listing) is
    # tuned per key on the first insertions; every shape gives the FP32
    # tiles' graph, and a key whose candidates are all timed uses the
    # fastest of them
    x = ev.image_like(n=6000, d=160, index=45)
    n, d = x.shape
    e1, e2 = capi.Engine(n, d, knn_mode=1), capi.Engine(n, d, knn_mode=2)
    for s in range(0, n, 250):
        for e in (e1, e2):
            e.append(x[s:s + 250])
            e.run(250, 250, 0)
    (i1, d1), (i2, d2) = e1.knn(), e2.knn()
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2)
    shapes = capi.screen_shapes()
    mine = [t for t in shapes if t["db"] == 160]
    assert mine, shapes
    tuned = 0
    for t in mine:
        cand = t["candidates"]
        assert len(cand) >= 1 and -1 <= t["best"] < len(cand)
        assert all(c["bsplit"] >= 1 for c in cand)
        if t["best"] >= 0:
            ms = [c["ms"] for c in cand]
            assert min(ms) >= 0 and ms[t["best"]] == min(ms), t
            tuned += 1
    # 24 insertions over the keys of n < 4096 and n >= 4096 (one batch tile)
    assert tuned >= 1, mine


def test_sgd_fused_equals_per_epoch_launches():
    # conflict-free edges, attraction only: the single cooperative launch and
    # the one-launch-per-epoch path compute the same epochs
    g = np.random.default_rng(6)
    n = 512
    emb = g.standard_normal((n, 2)).astype(np.float32)
    h = np.arange(0, n, 2, dtype=np.int32)
    t = h + 1
    w = g.uniform(0.5, 1, h.size).astype(np.float32)
    w[0] = 1
    fused = capi.sgd(emb, h, t, w, 2, 1, 1.577, 0.895)
    per = capi.sgd(emb, h, t, w, 2, 1, 1.577, 0.895, per_epoch=True)
    np.testing.assert_array_equal(fused, per)
    assert not np.array_equal(fused, emb)


def test_sgd_stream_equals_per_epoch_launches():
    # the streaming kernel (the engine's choice for edge sets beyond the
    # resident grid, forced here at a small size) computes the same epochs as
    # one launch per epoch: conflict-free edges of equal weight with 2 epochs
    # and nsr 1, so every edge takes exactly one attraction step (epoch 1)
    # and no negative sample; more slots than the grid has threads, so that
    # each wave walks several
    g = np.random.default_rng(16)
    n = 1 << 20
    emb = g.standard_normal((n, 2)).astype(np.float32)
    h = np.arange(0, n, 2, dtype=np.int32)
    t = h + 1
    w = np.ones(h.size, np.float32)
    stream = capi.sgd(emb, h, t, w, 2, 1, 1.577, 0.895, stream=True)
    per = capi.sgd(emb, h, t, w, 2, 1, 1.577, 0.895, per_epoch=True)
    np.testing.assert_array_equal(stream, per)
    assert (stream != emb).any(axis=1).mean() > 0.99


def test_sgd_stream_statistics_match_oracle():
    # the streaming kernel on a real fuzzy graph with negative sampling
    x = ev.gaussian10(n=2000, d=16)
    e, o = _lockstep(x, chunk=500, mode=1, check_every=10)
    batch = np.arange(2000, dtype=np.uint32)
    h, t, w, _, _ = e.fuzzy(batch)
    g = np.random.default_rng(8)
    emb0 = g.uniform(-10, 10, size=(2000, 2)).astype(np.float32)
    out_g = capi.sgd(emb0, h, t, w, 200, 5, 1.577, 0.895, stream=True)
    out_o = oracle.sgd(emb0, h, t, w, 200, 5, 1.577, 0.895)
    tg = ev.trustworthiness(x, out_g)
    to = ev.trustworthiness(x, out_o)
    assert abs(tg - to) <= 0.01, (tg, to)


def test_sgd_pull_equals_per_epoch_launches():
    # the row-owner kernel (the engine's choice for edge sets beyond the
    # resident grid) on the streaming test's conflict-free set: every edge
    # takes one attraction step, computed by both its rows' entries
    g = np.random.default_rng(16)
    n = 1 << 20
    emb = g.standard_normal((n, 2)).astype(np.float32)
    h = np.arange(0, n, 2, dtype=np.int32)
    t = h + 1
    w = np.ones(h.size, np.float32)
    pull = capi.sgd(emb, h, t, w, 2, 1, 1.577, 0.895, pull=True)
    per = capi.sgd(emb, h, t, w, 2, 1, 1.577, 0.895, per_epoch=True)
    np.testing.assert_array_equal(pull, per)
    assert (pull != emb).any(axis=1).mean() > 0.99


@pytest.mark.parametrize("c", [2, 3])
def test_sgd_pull_statistics_match_oracle_and_repeat_exactly(c):
    # the row-owner kernel on a real fuzzy graph with negative sampling: the
    # layout's trustworthiness within 0.01 of the oracle's, and the same
    # result bit for bit on a second run (every epoch reads the positions
    # the previous one left, and the integer sums do not depend on order)
    x = ev.gaussian10(n=2000, d=16)
    e, o = _lockstep(x, chunk=500, mode=1, check_every=10)
    batch = np.arange(2000, dtype=np.uint32)
    h, t, w, _, _ = e.fuzzy(batch)
    g = np.random.default_rng(8)
    emb0 = g.uniform(-10, 10, size=(2000, c)).astype(np.float32)
    out_g = capi.sgd(emb0, h, t, w, 200, 5, 1.577, 0.895, pull=True)
    again = capi.sgd(emb0, h, t, w, 200, 5, 1.577, 0.895, pull=True)
    np.testing.assert_array_equal(out_g, again)
    out_o = oracle.sgd(emb0, h, t, w, 200, 5, 1.577, 0.895)
    tg = ev.trustworthiness(x, out_g)
    to = ev.trustworthiness(x, out_o)
    assert abs(tg - to) <= 0.01, (tg, to)


def test_sgd_many_epochs_statistics_match_oracle():
    # a real fuzzy graph, full schedule with negative sampling: Hogwild runs
    # differ in detail, so compare what the layout preserves
    x = ev.gaussian10(n=2000, d=16)
    e, o = _lockstep(x, chunk=500, mode=1, check_every=10)
    idx, dist = e.knn()
    batch = np.arange(2000, dtype=np.uint32)
    h, t, w, _, _ = e.fuzzy(batch)
    g = np.random.default_rng(8)
    emb0 = g.uniform(-10, 10, size=(2000, 2)).astype(np.float32)
    out_g = capi.sgd(emb0, h, t, w, 200, 5, 1.577, 0.895)
    out_o = oracle.sgd(emb0, h, t, w, 200, 5, 1.577, 0.895)
    tg = ev.trustworthiness(x, out_g)
    to = ev.trustworthiness(x, out_o)
    assert abs(tg - to) <= 0.01, (tg, to)


@pytest.mark.parametrize("mode", MODES)
def test_gaussian10_quality_and_exact_graph(mode):
    """10,000 x 64 with the reference's op plan: 500-row insertions with
    run(500, 500, 500), then full passes until complete."""
    x = ev.gaussian10()
    n, d = x.shape
    e = capi.Engine(n, d, knn_mode=mode)
    o = oracle.Oracle(n, d, n_threads=16,
                      knn_mode=0 if mode in (0, 4) else 1)
    pos = 0
    while pos < n:
        e.append(x[pos:pos + 500])
        o.append(x[pos:pos + 500])
        pos += 500
        r1, r2 = e.run(500, 500, 500), o.run(500, 500, 500)
    for _ in range(60):
        if r1.update_completed:
            break
        r1, r2 = e.run(0, n, n), o.run(0, n, n)
    assert r1.update_completed and r2.update_completed
    i1, d1 = e.knn()
    i2, d2 = o.knn()
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2)
    r1, r2 = e.run(0, 0, n), o.run(0, 0, n)
    ei, ed = ev.exact_knn(x, 15)
    assert ev.recall_at_k(i1, x, ei, ed) >= 0.95
    t1 = ev.trustworthiness(x, e.embedding())
    t2 = ev.trustworthiness(x, o.embedding())
    assert np.isfinite(e.embedding()).all()
    assert abs(t1 - t2) <= 0.01, (t1, t2)
