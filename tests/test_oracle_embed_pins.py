"""Pins of the oracle's embedding-side steps and of its restatement of the
reference engine's orchestration, against what fixes each independently:

- the initial position of an inserted row (initalize_random_embeddings_kernel
  / initialize_embeddings_kernel, insertion_kernel.cuh:790-878): the textbook
  Box-Muller transform of the published Philox4x32-10 stream, or the
  exp(-(d - d_min)) weighted mean of its earlier neighbours' positions;
- the embedding noise (add_embedding_noise, embedding_kernel.cuh:953-1007):
  N(0, 0.001 x_range) with x_range the spread of column 0, drawn by
  Box-Muller from the stream keyed 42 ^ sweeps;
- the reposition of the rows outside an embedding op's batch
  (heuristic_neighbor_avg_kernel, :811-894): the membership-weighted mean of
  their embed-time neighbours' positions, in FP64 from the (pinned) sigma and
  rho;
- the result fields of a scripted run (impl.cuh:146-500): op counts, queue
  sizes, the embedding cursor and sweep count, the three completion flags,
  computed by hand below.
CPU only (the same script runs on the device in test_gpu_embed_parity.py)."""

import math

import numpy as np
import pytest

import oracle
from gpumap_amd import evaluation as ev

U32 = 2.0 ** -32
TAG_EMB_INIT, TAG_NOISE = 2, 3
FLT_MIN = float(np.finfo(np.float32).tiny)


def _philox(key64, ctr):
    return [int(v) for v in oracle.philox4x32(
        ctr, (key64 & 0xFFFFFFFF, key64 >> 32))]


def box_muller(w0, w1):
    """Textbook Box-Muller of two uniform 32-bit words: u1 in (0, 1],
    u2 in [0, 1)."""
    u1 = (w0 + 1) * U32
    u2 = w1 * U32
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def init_normals(row, c):
    """A new row's N(0, 1) start: key 42, counter (row, 0, block, 2)."""
    words = _philox(42, (row, 0, 0, TAG_EMB_INIT)) + \
        _philox(42, (row, 0, 1, TAG_EMB_INIT))
    return [box_muller(words[2 * d], words[2 * d + 1]) for d in range(c)]


def noise_normal(seed, t):
    """Noise draw of element t (row-major over the batch's coordinates):
    key seed, counter (t, t >> 32, 0, 3)."""
    w = _philox(seed, (t & 0xFFFFFFFF, t >> 32, 0, TAG_NOISE))
    return box_muller(w[0], w[1])


def membership(d, rho, sigma):
    """compute_membership_strength_kernel (embedding_kernel.cuh:214-250)."""
    if d - rho <= 0 or sigma == 0:
        return 1.0
    return max(math.exp(-(d - rho) / sigma), FLT_MIN)


def test_initial_positions_are_box_muller_draws_or_neighbour_means():
    n, d, k, c = 600, 8, 6, 2
    x = ev.gaussian10(n=n, d=d, index=31)
    o = oracle.Oracle(n, d, n_neighbors=k, n_components=c, knn_mode=1)
    o.append(x[:200])
    o.run(200, 0, 0)
    e0 = o.embedding()
    # first batch: no earlier rows, every start is a draw
    want = np.array([init_normals(r, c) for r in range(200)])
    np.testing.assert_allclose(e0, want, rtol=2e-6, atol=2e-6)
    assert abs(want.mean()) < 0.15 and abs(want.std() - 1) < 0.1
    o.append(x[200:400])
    o.run(200, 0, 0)
    e1 = o.embedding()
    idx, dist = o.knn()
    np.testing.assert_array_equal(e1[:200], e0)
    n_mean = 0
    for r in range(200, 400):
        earlier = [(int(j), float(dd)) for j, dd in zip(idx[r], dist[r])
                   if j < 200]
        if not earlier:
            np.testing.assert_allclose(e1[r], init_normals(r, c), rtol=2e-6,
                                       atol=2e-6)
            continue
        n_mean += 1
        dmin = min(dd for _, dd in earlier)
        wts = np.array([math.exp(-(dd - dmin)) for _, dd in earlier])
        pos = np.array([e0[j] for j, _ in earlier], np.float64)
        np.testing.assert_allclose(e1[r], wts @ pos / wts.sum(), rtol=1e-5,
                                   atol=1e-6)
    assert n_mean > 100


def _fuzzy_inputs(o, n, k):
    idx, dist = o.knn()
    _, _, _, sig, rho = oracle.fuzzy(idx, dist, k, True,
                                     np.arange(n, dtype=np.uint32), 0)
    return idx, dist, sig, rho


def test_noise_and_reposition_closed_forms():
    # max_epoch = 0: an embedding op is the noise on its batch rows and the
    # reposition of the others, with no SGD in between
    n, d, k, c = 600, 8, 6, 2
    x = ev.gaussian10(n=n, d=d, index=32)
    o = oracle.Oracle(n, d, n_neighbors=k, n_components=c, knn_mode=1,
                      max_epoch=0)
    o.append(x)
    r = o.run(n, n, 0)
    assert r.update_completed
    r = o.run(0, 0, n)                # a full sweep: counts sweep 1
    g = np.random.default_rng(33)
    e0 = g.uniform(-5, 5, size=(n, c)).astype(np.float32)
    o.set_embedding(e0)
    r = o.run(0, 0, 150)              # the cursor's next 150 rows: 0..149
    assert r.n_embedding_ops == 150
    e1 = o.embedding()
    # noise: N(0, 0.001 x_range) on the batch rows, key 42 ^ sweeps (1)
    scale = 0.001 * (float(e0[:, 0].max()) - float(e0[:, 0].min()))
    snap = e0.astype(np.float64).copy()
    for t in range(150 * c):
        row, dd = divmod(t, c)
        snap[row, dd] += noise_normal(42 ^ 1, t) * scale
    np.testing.assert_allclose(e1[:150], snap[:150], rtol=1e-6, atol=1e-6)
    # reposition: rows outside the batch, membership-weighted mean of their
    # embed-time neighbours (self and the first k - 1 stored, sqrt distances)
    idx, dist, sig, rho = _fuzzy_inputs(o, n, k)
    for i in range(150, n):
        nb = idx[i, :k - 1]
        dd = np.sqrt(dist[i, :k - 1].astype(np.float32)).astype(np.float64)
        m = np.array([membership(v, rho[i], sig[i]) for v in dd])
        want = m @ snap[nb] / m.sum()
        np.testing.assert_allclose(e1[i], want, rtol=1e-5, atol=1e-6)
    # the repositioned rows moved toward their neighbours
    assert (np.abs(e1[150:] - e0[150:]).sum(1) > 0).mean() > 0.95


# Scripted run of the reference engine (impl.cuh:146-500), N = 12, k = 3,
# exact insertion (the update only drains its queue), computed by hand:
#  1. append 3, run(3, 0, 0): rows 0-2 inserted; update queue {0, 1, 2}.
#  2. append 2, run(2, 2, 2): rows 3-4 inserted and every earlier pool
#     changes (none is full): queue 0-4; the update pops 2 (ops < n and
#     < queue); the embedding drains its new-row queue: all 5 appended rows.
#  3. append 7, run(7, 20, 3): rows 5-11; insertion complete (n = N); the
#     update takes all 12 rows (ops >= n), empties the queue and completes;
#     the embedding drains the 7 new rows (more than its 3 ops).
#  4. run(0, 0, 5): the cursor sweep starts at 0: rows 0-4.
#  5. run(0, 0, 8): rows 5-11 (the cursor passes N - 1 with the update
#     complete: sweep 1), then row 0.
#  6. run(0, 0, 12): ops >= n: every row, cursor N - 1: sweep 2.
#  7. run(0, 0, 12): sweep 3 > 2: the embedding completes (it ran).
#  8. run(0, 0, 12): completed, so no embedding op.
SCRIPT = [(3, (3, 0, 0)), (2, (2, 2, 2)), (7, (7, 20, 3)), (0, (0, 0, 5)),
          (0, (0, 0, 8)), (0, (0, 0, 12)), (0, (0, 0, 12)), (0, (0, 0, 12))]
FIELDS = ["n_insertion_ops", "n_update_ops", "n_embedding_ops",
          "insertion_completed", "update_completed", "embedding_completed",
          "n_inserted_instances", "insertion_queue_size", "update_queue_size",
          "embedding_queue_size", "embedding_pending_edges",
          "n_update_changed_edges"]
EXPECTED = [
    (3, 0, 0, 0, 0, 0, 3, 0, 3, 3, 3, 0),
    (2, 2, 5, 0, 0, 0, 5, 0, 3, 5, 5, 0),
    (7, 12, 7, 1, 1, 0, 12, 0, 0, 12, 12, 0),
    (0, 0, 5, 1, 1, 0, 12, 0, 0, 12, 12, 0),
    (0, 0, 8, 1, 1, 0, 12, 0, 0, 12, 12, 0),
    (0, 0, 12, 1, 1, 0, 12, 0, 0, 12, 12, 0),
    (0, 0, 12, 1, 1, 1, 12, 0, 0, 12, 12, 0),
    (0, 0, 0, 1, 1, 1, 12, 0, 0, 12, 12, 0),
]


def scripted_run(engine):
    x = ev.gaussian10(n=12, d=2, index=34)
    pos, out = 0, []
    for app, ops in SCRIPT:
        if app:
            engine.append(x[pos:pos + app])
            pos += app
        r = engine.run(*ops)
        out.append(tuple(int(getattr(r, f)) for f in FIELDS))
    return out


def test_orchestration_follows_the_hand_computed_sequence():
    o = oracle.Oracle(12, 2, n_neighbors=3, knn_mode=1, max_epoch=0)
    got = scripted_run(o)
    for step, (g, w) in enumerate(zip(got, EXPECTED)):
        assert g == w, (step + 1, dict(zip(FIELDS, g)), dict(zip(FIELDS, w)))
