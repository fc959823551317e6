"""Pins of the CPU oracle (oracle/oracle.cpp) against what fixes each of its
functions independently: published Philox vectors, FP64 arithmetic, an FP64
brute-force kNN and scipy.sparse algebra.  CPU only."""

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from gpumap_amd import evaluation as ev

# Random123 known-answer vectors for philox4x32-10 (kat_vectors file of the
# Random123 distribution): counter, key -> output.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2,
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
     (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(int(v) for v in oracle.philox4x32(ctr, key)) == want


def test_canonical_distance_matches_fp64():
    g = np.random.default_rng(0)
    for d in (1, 4, 37, 64, 784):
        a = g.standard_normal(d).astype(np.float32)
        b = g.standard_normal(d).astype(np.float32)
        exact = float(((a.astype(np.float64) - b) ** 2).sum())
        assert oracle.canon_dist(a, b) == pytest.approx(exact, rel=1e-5)
    z = np.zeros(8, np.float32)
    assert oracle.canon_dist(z, z) == 0.0


def test_canonical_distance_is_sequential_fma():
    # fma(t, t, acc) in float32 equals round32(t*t + acc) computed exactly in
    # float64 whenever that sum needs no more than 53 bits, which holds for
    # these small integers; a reordered or non-fused sum gives other bits on
    # the second vector.
    a = np.array([3.0, 1e-4, 2.0 ** -12, 7.0], np.float32)
    b = np.zeros(4, np.float32)
    acc = np.float32(0.0)
    for t in (a - b):
        acc = np.float32(np.float64(t) * np.float64(t) + np.float64(acc))
    assert oracle.canon_dist(a, b) == float(acc)


def _run_oracle(x, k=15, chunk=250, metric=0, threads=4):
    n, d = x.shape
    o = oracle.Oracle(n, d, n_neighbors=k, metric=metric, n_threads=threads)
    pos = 0
    while pos < n:
        o.append(x[pos:pos + chunk])
        pos += chunk
        r = o.run(chunk, chunk, 0)
    it = 0
    while not r.update_completed and it < 60:
        r = o.run(0, n, 0)
        it += 1
    return o, r


@pytest.fixture(scope="module")
def small_run():
    x = ev.gaussian10(n=2000, d=16)
    o, r = _run_oracle(x)
    idx, dist = o.knn()
    return x, o, r, idx, dist


def test_oracle_knn_recall_vs_fp64(small_run):
    x, o, r, idx, dist = small_run
    assert r.update_completed
    ei, ed = ev.exact_knn(x, 15)
    assert ev.recall_at_k(idx, x, ei, ed) >= 0.98


def test_oracle_knn_invariants(small_run):
    x, o, r, idx, dist = small_run
    n = x.shape[0]
    assert (idx < n).all()
    assert (idx != np.arange(n)[:, None]).all()
    for i in range(0, n, 97):
        assert len(set(idx[i].tolist())) == 15
        assert (np.diff(dist[i]) >= 0).all()
        for j in range(15):
            assert dist[i, j] == oracle.canon_dist(x[i], x[idx[i, j]])
    # stored distances are squared L2 (reference notebook test.ipynb:299)
    d64 = ((x[idx].astype(np.float64) - x[:, None, :]) ** 2).sum(-1)
    np.testing.assert_allclose(dist, d64, rtol=1e-5, atol=1e-6)


def test_oracle_pool_keys_are_canonical(small_run):
    x, o, r, idx, dist = small_run
    pool = o.pool()
    for i in range(0, x.shape[0], 131):
        for key in pool[i]:
            if key == np.uint64(0xFFFFFFFFFFFFFFFF):
                continue
            j = int(key) & 0x7FFFFFFF
            dbits = np.uint32(int(key) >> 32)
            assert dbits.view(np.float32) == np.float32(oracle.canon_dist(x[i], x[j]))


def _fp64_smooth(dk, k):
    """Independent FP64 smooth_knn_dist: rho, sigma of each embed-time row."""
    n = dk.shape[0]
    target = np.log2(k - 1)
    valid = np.isfinite(dk) & (dk >= 0) & (dk < np.finfo(np.float32).max / 2)
    dz = np.where(valid, dk, 0.0)
    mean_all = dk[valid].mean() if valid.any() else 1.0
    rho = np.zeros(n)
    sig = np.zeros(n)
    for i in range(n):
        nz = np.nonzero(dz[i] > 0)[0]
        if nz.size == 0:
            sig[i] = 1e-3 * mean_all
            continue
        rho[i] = dk[i, nz[0]]
        lo, hi, mid = 0.0, np.inf, 1.0
        for _ in range(64):
            x = dz[i, 2:] - rho[i]
            psum = np.where(x > 0, np.exp(-np.maximum(x, 0) / mid), 1.0).sum()
            if abs(psum - target) < 1e-5:
                break
            if psum > target:
                hi = mid
                mid = (lo + hi) / 2
            else:
                lo = mid
                mid = mid * 2 if np.isinf(hi) else (lo + hi) / 2
        sig[i] = max(mid, 1e-3 * dz[i].mean())
    return rho, sig


def test_oracle_fuzzy_vs_scipy_union(small_run):
    x, o, r, idx, dist = small_run
    n, k = idx.shape
    batch = np.arange(0, n, 3, dtype=np.uint32)
    head, tail, w, sig, rho = oracle.fuzzy(idx, dist, k, True, batch, 200)
    # embed-time kNN: self + first k-1 neighbours, sqrt of squared L2
    ei = np.concatenate([np.arange(n)[:, None], idx[:, :k - 1]], 1)
    # embed-time distances: float32 sqrt of the stored squared distance, as
    # the reference's sqrtf (embedding_kernel.cuh:627-629), then FP64
    ed = np.concatenate([np.zeros((n, 1)),
                         np.sqrt(dist[:, :k - 1].astype(np.float32)).astype(np.float64)], 1)
    rho64, sig64 = _fp64_smooth(ed, k)
    np.testing.assert_allclose(rho, rho64, rtol=1e-12)
    np.testing.assert_allclose(sig, sig64, rtol=1e-6)
    # bisection invariant: the sum hits log2(k-1) unless the floor is active
    x_ = ed[:, 2:] - rho[:, None]
    psum = np.where(x_ > 0, np.exp(-np.maximum(x_, 0) / sig[:, None]), 1.0).sum(1)
    floor = np.isclose(sig, 1e-3 * ed.mean(1), rtol=1e-12)
    assert (np.abs(psum - np.log2(k - 1)) < 2e-5)[~floor].all()
    # union by sparse algebra: W = P + P^T - P o P^T
    p = np.where(ed - rho[:, None] <= 0, 1.0, np.exp(-(ed - rho[:, None]) / sig[:, None]))
    p = np.maximum(p, np.finfo(np.float32).tiny)
    p[:, 0] = 0.0
    P = sp.csr_matrix((p.ravel(), (np.repeat(np.arange(n), k), ei.ravel())), shape=(n, n))
    P.eliminate_zeros()
    Wm = (P + P.T - P.multiply(P.T)).tocsr()
    inb = np.zeros(n, bool)
    inb[batch] = True
    expect = {}
    PT = P.T.tocsr()
    for i in batch:
        for j in ei[i, 1:]:
            wij = Wm[i, j]
            expect[(i, j)] = wij
            if PT[i, j] == 0:
                expect[(j, i)] = wij
    maxw = max(expect.values())
    keep = {e: v for e, v in expect.items() if np.float32(v) >= np.float32(maxw) / np.float32(200)}
    got = {(int(h), int(t)): float(v) for h, t, v in zip(head, tail, w)}
    assert set(got) == set(keep)
    for e, v in keep.items():
        assert got[e] == pytest.approx(v, rel=1e-6)


def test_oracle_sgd_single_edge_closed_form():
    a, b = 1.577, 0.895
    emb = np.array([[0.0, 0.0], [1.5, -2.0]], np.float32)
    # w = max_w -> epochs_per_sample 1: the edge is first sampled at epoch 1;
    # nsr = 1 puts the first negative sample at epoch 2, so epochs = 2 is
    # exactly one attraction step with alpha = alpha0
    out = oracle.sgd(emb, np.array([0]), np.array([1]), np.array([1.0], np.float32),
                     n_epochs=2, nsr=1, a=a, b=b, alpha0=1.0)
    y0, y1 = emb.astype(np.float32)
    d2 = np.float32(((y0 - y1) ** 2).sum())
    a32, b32 = np.float32(a), np.float32(b)
    attr = np.float32(-2.0) * a32 * b32 * np.float32(d2 ** (b32 - 1)) / (a32 * np.float32(d2 ** b32) + 1)
    g = np.clip(attr * (y0 - y1), -4, 4).astype(np.float32)
    # rounding factor 2^ceil-ish of 4*max_deg*alpha: max_deg = 1 -> 8
    r = np.float32(8.0)
    np.testing.assert_allclose(out[0], y0 + ((r + g) - r), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(out[1], y1 + ((r - g) - r), rtol=1e-6, atol=1e-7)


def test_oracle_embedding_trustworthiness(small_run):
    x, o, r, idx, dist = small_run
    r = o.run(0, 0, x.shape[0])
    emb = o.embedding()
    assert np.isfinite(emb).all()
    assert ev.trustworthiness(x, emb, sample=2000) > 0.90
