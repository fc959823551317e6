"""Synthetic workloads and quality scores for the progressive UMAP engine.

Workloads (all float32, row-major, seeded with PCG64(1000 + index)):
  gaussian10   10,000 x 64: 10 centres ~ U[-10,10]^64, 1,000 points each,
               x = centre + N(0, I), rows shuffled
  image_like   70,000 x 784: 10 clusters of 7,000; per cluster a fixed mask of
               20 % active features with mean mu ~ Beta(2,5); active feature
               x = round(255 * clip(mu + (Beta(2,5) - 2/7) * 0.5, 0, 1)) / 255,
               inactive features exactly 0; rows shuffled
  vmf50        200,000 x 300 unit vectors, 50 von Mises-Fisher clusters
               (kappa = 50) with Zipf(1.2) sizes
  stream100    2,000,000 x 128, 100 Gaussian clusters, sigma ~ U[0.5, 2],
               centres ~ U[-20, 20]^128

Scores follow the reference notebook's method (tests/test.ipynb:265,299,
350-363): exact kNN without the self column, recall@k = mean |exact & approx|
/ k, stored distances compared with squared exact distances.  Recall here is
tie-aware: an approximate neighbour counts when its exact distance is not
above the exact k-th distance.
"""

from __future__ import annotations

import numpy as np


def _rng(index: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(1000 + index))


def gaussian10(n: int = 10_000, d: int = 64, n_clusters: int = 10,
               index: int = 0) -> np.ndarray:
    g = _rng(index)
    centres = g.uniform(-10.0, 10.0, size=(n_clusters, d))
    labels = np.repeat(np.arange(n_clusters), n // n_clusters)
    labels = np.concatenate([labels, np.arange(n - labels.size) % n_clusters])
    x = centres[labels] + g.standard_normal((n, d))
    perm = g.permutation(n)
    return np.ascontiguousarray(x[perm], dtype=np.float32)


def image_like(n: int = 70_000, d: int = 784, n_clusters: int = 10,
               index: int = 1) -> np.ndarray:
    g = _rng(index)
    n_active = int(round(0.2 * d))
    x = np.zeros((n, d), dtype=np.float32)
    per = n // n_clusters
    sizes = [per] * n_clusters
    sizes[-1] += n - per * n_clusters
    row = 0
    for c in range(n_clusters):
        mask = g.choice(d, size=n_active, replace=False)
        mu = g.beta(2.0, 5.0, size=n_active)
        noise = g.beta(2.0, 5.0, size=(sizes[c], n_active))
        v = np.clip(mu[None, :] + (noise - 2.0 / 7.0) * 0.5, 0.0, 1.0)
        x[row:row + sizes[c], mask] = (np.round(255.0 * v) / 255.0).astype(np.float32)
        row += sizes[c]
    perm = g.permutation(n)
    return np.ascontiguousarray(x[perm])


def vmf50(n: int = 200_000, d: int = 300, n_clusters: int = 50,
          kappa: float = 50.0, index: int = 2) -> np.ndarray:
    from scipy.stats import vonmises_fisher

    g = _rng(index)
    ranks = np.arange(1, n_clusters + 1, dtype=np.float64)
    share = ranks ** -1.2
    share /= share.sum()
    raw = share * n
    sizes = np.floor(raw).astype(np.int64)
    rem = n - sizes.sum()
    sizes[np.argsort(-(raw - sizes))[:rem]] += 1
    mus = g.standard_normal((n_clusters, d))
    mus /= np.linalg.norm(mus, axis=1, keepdims=True)
    parts = [vonmises_fisher(mus[c], kappa).rvs(int(sizes[c]), random_state=g)
             for c in range(n_clusters)]
    x = np.concatenate(parts, axis=0)
    x = x[g.permutation(n)]
    return np.ascontiguousarray(x, dtype=np.float32)


def stream100(n: int = 2_000_000, d: int = 128, n_clusters: int = 100,
              index: int = 4) -> np.ndarray:
    g = _rng(index)
    centres = g.uniform(-20.0, 20.0, size=(n_clusters, d)).astype(np.float32)
    sig = g.uniform(0.5, 2.0, size=n_clusters).astype(np.float32)
    labels = np.repeat(np.arange(n_clusters), n // n_clusters)
    x = np.empty((n, d), dtype=np.float32)
    step = 200_000
    for s in range(0, n, step):
        lab = labels[s:s + step]
        x[s:s + step] = centres[lab] + sig[lab, None] * g.standard_normal(
            (lab.size, d), dtype=np.float32)
    return x[g.permutation(n)]


WORKLOADS = {
    "gaussian10": gaussian10,
    "image_like": image_like,
    "vmf50": vmf50,
    "stream100": stream100,
}


def exact_knn(x: np.ndarray, k: int, queries: np.ndarray | None = None,
              metric: str = "euclidean", block: int = 2048, device=None):
    """Exact FP64 kNN (self excluded).  Returns (idx, sqdist) of the queries.

    Distances are squared L2 (Euclidean) or 2 - 2 cos (cosine), the stored
    distance convention of the engine.  ``device`` = a torch device runs the
    FP64 products there; otherwise numpy.
    """
    xd = np.asarray(x, dtype=np.float64)
    if metric == "cosine":
        xd = xd / np.maximum(np.linalg.norm(xd, axis=1, keepdims=True), 1e-300)
    q = np.arange(xd.shape[0]) if queries is None else np.asarray(queries)
    idx = np.empty((q.size, k), dtype=np.int64)
    dd = np.empty((q.size, k), dtype=np.float64)
    if device is not None:
        import torch

        X = torch.from_numpy(xd).to(device)
        n2 = (X * X).sum(1)
        for s in range(0, q.size, block):
            qi = torch.from_numpy(q[s:s + block]).to(device)
            D = n2[qi, None] + n2[None, :] - 2.0 * (X[qi] @ X.T)
            D[torch.arange(qi.numel(), device=device), qi] = float("inf")
            v, i = torch.topk(D, k, dim=1, largest=False)
            idx[s:s + block] = i.cpu().numpy()
            dd[s:s + block] = np.maximum(v.cpu().numpy(), 0.0)
        return idx, dd
    n2 = (xd * xd).sum(1)
    for s in range(0, q.size, block):
        qi = q[s:s + block]
        D = n2[qi, None] + n2[None, :] - 2.0 * (xd[qi] @ xd.T)
        D[np.arange(qi.size), qi] = np.inf
        part = np.argpartition(D, k, axis=1)[:, :k]
        pd = np.take_along_axis(D, part, 1)
        o = np.argsort(pd, axis=1, kind="stable")
        idx[s:s + block] = np.take_along_axis(part, o, 1)
        dd[s:s + block] = np.maximum(np.take_along_axis(pd, o, 1), 0.0)
    return idx, dd


def recall_at_k(approx_idx: np.ndarray, x: np.ndarray, exact_idx: np.ndarray,
                exact_d: np.ndarray, queries: np.ndarray | None = None,
                metric: str = "euclidean") -> float:
    """Tie-aware recall@k of approximate rows against the exact kNN."""
    k = exact_idx.shape[1]
    q = np.arange(exact_idx.shape[0]) if queries is None else np.asarray(queries)
    a = np.asarray(approx_idx)[q, :k].astype(np.int64)
    xd = np.asarray(x, dtype=np.float64)
    if metric == "cosine":
        xd = xd / np.maximum(np.linalg.norm(xd, axis=1, keepdims=True), 1e-300)
    valid = (a >= 0) & (a < xd.shape[0]) & (a != q[:, None])
    a_safe = np.where(valid, a, 0)
    diff = xd[a_safe] - xd[q][:, None, :]
    da = (diff * diff).sum(-1)
    kth = exact_d[:, k - 1:k]
    tol = 1e-6 * np.maximum(kth, 1e-30)
    hit = valid & (da <= kth + tol)
    # one hit per distinct id
    hits = 0
    for r in range(a.shape[0]):
        hits += len(set(a[r][hit[r]].tolist()))
    return hits / float(a.shape[0] * k)


def trustworthiness(x: np.ndarray, emb: np.ndarray, n_neighbors: int = 15,
                    sample: int = 10_000, seed: int = 0,
                    metric: str = "euclidean") -> float:
    """sklearn trustworthiness@n_neighbors on a seeded row subsample."""
    from sklearn.manifold import trustworthiness as tw

    n = x.shape[0]
    if n > sample:
        sel = np.sort(np.random.Generator(np.random.PCG64(seed)).choice(
            n, size=sample, replace=False))
        x, emb = x[sel], emb[sel]
    return float(tw(np.asarray(x, np.float64), np.asarray(emb, np.float64),
                    n_neighbors=n_neighbors, metric=metric))
