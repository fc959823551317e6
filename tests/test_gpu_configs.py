"""The benchmark configurations under test at reduced size, against the CPU
oracle:

- the serving mix (Poisson arrivals of sessions of log-uniform sizes, dims in
  {32, 128, 784}, each looping run(target_latency=0.5) through one Server):
  every session's graph equals the oracle's and its run alone bit for bit
  under the same op sequence, and its embedding's trustworthiness is within
  0.01 of the oracle's;
- the 70,000 x 784 image-like workload's op plan on a prefix: the
  BF16-screened session's trustworthiness is within 0.01 of the oracle's."""

import threading
import time

import numpy as np
import pytest

import oracle
from gpumap_amd import evaluation as ev

pytestmark = pytest.mark.gpu


def _params(n, d, mode="exact_bf16"):
    from gpumap_amd import GPUMAPParams

    p = GPUMAPParams()
    p.n_instances, p.n_features = n, d
    p.knn_mode = mode
    return p


def _replay_oracle(x, ops_history, a, b):
    n, d = x.shape
    o = oracle.Oracle(n, d, a=a, b=b, n_threads=16, knn_mode=1)
    pos = 0
    for ni, nu, ne in ops_history:
        if ni:
            o.append(x[pos:pos + ni])
            pos += ni
        o.run(ni, nu, ne)
    return o


def test_serving_mix_sessions_equal_the_oracle_and_their_run_alone():
    from gpumap_amd import GPUMAP
    from gpumap_amd.data_loader import ArrayDataLoader
    from gpumap_amd.serving import Server

    g = np.random.default_rng(1004)
    n_s = 10
    sizes = np.exp(g.uniform(np.log(1e3), np.log(5e3), n_s)).astype(int)
    dims = g.choice([32, 128, 784], n_s)
    dims[:3] = [32, 128, 784]            # every dimension present
    arrive = np.cumsum(g.exponential(1.0 / 8.0, n_s))
    xs = [ev.gaussian10(n=int(s), d=int(d), index=200 + i)
          for i, (s, d) in enumerate(zip(sizes, dims))]
    print(f"serving mix: sizes {sizes.tolist()} dims {dims.tolist()}",
          flush=True)
    srv = Server(batch_window=0.001)
    sess = [None] * n_s
    logs = [None] * n_s
    done = [False] * n_s
    errs = []

    def client(i, t0):
        time.sleep(max(0.0, t0 + arrive[i] - time.perf_counter()))
        try:
            s = srv.open(ArrayDataLoader(xs[i]), _params(*xs[i].shape))
            sess[i] = s
            log = []
            for _ in range(400):
                r = s.run(target_latency=0.5)
                log.append(r)
                if (r.insertion_completed and r.update_completed
                        and r.embedding_completed):
                    done[i] = True
                    break
            logs[i] = log
        except BaseException as e:
            errs.append(f"session {i}: {e!r}")

    t0 = time.perf_counter()
    th = [threading.Thread(target=client, args=(i, t0)) for i in range(n_s)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=300)
    assert not errs, errs
    assert all(done), done
    assert srv.max_batch >= 2            # sessions shared cycles
    for i, (x, s) in enumerate(zip(xs, sess)):
        hist = s.ops_history
        assert len(hist) == len(logs[i])
        idx = s.get_neighbor_indices().cpu().numpy()
        dist = s.get_neighbor_dists().cpu().numpy()
        emb = s.get_embedding().cpu().numpy()
        p = s.params
        # the oracle replaying the session's op sequence
        o = _replay_oracle(x, hist, p.a, p.b)
        i2, d2 = o.knn()
        assert np.array_equal(idx, i2), i
        assert np.array_equal(dist.view(np.uint32), d2.view(np.uint32)), i
        # the same session alone
        m = GPUMAP(ArrayDataLoader(x), _params(*x.shape))
        for ops in hist:
            m.run_with_n_ops(*ops)
        assert np.array_equal(m.get_neighbor_indices().cpu().numpy(), idx), i
        m.close()
        tg = ev.trustworthiness(x, emb, sample=2000)
        to = ev.trustworthiness(x, o.embedding(), sample=2000)
        print(f"session {i} {x.shape}: {len(hist)} cycles, trust {tg:.4f} "
              f"vs oracle {to:.4f}", flush=True)
        assert abs(tg - to) <= 0.01, (i, tg, to)
        s.close()
    srv.close()


def test_image_like_op_plan_trustworthiness_matches_oracle():
    # the headline workload's data and op plan on its first 8,000 rows:
    # run(500, 500, 500) per chunk, then run(0, N, N) until the update has
    # completed, then one full embedding sweep
    from gpumap_amd import capi

    x = ev.image_like(n=8000, d=784)
    n, d = x.shape
    e = capi.Engine(n, d, knn_mode=2)
    o = oracle.Oracle(n, d, n_threads=16, knn_mode=1)
    pos = 0
    while pos < n:
        for m in (e, o):
            m.append(x[pos:pos + 500])
        pos += 500
        r1, r2 = e.run(500, 500, 500), o.run(500, 500, 500)
    while not r1.update_completed:
        r1, r2 = e.run(0, n, n), o.run(0, n, n)
    r1, r2 = e.run(0, 0, n), o.run(0, 0, n)
    i1, d1 = e.knn()
    i2, d2 = o.knn()
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2)
    tg = ev.trustworthiness(x, e.embedding(), sample=4000)
    to = ev.trustworthiness(x, o.embedding(), sample=4000)
    print(f"image-like 8000x784: trust GPU {tg:.4f} oracle {to:.4f}",
          flush=True)
    assert abs(tg - to) <= 0.01, (tg, to)
