"""Batched cycles of several sessions (gm_run_batch, the serving form of
run()): each session's results equal its results alone."""

import numpy as np
import pytest

from gpumap_amd import capi
from gpumap_amd import evaluation as ev

pytestmark = pytest.mark.gpu

RESULT_FIELDS = ["n_insertion_ops", "n_update_ops", "n_embedding_ops",
                 "n_inserted_instances", "n_update_changed_edges",
                 "embedding_pending_edges", "insertion_queue_size",
                 "update_queue_size", "embedding_queue_size",
                 "insertion_completed", "update_completed",
                 "embedding_completed"]


class Plan:
    """The bench's op plan for one session: run(chunk, chunk, chunk) per
    chunk, then run(0, N, N) until the update has completed and one full
    embedding sweep has run after it."""

    def __init__(self, engine, x, chunk):
        self.e, self.x, self.chunk = engine, x, chunk
        self.pos, self.done = 0, False

    def next_ops(self):
        n = len(self.x)
        ins = min(self.chunk, n - self.pos)
        if ins:
            self.e.append(self.x[self.pos:self.pos + ins])
            self.pos += ins
            return (ins, self.chunk, self.chunk)
        return (0, n, n)

    def record(self, ops, r):
        n = len(self.x)
        self.done = self.pos == n and r.update_completed and ops[2] >= n


def _sessions(specs):
    out = []
    for x, mode, k in specs:
        n, d = x.shape
        out.append((capi.Engine(n, d, n_neighbors=k, knn_mode=mode), x))
    return out


def _run_alone(specs, chunk):
    logs = []
    for e, x in _sessions(specs):
        p, log = Plan(e, x, chunk), []
        while not p.done:
            ops = p.next_ops()
            r = e.run(*ops)
            log.append([getattr(r, f) for f in RESULT_FIELDS])
            p.record(ops, r)
        logs.append((e, log))
    return logs


def _run_mixed(specs, chunk):
    plans = [Plan(e, x, chunk) for e, x in _sessions(specs)]
    logs = [[] for _ in plans]
    while not all(p.done for p in plans):
        live = [i for i, p in enumerate(plans) if not p.done]
        ops = [plans[i].next_ops() for i in live]
        rs = capi.Engine.run_batch([plans[i].e for i in live], ops)
        for i, o, r in zip(live, ops, rs):
            logs[i].append([getattr(r, f) for f in RESULT_FIELDS])
            plans[i].record(o, r)
    return [(p.e, log) for p, log in zip(plans, logs)]


def test_batched_cycles_equal_each_session_alone():
    # uneven sessions: sizes, dimensions, k and insertion methods differ, so
    # sessions finish at different cycles and leave the batch one by one
    specs = [(ev.gaussian10(n=2500, d=32), 2, 15),
             (ev.image_like(n=2000, d=120), 1, 15),
             (ev.gaussian10(n=1500, d=64, index=3), 0, 10),
             (ev.gaussian10(n=3000, d=16, index=4), 2, 15)]
    alone = _run_alone(specs, 500)
    mixed = _run_mixed(specs, 500)
    for (ea, la), (em, lm), (x, _, k) in zip(alone, mixed, specs):
        # the same result fields every cycle, the same graph bit for bit
        assert la == lm
        ia, da = ea.knn()
        im, dm = em.knn()
        assert np.array_equal(ia, im)
        assert np.array_equal(da.view(np.uint32), dm.view(np.uint32))
        # Hogwild SGD: the layouts agree in what they preserve
        ta = ev.trustworthiness(x, ea.embedding(), sample=1500)
        tm = ev.trustworthiness(x, em.embedding(), sample=1500)
        assert abs(ta - tm) < 0.01, (ta, tm)
        assert np.isfinite(em.embedding()).all()


def test_own_stream_sgd_after_a_batched_launch_in_the_same_cycle():
    # 16 sessions fill batched SGD launches (by job and wave count), which
    # are issued before the 17th session's op: a full sweep too large for
    # the batched kernel, run on the session's own stream.  That grid-wide
    # kernel must start after the batched ones (two at once could each hold
    # CUs the other's blocks wait for); the cycle ends without a barrier
    # time-out.
    small = [ev.gaussian10(n=600, d=8, index=10 + i) for i in range(17)]
    big = ev.gaussian10(n=3000, d=16, index=40)
    es = [capi.Engine(*x.shape, knn_mode=2) for x in small]
    for e, x in zip(es, small):
        e.append(x)
    eb = capi.Engine(*big.shape, knn_mode=2)
    eb.append(big)
    r = eb.run(3000, 3000, 0)
    assert r.update_completed
    order = es[:16] + [eb] + es[16:]
    ops = [(600, 0, 500)] * 16 + [(0, 0, 3000)] + [(600, 0, 500)]
    rs = capi.Engine.run_batch(order, ops)
    # the new rows' queue is drained whole (600 rows for 500 ops)
    assert [q.n_embedding_ops for q in rs] == [600] * 16 + [3000] + [600]
    for e in order:
        assert np.isfinite(e.embedding()).all()


def test_sessions_run_from_threads_equal_each_session_alone():
    # run() on several sessions from several threads at once (no Server):
    # the phases interleave on the device under the process-wide device
    # lock and every session's results equal its run alone
    import threading

    specs = [(ev.gaussian10(n=2000, d=24, index=20 + i), 2, 15)
             for i in range(3)]
    alone = _run_alone(specs, 400)
    sessions = _sessions(specs)
    logs = [None] * len(sessions)

    def drive(i):
        e, x = sessions[i]
        p, log = Plan(e, x, 400), []
        while not p.done:
            ops = p.next_ops()
            r = e.run(*ops)
            log.append([getattr(r, f) for f in RESULT_FIELDS])
            p.record(ops, r)
        logs[i] = log

    th = [threading.Thread(target=drive, args=(i,))
          for i in range(len(sessions))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for (ea, la), (em, _), lm in zip(alone, sessions, logs):
        assert la == lm
        ia, da = ea.knn()
        im, dm = em.knn()
        assert np.array_equal(ia, im)
        assert np.array_equal(da.view(np.uint32), dm.view(np.uint32))
        assert np.isfinite(em.embedding()).all()


def test_batch_rejects_duplicate_sessions():
    x = ev.gaussian10(n=600, d=8)
    e = capi.Engine(*x.shape)
    e.append(x)
    with pytest.raises(Exception):
        capi.Engine.run_batch([e, e], [(100, 100, 100), (100, 100, 100)])


def _params(n, d, k=15):
    from gpumap_amd import GPUMAPParams

    p = GPUMAPParams()
    p.n_instances, p.n_features, p.n_neighbors = n, d, k
    return p


def _drive(session, n, chunk=400):
    """A client: run_with_n_ops per chunk, then until the update completes,
    then one full sweep; returns the result log."""
    log, pos = [], 0
    while True:
        ins = min(chunk, n - pos)
        r = session.run_with_n_ops(ins, chunk if ins else n,
                                   chunk if ins else n)
        pos += ins
        log.append([getattr(r, f) for f in RESULT_FIELDS])
        if pos == n and r.update_completed and not ins:
            return log


def test_server_clients_in_threads_equal_sessions_alone():
    import threading

    import torch

    from gpumap_amd import GPUMAP
    from gpumap_amd.data_loader import ArrayDataLoader
    from gpumap_amd.serving import Server

    xs = [ev.gaussian10(n=1600, d=24, index=7),
          ev.image_like(n=1200, d=96, index=8),
          ev.gaussian10(n=2000, d=48, index=9)]
    alone = []
    for x in xs:
        m = GPUMAP(ArrayDataLoader(x), _params(*x.shape))
        log = _drive(m, len(x))
        alone.append((log, m.get_neighbor_indices().cpu().numpy(),
                      m.get_embedding().cpu().numpy()))
        m.close()
    with Server(max_sessions=8) as srv:
        sess = [srv.open(ArrayDataLoader(x), _params(*x.shape)) for x in xs]
        logs = [None] * len(xs)

        def client(i):
            logs[i] = _drive(sess[i], len(xs[i]))

        th = [threading.Thread(target=client, args=(i,))
              for i in range(len(xs))]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
        assert all(log is not None for log in logs)
        # the three clients shared cycles
        assert srv.cycles < sum(len(log) for log in logs)
        for i, x in enumerate(xs):
            assert logs[i] == alone[i][0]
            idx = sess[i].get_neighbor_indices().cpu().numpy()
            assert np.array_equal(idx, alone[i][1])
            emb = sess[i].get_embedding().cpu().numpy()
            ta = ev.trustworthiness(x, alone[i][2], sample=1000)
            tm = ev.trustworthiness(x, emb, sample=1000)
            assert abs(ta - tm) < 0.01, (i, ta, tm)
        for s in sess:
            s.close()
        assert srv.admission.sessions == 0 and srv.admission.used == 0
    torch.cuda.synchronize()


def test_server_admits_a_waiting_session_when_one_closes():
    import threading

    from gpumap_amd.data_loader import ArrayDataLoader
    from gpumap_amd.serving import Server, session_footprint

    x = ev.gaussian10(n=800, d=16, index=11)
    fp = session_footprint(800, 16, max_batch_rows=1000)
    with Server(hbm_budget_bytes=fp * 2 + fp // 2, max_sessions=8) as srv:
        a = srv.open(ArrayDataLoader(x), _params(800, 16))
        b = srv.open(ArrayDataLoader(x), _params(800, 16))
        admitted = []

        def late():
            admitted.append(srv.open(ArrayDataLoader(x), _params(800, 16)))

        t = threading.Thread(target=late)
        t.start()
        a.run_with_n_ops(400, 400, 400)
        t.join(timeout=0.2)
        assert not admitted            # no room for a third session yet
        a.close()
        t.join(timeout=60)
        assert admitted
        c = admitted[0]
        r1 = b.run_with_n_ops(400, 400, 400)
        r2 = c.run_with_n_ops(400, 400, 400)
        assert r1.n_inserted_instances == r2.n_inserted_instances == 400
        b.close()
        c.close()


def test_batched_launch_closed_mid_loop_waits_for_later_insertions():
    # 16 small sessions fill a batched SGD launch, which closes while the
    # loop has not yet reached two later sessions whose large FP32-tile
    # insertions were enqueued at the start of the cycle.  The launch waits
    # for them (its waves would otherwise wait at the epoch boundaries for
    # CUs the insertion kernels hold); the cycle completes and the large
    # sessions' results equal their run alone.
    small = [ev.gaussian10(n=600, d=8, index=60 + i) for i in range(16)]
    big = [ev.gaussian10(n=24_000, d=256, index=80 + i) for i in range(2)]
    es = [capi.Engine(*x.shape, knn_mode=2) for x in small]
    for e, x in zip(es, small):
        e.append(x)

    def big_sessions():
        out = []
        for x in big:
            e = capi.Engine(*x.shape, knn_mode=1)
            e.append(x[:20_000])
            e.run(20_000, 0, 0)
            e.append(x[20_000:])
            out.append(e)
        return out

    ref = big_sessions()
    ref_rs = [e.run(4000, 0, 0) for e in ref]
    bs = big_sessions()
    ops = [(600, 0, 600)] * 16 + [(4000, 0, 0)] * 2
    rs = capi.Engine.run_batch(es + bs, ops)
    assert [q.n_embedding_ops for q in rs[:16]] == [600] * 16
    for e in es:
        assert np.isfinite(e.embedding()).all()
    for r, q, e, er in zip(ref_rs, rs[16:], ref, bs):
        assert [getattr(r, f) for f in RESULT_FIELDS] == \
            [getattr(q, f) for f in RESULT_FIELDS]
        assert np.array_equal(e.knn()[0], er.knn()[0])


def _free_bytes():
    import torch

    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


def test_footprint_covers_what_a_session_allocates_and_release_frees_it():
    # hipMemGetInfo before and after creating a session and running one
    # insertion of the largest batch the footprint allows for: the estimate
    # is at least what the session took.  gm_release frees it once the
    # tensor exported by a getter is gone.
    import gc

    from gpumap_amd import GPUMAP
    from gpumap_amd.data_loader import ArrayDataLoader
    from gpumap_amd.serving import session_footprint

    x = ev.image_like(n=40_000, d=784, index=5)
    n, d = x.shape
    for mode in ("exact_bf16", "exact_fp32", "nndescent"):
        p = _params(n, d)
        p.knn_mode = mode
        code = {"nndescent": 0, "exact_fp32": 1, "exact_bf16": 2}[mode]
        f0 = _free_bytes()
        m = GPUMAP(ArrayDataLoader(x), p)
        m.run_with_n_ops(1000, 1000, 1000)
        m.run_with_n_ops(1000, 1000, 1000)
        used = f0 - _free_bytes()
        fp = session_footprint(n, d, 15, 2, code, max_batch_rows=1000)
        assert fp >= used, (mode, fp, used)
        emb = m.get_embedding()
        assert emb.shape == (2000, 2)
        m.release()
        assert not m.released               # the exported tensor holds it
        with pytest.raises(RuntimeError):
            m.run_with_n_ops(0, 0, 100)
        held = f0 - _free_bytes()
        del emb
        gc.collect()
        assert m.released
        freed = f0 - _free_bytes()
        assert held - freed >= 0.9 * used, (mode, held, freed, used)
        del m
        gc.collect()


def test_server_close_returns_the_memory_before_admitting():
    from gpumap_amd.data_loader import ArrayDataLoader
    from gpumap_amd.serving import Server

    x = ev.gaussian10(n=30_000, d=128, index=12)
    with Server(max_sessions=8) as srv:
        f0 = _free_bytes()
        a = srv.open(ArrayDataLoader(x), _params(*x.shape))
        a.run_with_n_ops(1000, 1000, 1000)
        held = f0 - _free_bytes()
        assert held > 30_000 * 128 * 4
        a.close()
        assert srv.admission.pending_release() == 0
        assert srv.admission.used == 0
        assert f0 - _free_bytes() < 0.1 * held
