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


def test_batch_rejects_duplicate_sessions():
    x = ev.gaussian10(n=600, d=8)
    e = capi.Engine(*x.shape)
    e.append(x)
    with pytest.raises(Exception):
        capi.Engine.run_batch([e, e], [(100, 100, 100), (100, 100, 100)])
