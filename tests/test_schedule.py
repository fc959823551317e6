"""Known-answer tests of the scheduler and cost models (hand-computed values;
reference behaviour python/gpumap/schedule/*.py).  CPU only."""

import math
from types import SimpleNamespace

import numpy as np
import pytest

from gpumap_amd.schedule import (DefaultLinearEstimator, MeanRateEstimator,
                                 PersistentEstimator,
                                 SequentialPriorityStrategy, WeightedStrategy,
                                 clip_ops)
from gpumap_amd.schedule.strategy import _estimate_or_fallback, _jitter

Z95 = 1.6448536269514722  # standard normal 0.95 quantile


def res(ni=0, nu=0, ne=0, ti=0.0, tu=0.0, te=0.0, inserted=0, uq=0,
        pending=0, eq=0, done=(False, False, False)):
    return SimpleNamespace(
        n_insertion_ops=ni, n_update_ops=nu, n_embedding_ops=ne,
        insertion_time=ti, update_time=tu, embedding_time=te,
        n_inserted_instances=inserted, update_queue_size=uq,
        embedding_pending_edges=pending, embedding_queue_size=eq,
        insertion_completed=done[0], update_completed=done[1],
        embedding_completed=done[2])


def test_mean_rate_estimator():
    e = MeanRateEstimator()
    with pytest.raises(ValueError):
        e.estimate_ops(1.0)
    for ops, t in ((100, 1.0), (200, 2.2), (300, 2.7)):
        e.add_sample(ops, t)
    e.add_sample(0, 5.0)    # ignored
    e.add_sample(10, -1.0)  # ignored
    fr = e.fit_result()
    assert fr.n_samples == 3
    assert fr.mean_rate == pytest.approx(0.01)
    assert fr.sigma_rate == pytest.approx(0.001)
    assert e.estimate_ops(1.0) == int(1.0 / (0.01 + Z95 * 0.001))  # 85
    assert e.estimate_ops(1.0) == 85
    assert e.estimate_ops(0.0) == 0
    assert e.predict_time(100) == pytest.approx(100 * (0.01 + Z95 * 0.001))
    assert e.predict_time(0) == 0.0


def test_mean_rate_window():
    e = MeanRateEstimator(window_size=2, min_samples=1)
    e.add_sample(10, 10.0)
    e.add_sample(10, 1.0)
    e.add_sample(10, 1.0)
    assert e.fit_result().mean_rate == pytest.approx(0.1)
    assert e.fit_result().sigma_rate == 0.0


def test_linear_estimator_fit_and_cold_start():
    e = DefaultLinearEstimator()
    assert [e.estimate_ops(1.0) for _ in range(4)] == [1000, 2000, 3000, 3000]
    for ops in (1, 2, 3):
        e.add_sample(ops, 0.25 * ops + 0.5)
    fr = e.fit_result()
    assert (fr.slope_a, fr.intercept_b, fr.sigma) == pytest.approx((0.25, 0.5, 0.0))
    assert e.estimate_ops(10.5) == 40
    assert e.predict_time(40) == pytest.approx(10.5)


def test_linear_estimator_fallbacks():
    e = DefaultLinearEstimator()
    for ops, t in ((100, 2.0), (200, 1.0), (300, 0.5)):  # negative slope
        e.add_sample(ops, t)
    # slowest recent rate among runs >= 1 % of the largest: 2.0 / 100
    assert e.estimate_ops(1.0) == 50
    d = DefaultLinearEstimator()
    for _ in range(3):
        d.add_sample(10, 1.0)
    with pytest.raises(ValueError):
        d.estimate_ops(1.0)


def test_estimate_or_fallback_and_jitter():
    np.random.seed(123)
    expect = 100 + int(np.random.randint(-10, 10))
    np.random.seed(123)
    assert _jitter(100) == expect
    assert _estimate_or_fallback(MeanRateEstimator(), 0.0, 0.95, 100) == 0
    np.random.seed(123)
    assert _estimate_or_fallback(MeanRateEstimator(), 1.0, 0.95, 100) == expect
    assert clip_ops(5, (10, 20)) == 10 and clip_ops(50, (10, 20)) == 20


def test_weighted_strategy_split_and_allocate():
    s = WeightedStrategy(n_instances=10_000)
    assert s._split(1.1) == pytest.approx((0.25, 0.5, 0.25))
    for est, rate in ((s.insertion_estimator, 0.001), (s.update_estimator, 0.002),
                      (s.embedding_estimator, 0.0005)):
        for _ in range(3):
            est.add_sample(1000, 1000 * rate)
    ins, upd, emb = s.allocate(1.1)
    assert ins == 500  # 0.25 / 0.001 = 250 -> floor of min_insertion_ops
    assert upd == 250  # 0.5 / 0.002
    assert emb == 500  # 0.25 / 0.0005
    s.on_operation_complete("insertion")
    assert s.weights[0] == 0.0 and s.completeness[0]
    assert s._split(0.33) == pytest.approx((0.0, 0.2, 0.1))  # 0.33 * w / 3.3
    with pytest.raises(ValueError):
        s.on_operation_complete("nope")


def test_strategy_update_and_progress():
    s = WeightedStrategy(n_instances=1000)
    s.update(res(ni=500, ti=0.5, nu=100, tu=0.1, ne=200, te=0.2, inserted=500,
                 uq=50, pending=500, eq=500))
    assert s.get_progress() == pytest.approx((0.5, 0.9, 0.0))
    assert s.insertion_estimator.n_samples == 1
    s.on_operation_complete("insertion")
    s.update(res(ni=10, ti=0.01, inserted=1000, pending=1000, eq=1000))
    assert s.insertion_estimator.n_samples == 1  # frozen once complete
    assert s.update_estimator.n_samples == 1


def test_sequential_strategy_cascade():
    s = SequentialPriorityStrategy(n_instances=100)
    m = 2**32 - 1
    assert s.allocate(1.0) == (m, 0, 0)
    s.on_operation_complete("insertion")
    assert s.allocate(1.0) == (0, m, 0)
    s.on_operation_complete("update")
    assert s.allocate(1.0) == (0, 0, m)
    s.on_operation_complete("embedding")
    assert s.allocate(1.0) == (0, 0, m)


def test_base_min_insertion_default():
    from gpumap_amd.schedule.strategy import BaseStrategy

    class S(BaseStrategy):
        def allocate(self, t):
            return (0, 0, 0)

    assert S(n_instances=20_000).min_insertion_ops == 200


def test_persistent_estimator_roundtrip(tmp_path):
    db = str(tmp_path / "sub" / "est.sqlite")
    p = PersistentEstimator(MeanRateEstimator(), db, "update")
    for ops, t in ((100, 1.0), (200, 2.2), (300, 2.7)):
        p.add_sample(ops, t)
    p.add_sample(-1, 1.0)
    p.log_run("gaussian10", 64, 10_000)
    p.close()
    q = PersistentEstimator(MeanRateEstimator(), db, "update")
    assert q.estimate_ops(1.0) == 85
    other = PersistentEstimator(MeanRateEstimator(), db, "insertion")
    with pytest.raises(ValueError):
        other.estimate_ops(1.0)
    import sqlite3

    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute("PRAGMA table_info(samples)")]
    assert cols == ["id", "kind", "ops", "elapsed", "ts"]
    cols = [r[1] for r in con.execute("PRAGMA table_info(runs)")]
    assert cols == ["id", "kind", "data_name", "n_features",
                    "n_inserted_instances", "ts"]
    assert con.execute("SELECT COUNT(*) FROM runs").fetchone()[0] == 1
