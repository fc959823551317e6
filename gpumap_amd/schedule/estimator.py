"""Per-operation cost models of the progressive scheduler.

Each model learns from the (ops, seconds) pairs that run() reports per
operation (GPUMAPResult's n_*_ops and *_time) and answers two questions for
the next cycle: how many ops fit a time budget at a confidence level, and how
long a given op count takes.  Behaviour follows the reference's estimators
(python/gpumap/schedule/estimator.py): MeanRateEstimator (:40-101),
DefaultLinearEstimator (:104-227) and PersistentEstimator (:230-325, whose
SQLite tables are kept column for column so that stored samples stay
readable by either implementation).
"""

from __future__ import annotations

import itertools
import os
import sqlite3
import statistics
import time
from collections import deque
from dataclasses import dataclass
from typing import Iterator, Optional


def _quantile(confidence: float) -> float:
    """Standard normal quantile: the margin, in standard deviations, that a
    one-sided `confidence` level asks for."""
    return statistics.NormalDist().inv_cdf(confidence)


class Estimator:
    """Interface of a cost model."""

    def add_sample(self, ops: int, elapsed_sec: float):
        raise NotImplementedError()

    def estimate_ops(self, time_budget_sec: float,
                     confidence: float = 0.95) -> int:
        raise NotImplementedError()

    def predict_time(self, ops: int, quantile: float = 0.95) -> float:
        raise NotImplementedError()


@dataclass
class FitResult:
    """t = slope_a * ops + intercept_b, residual deviation sigma."""
    slope_a: float
    intercept_b: float
    sigma: float
    n_samples: int


@dataclass
class RateFitResult:
    """Seconds per op: mean and sample deviation over the window."""
    mean_rate: float
    sigma_rate: float
    n_samples: int


class _Window:
    """The most recent `size` valid samples (ops > 0, seconds >= 0); a
    version number tells the models when to refit."""

    def __init__(self, size: int):
        self._q: deque = deque(maxlen=size)
        self.version = 0

    def push(self, ops: int, seconds: float) -> bool:
        if ops <= 0 or seconds < 0:
            return False
        self._q.append((int(ops), float(seconds)))
        self.version += 1
        return True

    def __len__(self) -> int:
        return len(self._q)

    def __iter__(self):
        return iter(self._q)


class _WindowedModel(Estimator):
    """A model fitted lazily to a sample window."""

    def __init__(self, window_size: int, min_samples: int):
        self.window_size = window_size
        self.min_samples = min_samples
        self._window = _Window(window_size)
        self._fitted_at = -1
        self._fitted = None

    @property
    def n_samples(self) -> int:
        """Samples currently in the window."""
        return len(self._window)

    def samples(self) -> list[tuple[int, float]]:
        return list(self._window)

    def add_sample(self, ops: int, elapsed_sec: float):
        self._window.push(ops, elapsed_sec)

    def _compute_fit(self):
        raise NotImplementedError()

    def fit_result(self):
        if self._fitted_at != self._window.version:
            self._fitted = self._compute_fit()
            self._fitted_at = self._window.version
        return self._fitted

    fit = fit_result


class MeanRateEstimator(_WindowedModel):
    """Proportional model t = rate * ops.  The op count for a budget T is
    floor(T / (mean + z sd)) over the window's per-op rates, z the normal
    quantile of the confidence; it needs `min_samples` samples first."""

    def __init__(self, window_size: int = 20, min_samples: int = 3):
        super().__init__(window_size, min_samples)

    def _compute_fit(self) -> RateFitResult:
        rates = [t / o for o, t in self._window]
        if len(rates) < self.min_samples:
            raise ValueError("Not enough samples.")
        mean = statistics.fmean(rates)
        if mean <= 0:
            raise ValueError("Non-positive mean rate.")
        sd = statistics.stdev(rates, mean) if len(rates) > 1 else 0.0
        return RateFitResult(mean, sd, len(rates))

    def _rate_at(self, q: float) -> float:
        fr = self.fit_result()
        return fr.mean_rate + _quantile(q) * fr.sigma_rate

    def estimate_ops(self, time_budget_sec: float,
                     confidence: float = 0.95) -> int:
        if time_budget_sec <= 0:
            return 0
        rate = self._rate_at(confidence)
        return int(time_budget_sec / rate) if rate > 0 else 0

    def predict_time(self, ops: int, quantile: float = 0.95) -> float:
        return ops * max(0.0, self._rate_at(quantile)) if ops > 0 else 0.0


class DefaultLinearEstimator(_WindowedModel):
    """Affine model t = a * ops + b by least squares over the window.  The op
    count for a budget T is (T - b - z sigma) / a.  Before `min_samples`
    samples it proposes 1000, 2000, 3000 ops (then 3000); when the fit has a
    non-positive slope or leaves no budget it falls back to the slowest
    per-op rate among the window's samples of at least 1 % of its largest
    op count."""

    COLD_START = (1000, 2000, 3000)

    def __init__(self, window_size: int = 10, min_samples: int = 3):
        super().__init__(window_size, min_samples)
        self._cold: Iterator[int] = itertools.chain(
            self.COLD_START, itertools.repeat(self.COLD_START[-1]))

    def _compute_fit(self) -> Optional[FitResult]:
        pts = list(self._window)
        if len(pts) < self.min_samples:
            return None
        xs = [float(o) for o, _ in pts]
        ys = [t for _, t in pts]
        if len(set(xs)) == 1:
            raise ValueError("Degenerate samples.")
        a, b = statistics.linear_regression(xs, ys)
        resid = [y - (a * x + b) for x, y in zip(xs, ys)]
        sigma = statistics.stdev(resid) if len(resid) > 1 else 0.0
        return FitResult(a, b, sigma, len(pts))

    def _slowest_rate_ops(self, budget: float) -> int:
        if budget <= 0:
            return 0
        pts = list(self._window)
        if pts:
            floor = 0.01 * max(o for o, _ in pts)
            rates = [t / o for o, t in pts if o >= floor and t > 0]
            if rates and max(rates) > 0:
                return max(1, int(budget / max(rates)))
        return next(self._cold)

    def estimate_ops(self, time_budget_sec: float,
                     confidence: float = 0.95) -> int:
        fr = self.fit_result()
        if fr is None:
            return next(self._cold)
        room = time_budget_sec - fr.intercept_b - _quantile(confidence) * fr.sigma
        if fr.slope_a > 0 and room > 0 and room / fr.slope_a > 0:
            return int(room / fr.slope_a)
        return self._slowest_rate_ops(time_budget_sec)

    def predict_time(self, ops: int, quantile: float = 0.95) -> float:
        fr = self.fit_result()
        if fr is None:
            return float("inf")
        return fr.slope_a * ops + fr.intercept_b + _quantile(quantile) * fr.sigma


# The on-disk store: one row per sample and one per logged run, each tagged
# with the operation kind ("insertion", "update", "embedding").
_SCHEMA = (
    "CREATE TABLE IF NOT EXISTS samples ("
    "id INTEGER PRIMARY KEY AUTOINCREMENT, kind TEXT NOT NULL, "
    "ops INTEGER NOT NULL, elapsed REAL NOT NULL, ts REAL NOT NULL)",
    "CREATE INDEX IF NOT EXISTS idx_kind_ts ON samples(kind, ts)",
    "CREATE TABLE IF NOT EXISTS runs ("
    "id INTEGER PRIMARY KEY AUTOINCREMENT, kind TEXT NOT NULL, "
    "data_name TEXT NOT NULL, n_features INTEGER NOT NULL, "
    "n_inserted_instances INTEGER NOT NULL, ts REAL NOT NULL)",
)


class PersistentEstimator(Estimator):
    """A cost model whose samples outlive the process: every sample is also
    written to a SQLite file, and on opening, the `load_limit` most recent
    samples of this `kind` are replayed into the wrapped model (oldest
    first)."""

    def __init__(self, base: Estimator, db_path: str, kind: str,
                 load_limit: int = 100):
        self._base = base
        self._kind = kind
        folder = os.path.dirname(os.path.abspath(db_path))
        os.makedirs(folder, exist_ok=True)
        self._db = sqlite3.connect(db_path, check_same_thread=False)
        with self._db:
            for stmt in _SCHEMA:
                self._db.execute(stmt)
        recent = self._db.execute(
            "SELECT ops, elapsed FROM samples WHERE kind = ? "
            "ORDER BY ts DESC LIMIT ?", (kind, load_limit)).fetchall()
        for ops, seconds in recent[::-1]:
            base.add_sample(int(ops), float(seconds))

    @property
    def base(self) -> Estimator:
        return self._base

    def add_sample(self, ops: int, elapsed_sec: float) -> None:
        if ops <= 0 or elapsed_sec < 0:
            return
        self._base.add_sample(ops, elapsed_sec)
        with self._db:
            self._db.execute(
                "INSERT INTO samples (kind, ops, elapsed, ts) VALUES (?, ?, ?, ?)",
                (self._kind, int(ops), float(elapsed_sec), time.time()))

    def estimate_ops(self, time_budget_sec: float,
                     confidence: float = 0.95) -> int:
        return self._base.estimate_ops(time_budget_sec, confidence)

    def predict_time(self, ops: int, quantile: float = 0.95) -> float:
        return self._base.predict_time(ops, quantile)

    def log_run(self, data_name: str, n_features: int,
                n_inserted_instances: int) -> None:
        with self._db:
            self._db.execute(
                "INSERT INTO runs (kind, data_name, n_features, "
                "n_inserted_instances, ts) VALUES (?, ?, ?, ?, ?)",
                (self._kind, data_name, int(n_features),
                 int(n_inserted_instances), time.time()))

    def close(self) -> None:
        try:
            self._db.close()
        except sqlite3.Error:
            pass
