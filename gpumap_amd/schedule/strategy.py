"""How one run() cycle's time budget is shared by insertion, update and
embedding.

Behaviour of the reference's BaseStrategy (python/gpumap/schedule/
strategy.py:14-160): the budget T gives operation i the share
T * w_i / (sum(w) * (1 + reserved)); its cost model turns the share into an
op count; a cold or failing model, or one that answers 0 for a positive
share, is replaced by `default_n_ops` varied by up to 10 % with numpy's
global generator; a completed operation's weight becomes 0.  Progress per
operation is inserted / total, 1 - update queue / inserted and
1 - embedding queue / pending rows.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .estimator import Estimator, MeanRateEstimator

OPERATIONS = ("insertion", "update", "embedding")
UINT32_MAX = 2**32 - 1


def clip_ops(n_ops: int, ops_range: tuple[int, int]) -> int:
    """n_ops limited to [low, high]."""
    low, high = ops_range
    return min(max(n_ops, low), high)


def _jitter(base: int) -> int:
    """`base` moved by a draw in [-s, s) with s = max(1, base // 10)."""
    spread = max(1, base // 10)
    return base + int(np.random.randint(-spread, spread))


def _estimate_or_fallback(estimator: Estimator, budget_sec: float,
                          confidence: float, default_ops: int) -> int:
    """The model's op count for a budget; a positive budget always gets a
    positive count (the jittered default when the model has none)."""
    if budget_sec <= 0.0:
        return 0
    try:
        ops = int(estimator.estimate_ops(budget_sec, confidence))
    except ValueError:  # a model without enough samples
        ops = 0
    return ops if ops > 0 else _jitter(default_ops)


@dataclass
class _Operation:
    """One of the three operations as a strategy sees it."""
    name: str
    weight: float
    model: Estimator
    done: bool = False
    progress: float = 0.0


class BaseStrategy(ABC):
    """Budget split and progress tracking; `allocate` is the policy."""

    def __init__(self, weights: tuple[float, float, float] = (1.0, 1.0, 1.0),
                 reserved: float = 0.2,
                 insertion_estimator: Optional[Estimator] = None,
                 update_estimator: Optional[Estimator] = None,
                 embedding_estimator: Optional[Estimator] = None,
                 n_instances: int = 0, n_neighbors: int = 15,
                 default_n_ops: int = 100, confidence: float = 0.95,
                 min_insertion_ops: Optional[int] = None):
        models = (insertion_estimator, update_estimator, embedding_estimator)
        self._ops = tuple(
            _Operation(name, float(w), m if m is not None else MeanRateEstimator())
            for name, w, m in zip(OPERATIONS, weights, models))
        self.reserved = float(reserved)
        self.n_instances = int(n_instances)
        self.n_neighbors = int(n_neighbors)
        self.default_n_ops = int(default_n_ops)
        self.confidence = float(confidence)
        # by default an insertion takes at least 1 % of the rows per cycle
        self.min_insertion_ops = (min_insertion_ops if min_insertion_ops
                                  else int(0.01 * n_instances))
        self.n_inserted = 0

    # -- the three operations' models, weights and state ---------------
    @property
    def insertion_estimator(self) -> Estimator:
        return self._ops[0].model

    @property
    def update_estimator(self) -> Estimator:
        return self._ops[1].model

    @property
    def embedding_estimator(self) -> Estimator:
        return self._ops[2].model

    @property
    def weights(self) -> tuple[float, float, float]:
        return tuple(op.weight for op in self._ops)

    @property
    def completeness(self) -> tuple[bool, bool, bool]:
        return tuple(op.done for op in self._ops)

    @property
    def insertion_progress(self) -> float:
        return self._ops[0].progress

    @property
    def update_progress(self) -> float:
        return self._ops[1].progress

    @property
    def embedding_progress(self) -> float:
        return self._ops[2].progress

    def get_progress(self) -> tuple[float, float, float]:
        return tuple(op.progress for op in self._ops)

    # -- policy ---------------------------------------------------------
    @abstractmethod
    def allocate(self, target_latency: float) -> tuple[int, int, int]:
        """(insertion, update, embedding) ops for one cycle within
        `target_latency` seconds."""

    def _split(self, total: float) -> tuple[float, float, float]:
        scale = sum(op.weight for op in self._ops) * (1.0 + self.reserved)
        if scale <= 0:
            return (0.0, 0.0, 0.0)
        return tuple(total * op.weight / scale for op in self._ops)

    # -- feedback from each cycle ------------------------------------------
    def update(self, result) -> None:
        """Learn the cycle's per-operation costs and track progress.  The
        insertion's model stops learning once the insertion has completed
        (its last batches are partial)."""
        observed = ((result.n_insertion_ops, result.insertion_time),
                    (result.n_update_ops, result.update_time),
                    (result.n_embedding_ops, result.embedding_time))
        for op, (n, secs) in zip(self._ops, observed):
            if n > 0 and not (op.name == "insertion" and op.done):
                op.model.add_sample(n, secs)
        self.n_inserted = result.n_inserted_instances
        ins, upd, emb = self._ops
        ins.progress = (self.n_inserted / self.n_instances
                        if self.n_instances > 0 else 0.0)
        upd.progress = (1.0 - result.update_queue_size / self.n_inserted
                        if self.n_inserted > 0 else 0.0)
        pending = result.embedding_pending_edges
        emb.progress = (1.0 - result.embedding_queue_size / pending
                        if pending > 0 else 0.0)

    def on_operation_complete(self, operation: str) -> None:
        if operation not in OPERATIONS:
            raise ValueError(f"Unknown operation: {operation}")
        op = self._ops[OPERATIONS.index(operation)]
        op.done = True
        op.weight = 0.0
