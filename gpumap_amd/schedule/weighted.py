"""The default strategy: the budget shared 1 : 2 : 1 by insertion, update and
embedding with 10 % held back, each share turned into ops by its cost model,
and at least `min_insertion_ops` rows inserted per cycle (behaviour of the
reference's WeightedStrategy, python/gpumap/schedule/weighted.py:9-59)."""

from __future__ import annotations

from typing import Optional

from .estimator import Estimator
from .strategy import UINT32_MAX, BaseStrategy, _estimate_or_fallback, clip_ops


class WeightedStrategy(BaseStrategy):
    def __init__(self, n_instances: int, n_neighbors: int = 15,
                 weights: tuple[float, float, float] = (1.0, 2.0, 1.0),
                 reserved: float = 0.1,
                 insertion_estimator: Optional[Estimator] = None,
                 update_estimator: Optional[Estimator] = None,
                 embedding_estimator: Optional[Estimator] = None,
                 default_n_ops: int = 100, confidence: float = 0.95,
                 min_insertion_ops: int = 500):
        super().__init__(weights, reserved, insertion_estimator,
                         update_estimator, embedding_estimator, n_instances,
                         n_neighbors, default_n_ops, confidence,
                         min_insertion_ops)

    def allocate(self, target_latency: float) -> tuple[int, int, int]:
        shares = self._split(max(0.0, target_latency))
        floors = (self.min_insertion_ops, 0, 0)
        out = []
        for op, share, floor in zip(self._ops, shares, floors):
            n = _estimate_or_fallback(op.model, share, self.confidence,
                                      self.default_n_ops)
            out.append(clip_ops(n, (floor, UINT32_MAX)))
        return tuple(out)
