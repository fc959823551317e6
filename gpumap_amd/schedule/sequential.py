"""One operation at a time with an unbounded op count: insertion until it
completes, then the update, then the embedding (behaviour of the reference's
SequentialPriorityStrategy, python/gpumap/schedule/sequential.py:9-52).  The
engine caps each count at what is left, so every cycle runs the whole
remaining operation."""

from __future__ import annotations

from typing import Optional

from .estimator import Estimator
from .strategy import UINT32_MAX, BaseStrategy


class SequentialPriorityStrategy(BaseStrategy):
    def __init__(self, n_instances: int, n_neighbors: int = 15,
                 insertion_estimator: Optional[Estimator] = None,
                 update_estimator: Optional[Estimator] = None,
                 embedding_estimator: Optional[Estimator] = None,
                 default_n_ops: int = 100, confidence: float = 0.95,
                 min_insertion_ops: int = 500):
        super().__init__((1.0, 1.0, 1.0), 0.1, insertion_estimator,
                         update_estimator, embedding_estimator, n_instances,
                         n_neighbors, default_n_ops, confidence,
                         min_insertion_ops)

    def allocate(self, target_latency: float) -> tuple[int, int, int]:
        # the first unfinished operation, the embedding once all are done
        first = next((i for i, op in enumerate(self._ops) if not op.done), 2)
        return tuple(UINT32_MAX if i == first else 0 for i in range(3))
