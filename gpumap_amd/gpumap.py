"""Python facade of the progressive UMAP engine.

Same surface as the reference facade (python/gpumap/gpumap.py:20-162):
``GPUMAP(data_loader, params, strategy=None, insertion_estimator=None,
update_estimator=None, embedding_estimator=None, default_n_ops=100)`` with
``run(target_latency)``, ``run_with_n_ops(i, u, e)``, ``get_embedding()``,
``get_neighbor_indices()``, ``get_neighbor_dists()``, ``close()``.  The getters
return device tensors imported through DLPack (torch tensors on ROCm) instead
of CuPy arrays; they are views of the engine's HBM arrays, valid while the
returned tensor lives.
"""

from __future__ import annotations

import numpy as np

from ._native import _core
from .data_loader import DataLoader
from .schedule import BaseStrategy, Estimator, WeightedStrategy
from .validators import as_f32_c_contig_2d, validate_data_chunk


def find_ab_params(spread: float, min_dist: float) -> tuple[float, float]:
    """Fit 1 / (1 + a x^(2b)) to the UMAP target membership curve on 300
    points of [0, 3 spread] (reference gpumap.py:20-29)."""
    from scipy.optimize import curve_fit

    def curve(x, a, b):
        return 1.0 / (1.0 + a * x ** (2 * b))

    xv = np.linspace(0.0, 3.0 * spread, 300)
    yv = np.where(xv < min_dist, 1.0, np.exp(-(xv - min_dist) / spread))
    (a, b), _ = curve_fit(curve, xv, yv)
    return float(a), float(b)


def _device_rows(chunk):
    """A torch tensor on a GPU as (pointer, rows, cols) of C-contiguous
    float32 rows, or None for anything else (host arrays take the reference's
    path)."""
    if not getattr(chunk, "is_cuda", False):
        return None
    import torch

    if chunk.dim() != 2:
        raise ValueError("data_chunk must be 2D array")
    chunk = chunk.to(torch.float32).contiguous()
    # the rows' producer is torch's current stream.  The session's stream is
    # a blocking stream, so it already waits for work on the null stream; a
    # side stream must finish first.
    producer = torch.cuda.current_stream(chunk.device)
    if producer.cuda_stream != 0:
        producer.synchronize()
    return chunk, chunk.data_ptr(), int(chunk.shape[0]), int(chunk.shape[1])


def _to_device_tensor(arr):
    try:
        import torch
    except ImportError:  # pragma: no cover - torch is part of the image
        return arr
    return torch.from_dlpack(arr)


class GPUMAP:
    def __init__(self, data_loader: DataLoader, params, strategy:
                 BaseStrategy | None = None,
                 insertion_estimator: Estimator | None = None,
                 update_estimator: Estimator | None = None,
                 embedding_estimator: Estimator | None = None,
                 default_n_ops: int = 100):
        self.data_loader = data_loader
        self._params = params
        if self.params.min_dist > 1.0:
            raise ValueError("min_dist must be <= 1.0")
        self.params.a, self.params.b = find_ab_params(1.0, self.params.min_dist)
        if not self.params.n_instances:
            self.params.n_instances = data_loader.n_instances
        if not self.params.n_features:
            self.params.n_features = data_loader.n_features
        try:
            self._impl = _core.GPUMAP(self.params)
        except MemoryError:
            # one retry, as the reference does for a transient first allocation
            self._impl = _core.GPUMAP(self.params)
        self._closed = False
        if strategy is not None:
            self._strategy = strategy
        else:
            self._strategy = WeightedStrategy(
                n_instances=data_loader.n_instances,
                n_neighbors=params.n_neighbors,
                insertion_estimator=insertion_estimator,
                update_estimator=update_estimator,
                embedding_estimator=embedding_estimator,
                default_n_ops=default_n_ops)
        self._inserted_instances = 0
        self._impl.initialize(self.data_loader.n_instances,
                              self.data_loader.n_features)

    @property
    def params(self):
        return self._params

    @property
    def strategy(self) -> BaseStrategy:
        return self._strategy

    @property
    def n_inserted_instances(self) -> int:
        return self._inserted_instances

    def close(self):
        if not self._closed:
            self._impl.close()
            self._closed = True

    def release(self):
        """Close the session and free its device memory as soon as no tensor
        returned by a getter holds it (no counterpart in the reference, whose
        storage lives until the object is destroyed); the session accepts no
        further calls."""
        self._impl.release()
        self._closed = True

    @property
    def released(self) -> bool:
        """True once a released session's device memory has been freed."""
        return self._impl.released

    def _feed(self, n_insertion_ops: int) -> None:
        if n_insertion_ops <= 0:
            return
        chunk = self.data_loader.get_next_chunk(n_insertion_ops)
        dev = _device_rows(chunk)
        if dev is not None:
            # rows resident in HBM (a loader over a torch GPU tensor): appended
            # by a device copy, no host round trip
            _, ptr, rows, cols = dev
            if rows:
                self._impl.append_data_device(ptr, rows, cols)
            return
        try:
            validate_data_chunk(chunk, self._params.n_features)
            chunk = as_f32_c_contig_2d(chunk)
        except ValueError as e:
            # the reference reports and still appends (gpumap.py:93-99); the
            # binding then rejects a malformed chunk with its own error
            print(f"Error validating data chunk: {e}")
        self._impl.append_data(chunk)

    def _finish(self, res):
        self._inserted_instances = res.n_inserted_instances
        self._strategy.update(res)
        if res.insertion_completed:
            self._strategy.on_operation_complete("insertion")
        if res.update_completed:
            self._strategy.on_operation_complete("update")
        if res.embedding_completed:
            self._strategy.on_operation_complete("embedding")
        return res

    def run_with_n_ops(self, n_insertion_ops: int, n_update_ops: int,
                       n_embedding_ops: int):
        self._feed(n_insertion_ops)
        res = self._impl.run(n_insertion_ops, n_update_ops, n_embedding_ops)
        return self._finish(res)

    def run(self, target_latency: float = 10.0):
        allocation = self._strategy.allocate(target_latency)
        n_ins = min(allocation[0],
                    self.data_loader.n_instances - self._inserted_instances)
        self._feed(n_ins)
        res = self._impl.run(n_ins, allocation[1], allocation[2])
        return self._finish(res)

    def get_embedding(self):
        return _to_device_tensor(self._impl.get_embedding())

    def get_neighbor_indices(self):
        return _to_device_tensor(self._impl.get_neighbor_indices())

    def get_neighbor_dists(self):
        return _to_device_tensor(self._impl.get_neighbor_dists())
