"""Chunk sources for the facade (reference python/gpumap/data_loader.py:7-62).

``DataLoader.get_next_chunk(n)`` hands out the next ``n`` rows (fewer at the
end).  ``NumpyDataLoader`` reads an ``.npy`` file (no pickles) and
``ArrayDataLoader`` wraps an in-memory array.
"""

from __future__ import annotations

import os

import numpy as np


class DataLoader:
    n_instances: int
    n_features: int
    X: np.ndarray
    y: np.ndarray | None = None
    _current_index: int = 0

    def __init__(self):
        raise NotImplementedError("init not implemented")

    def get_next_chunk(self, chunk_size: int) -> np.ndarray:
        chunk_size = min(int(chunk_size), self.n_instances - self._current_index)
        chunk = self.X[self._current_index:self._current_index + chunk_size]
        self._current_index += chunk_size
        return chunk


def load_numpy_data(file_path: str) -> np.ndarray:
    if not os.path.exists(file_path):
        raise ValueError(f"File not found: {file_path}")
    return np.load(file_path, allow_pickle=False)


class NumpyDataLoader(DataLoader):
    def __init__(self, dataset_name: str, npy_path: str,
                 label_npy_path: str | None = None):
        self.dataset_name = dataset_name
        self.X = load_numpy_data(npy_path)
        self.y = load_numpy_data(label_npy_path) if label_npy_path else None
        self.n_instances = len(self.X)
        self.n_features = self.X.shape[1]
        self._current_index = 0


class ArrayDataLoader(DataLoader):
    """Chunks of an in-memory array: a host array, or a torch tensor on the
    GPU whose chunks the facade appends by a device copy."""

    def __init__(self, X: np.ndarray, y: np.ndarray | None = None):
        self.X = X if getattr(X, "is_cuda", False) else np.asarray(X)
        self.y = y
        self.n_instances = int(self.X.shape[0])
        self.n_features = int(self.X.shape[1])
        self._current_index = 0
