"""Chunk checks of the facade (reference python/gpumap/validators/
validate_data_chunk.py:4-16 and utils/as_f32_c_contig_2d.py:4-10)."""

from __future__ import annotations

import numpy as np


def validate_data_chunk(data_chunk: np.ndarray, feature_dim: int | None) -> int:
    shape = np.shape(data_chunk)
    if len(shape) != 2:
        raise ValueError("data_chunk must be 2D array")
    if feature_dim is not None and shape[1] != feature_dim:
        raise ValueError(
            f"data_chunk feature dimension {shape[1]} does not match "
            f"previous feature dimension {feature_dim}")
    return shape[1]


def as_f32_c_contig_2d(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x)
    if x.ndim != 2:
        raise ValueError("Input array must be 2D")
    if not x.flags["C_CONTIGUOUS"] or x.dtype != np.float32:
        x = np.ascontiguousarray(x, dtype=np.float32)
    return x
