"""Distance metric selector for ``GPUMAPParams.metric``.

Same values as the reference's gpumap.Metric (python/gpumap/metric.py:20-22,
cpp/include/gpumap/gpumap.cuh:13).  ``params.metric`` also accepts the names
"euclidean" / "cosine" (case-insensitive) or a plain int.
"""

from __future__ import annotations

import enum


class Metric(enum.IntEnum):
    Euclidean = 0
    Cosine = 1
