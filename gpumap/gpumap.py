"""gpumap.gpumap: the facade module (reference python/gpumap/gpumap.py)."""

import sys

from gpumap_amd import gpumap as _facade

sys.modules[__name__] = _facade
