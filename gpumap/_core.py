"""gpumap._core: the compiled engine's Python module (GPUMAP, GPUMAPParams,
the result type), as the reference's pybind11 module of that name
(python/gpumap/_core.pyi).  Importing it loads the gfx950 extension."""

import sys

from gpumap_amd._native import _core

sys.modules[__name__] = _core
