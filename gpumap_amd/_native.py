"""Loader of the compiled engine.

The product path has no CPU fallback: importing gpumap_amd without the built
extension raises ImportError naming the build command.
"""

from __future__ import annotations

import atexit
import ctypes
import os

from .build import LIB, core_path

if not (os.path.exists(LIB) and os.path.exists(core_path())):
    raise ImportError(
        "gpumap_amd's gfx950 extension is not built "
        f"(missing {os.path.basename(LIB)} or _core); run "
        "`python -m gpumap_amd.build`")


def _share_torch_runtime() -> None:
    """Load torch (when installed) before the engine.  torch carries its
    own HIP runtime; loaded first, its runtime is the one the engine binds
    to, so the DLPack getters and the engine share one runtime.  With the
    engine's runtime loaded first, torch's own initialisation later fails
    in the same process (it finds no devices)."""
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - torch is part of the image
        pass


_share_torch_runtime()

# the binding resolves libgpumap_amd.so through its rpath; loading it here
# first makes the failure message precise if the HIP runtime is missing
_lib = ctypes.CDLL(LIB, mode=ctypes.RTLD_GLOBAL)

from . import _core  # noqa: E402


def _teardown() -> None:
    """Free the sessions' device memory while the HIP runtime is still up.
    Registered after torch is loaded, so it runs before torch's own exit
    hooks; tensors exported by the getters that outlive it must not be read
    afterwards.  Without it, sessions held by objects that interpreter
    teardown collects late would be freed after the runtime (or a profiler
    hooked into it) has finalized."""
    _core.shutdown()


atexit.register(_teardown)

__all__ = ["_core", "LIB"]
