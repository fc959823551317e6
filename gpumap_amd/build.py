"""Build the gfx950 engine in-tree.

Produces, next to this file:
  libgpumap_amd.so            HIP kernels + C++ engine + C ABI (hipcc, gfx950)
  _core<EXT_SUFFIX>.so        pybind11 binding over the C ABI (g++)

Run ``python -m gpumap_amd.build`` or call :func:`build`.  Object files go to
``gpumap_amd/_build`` (git-ignored).
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpumap_amd.so")
HIP_SOURCES = ["nnd_kernels.hip", "exact_kernels.hip", "mfma_kernels.hip",
               "fp8_kernels.hip", "join_kernels.hip", "embed_kernels.hip",
               "engine.hip", "noalloc.hip"]
ARCH = os.environ.get("GPUMAP_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def core_path() -> str:
    return os.path.join(HERE, "_core" + sysconfig.get_config_var("EXT_SUFFIX"))


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([_hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-c",
                     os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
          "-o", tmp])
    os.replace(tmp, LIB)

    import pybind11

    py_inc = sysconfig.get_paths()["include"]
    core = core_path()
    tmp = core + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
          "-I", py_inc, "-I", pybind11.get_include(), "-I", INCLUDE,
          os.path.join(CSRC, "core_binding.cpp"), "-L", HERE, "-lgpumap_amd",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, core)
    if verbose:
        print(f"built {LIB}\nbuilt {core}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
