"""Loading of the in-tree native libraries.

``libgfaz_amd.so`` holds the HIP kernels for gfx950 and the C ABI declared in
``include/gfaz_amd.h``; ``libgfaz_synth.so`` is the host-only workload generator.
There is no fallback: if a library is missing the import fails with the build
command to run.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
CODEC_LIB = os.path.join(LIB_DIR, "libgfaz_amd.so")
SYNTH_LIB = os.path.join(LIB_DIR, "libgfaz_synth.so")

_cache: dict[str, ctypes.CDLL] = {}


def _load(path: str) -> ctypes.CDLL:
    lib = _cache.get(path)
    if lib is None:
        if not os.path.exists(path):
            raise RuntimeError(
                f"{os.path.basename(path)} is not built; run `python -m gfaz_amd._build` "
                "(hipcc --offload-arch=gfx950) first")
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        _cache[path] = lib
    return lib


def codec_lib() -> ctypes.CDLL:
    return _load(CODEC_LIB)


def synth_lib() -> ctypes.CDLL:
    return _load(SYNTH_LIB)
