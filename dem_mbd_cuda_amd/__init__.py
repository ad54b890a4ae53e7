"""dem_mbd_cuda_amd: the FP64 DEM step of EreYui/dem_mbd_cuda on MI355X.

libdemhip.so (HIP kernels for gfx950 behind the C-ABI in include/dem_hip.h)
is the product; this package loads it, mirrors the reference's
CudaDemSolver interface in Python, and builds scenes.  Loading fails loudly
when the native library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import os

from ._abi import DemError, Library, Solver  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdemhip.so")

_lib = None


def library() -> Library:
    """The loaded libdemhip.so (raises DemError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DemError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        _lib = Library(LIB_PATH, "dem_")
    return _lib


def CudaDemSolver() -> Solver:
    """A new solver on the HIP backend (mirror of the reference's class name)."""
    return library().solver()


def backend_available() -> bool:
    return bool(library()._fn["backend_available"]())


def device_description() -> str:
    return library()._fn["device_description"]().decode()
