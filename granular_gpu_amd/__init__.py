"""MI355X-native granular PBD solver (HIP kernels for gfx950 behind a C-ABI).

The compute path is libgranular_gpu.so; this package is the Python-side
binding and scene tooling.  Importing it does not touch the GPU.
"""
from . import scenes  # noqa: F401
from ._lib import LIB_PATH, declared_symbols, load  # noqa: F401

__all__ = ["scenes", "load", "declared_symbols", "LIB_PATH"]
