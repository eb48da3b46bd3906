"""rfd -- MI355X-native bonded-particle-model (BPM) step of RockFit-DEM.

Python host side over the C ABI of librfd.so (include/rfd.h).  Device memory
is held in torch tensors; all compute runs in the HIP library.
"""
from ._lib import LIB_PATH, RfdError, declared_symbols, lib  # noqa: F401
from .params import Compression, CellGrid, Material, SampleBox  # noqa: F401
from . import specimen  # noqa: F401

__all__ = ["lib", "RfdError", "LIB_PATH", "Material", "CellGrid", "SampleBox", "Compression",
           "specimen"]
