"""MI355X-native coupled WCSPH-DEM time step (hand-written HIP for gfx950).

Python host layer with the reference's interface (Solv, input files,
SPH_OneStep / SPHDEM_OneStep, HashGrid, per-kernel entries) over the C ABI
of libsphdem.so (include/sphdem.h).
"""
from .config import BaseSolv, Solv
from .struct import BNDptl, DEMBNDptl, DEMptl, SPHptl

__all__ = ["BaseSolv", "Solv", "SPHptl", "BNDptl", "DEMptl", "DEMBNDptl"]
