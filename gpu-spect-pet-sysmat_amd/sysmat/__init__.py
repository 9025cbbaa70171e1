"""MI355X-native SPECT/PET system-matrix generator (Python face).

The compute path is libsysmat_hip.so (hand-written gfx950 HIP kernels behind
the C ABI in include/sysmat.h); this package binds it with ctypes and mirrors
the reference's interfaces: Params_*.dat geometry (``params``), the
``PESysMatGen`` operator (``pe``) and the scatter operators (``scatter``);
``csr`` is the sparse output stage, ``matrix`` reads the dense ``.sysmat``
files, ``geocheck`` reports on a parameter set.
"""
from . import csr, geocheck, matrix, params
from ._lib import SysmatError, load
from .params import Geometry
from .pe import PePlan, PESysMatGen, pe_compute

__all__ = ["csr", "geocheck", "matrix", "params", "Geometry", "SysmatError", "load", "PePlan", "PESysMatGen", "pe_compute"]
