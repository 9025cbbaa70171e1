"""MI355X-native SPECT/PET system-matrix generator (Python face).

The compute path is libsysmat_hip.so (hand-written gfx950 HIP kernels behind
the C ABI in include/sysmat.h); this package binds it with ctypes and mirrors
the reference's interfaces: Params_*.dat geometry (``params``), the
``PESysMatGen`` operator (``pe``) and the scatter operators (``scatter``);
``csr`` is the sparse output stage.
"""
from . import csr, params
from ._lib import SysmatError, load
from .params import Geometry
from .pe import PePlan, PESysMatGen, pe_compute

__all__ = ["csr", "params", "Geometry", "SysmatError", "load", "PePlan", "PESysMatGen", "pe_compute"]
