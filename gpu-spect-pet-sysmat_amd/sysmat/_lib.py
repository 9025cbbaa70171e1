"""ctypes binding of libsysmat_hip.so (the C ABI declared in include/sysmat.h).

The library is loaded from the package directory (built in-tree by
``__graft_entry__.build()`` / ``make``).  Loading fails loudly: there is no CPU
fallback on the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("SYSMAT_LIB") or os.path.join(PKG_DIR, "libsysmat_hip.so")

SM_OK = 0
SM_E_ARG = -1
SM_E_DEVICE = -2
SM_E_NOMEM = -3
SM_E_OVERFLOW = -4
SM_E_IO = -5

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class SmPeOptions(ctypes.Structure):
    _fields_ = [("voxel_sub", ctypes.c_int), ("sub_x", ctypes.c_int), ("sub_y", ctypes.c_int),
                ("sub_z", ctypes.c_int), ("det_begin", ctypes.c_int), ("det_count", ctypes.c_int),
                ("z_begin", ctypes.c_int), ("z_count", ctypes.c_int), ("rotation", ctypes.c_int),
                ("tile_x", ctypes.c_int), ("tile_y", ctypes.c_int), ("tile_z", ctypes.c_int),
                ("bundle_filter", ctypes.c_int), ("device", ctypes.c_int)]


class SysmatError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sysmat error {code}: {msg}")
        self.code = code


# every symbol the C header declares, with (restype, argtypes)
_VP = ctypes.c_void_p
SIGNATURES = {
    "sm_version": (ctypes.c_int, []),
    "sm_last_error": (ctypes.c_char_p, []),
    "sm_device_count": (ctypes.c_int, []),
    "sm_pe_default_options": (ctypes.c_int, [ctypes.POINTER(SmPeOptions)]),
    "sm_PESysMatGen": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int]),
    "sm_pe_compute": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(SmPeOptions), _VP]),
    "sm_pe_compute_device": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(SmPeOptions), _VP, ctypes.c_longlong,
                                            _VP]),
    "sm_pe_noalloc_scratch_bytes": (ctypes.c_longlong, []),
    "sm_pe_compute_device_noalloc": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(SmPeOptions), _VP,
                                                    ctypes.c_longlong, _VP]),
    "sm_pe_autotune": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(SmPeOptions), ctypes.c_int, ctypes.c_char_p,
                                      ctypes.POINTER(SmPeOptions), ctypes.POINTER(ctypes.c_double)]),
    "sm_pe_plan_create": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(SmPeOptions), ctypes.POINTER(_VP)]),
    "sm_pe_plan_launch": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP, ctypes.c_longlong, _VP]),
    "sm_pe_plan_sync": (ctypes.c_int, [_VP]),
    "sm_pe_plan_destroy": (ctypes.c_int, [_VP]),
    "sm_pe_plan_subrays": (ctypes.c_longlong, [_VP]),
    "sm_crystal_bitmap_words": (ctypes.c_longlong, [ctypes.c_int]),
    "sm_collimator_bitmap_words": (ctypes.c_longlong, [ctypes.c_int, ctypes.c_int]),
    "sm_bitmap_crystal": (ctypes.c_int, [_VP, _VP, ctypes.c_longlong, ctypes.c_int]),
    "sm_bitmap_collimator": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_longlong, ctypes.c_int]),
    "sm_scatter": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, ctypes.c_char_p, ctypes.c_char_p, _VP, ctypes.c_int]),
    "sm_scatter_collimator_only": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, ctypes.c_char_p, _VP, ctypes.c_int]),
    "sm_scatter_crystal_only": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_char_p, _VP, ctypes.c_int]),
    "sm_scatter_crystal": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP,
                                          ctypes.c_int]),
    "sm_scatter_collimator": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, _VP, ctypes.c_int]),
    "sm_scatter_plan_create": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int, _VP, _VP, ctypes.c_int,
                                              ctypes.POINTER(_VP)]),
    "sm_scatter_plan_crystal": (ctypes.c_int, [_VP, _VP, ctypes.c_int, ctypes.c_int, _VP, ctypes.c_longlong, _VP]),
    "sm_scatter_plan_collimator": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP, ctypes.c_longlong, _VP]),
    "sm_scatter_plan_destroy": (ctypes.c_int, [_VP]),
    "sm_scatter_plan_create_ex": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _VP,
                                                 ctypes.c_int, ctypes.POINTER(_VP)]),
    "sm_scatter_plan_crystal_rows": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, _VP, ctypes.c_longlong, _VP]),
    "sm_scatter_plan_cells": (ctypes.c_int, [_VP, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int, _VP,
                                             ctypes.c_longlong, ctypes.c_int, _VP]),
    "sm_scatter_cell_table": (ctypes.c_longlong, [_VP, ctypes.c_float, ctypes.c_float, _VP, ctypes.c_longlong]),
    "sm_surrogate_tile_count": (ctypes.c_longlong, [_VP]),
    "sm_csr_beams_device": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, ctypes.c_double, _VP, _VP, _VP]),
    "sm_scatter_cells": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                        ctypes.c_int, _VP, ctypes.c_int]),
    "sm_surrogate_create": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.POINTER(_VP)]),
    "sm_surrogate_set_norm": (ctypes.c_int, [_VP, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                             ctypes.c_float]),
    "sm_surrogate_row_sites": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP, ctypes.c_int]),
    "sm_surrogate_near": (ctypes.c_int, [_VP, _VP, ctypes.c_int, ctypes.c_int, _VP, ctypes.c_longlong, _VP]),
    "sm_surrogate_exact_rows": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, _VP, ctypes.c_longlong, _VP]),
    "sm_surrogate_features": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, _VP,
                                             ctypes.c_int, _VP, _VP]),
    "sm_surrogate_apply": (ctypes.c_int, [_VP, _VP, ctypes.c_int, _VP, ctypes.c_int, ctypes.c_int, _VP,
                                          ctypes.c_longlong, ctypes.c_float, ctypes.c_float, _VP, _VP]),
    "sm_surrogate_fallback": (ctypes.c_longlong, [_VP, _VP, ctypes.c_int, ctypes.c_int, _VP, _VP, ctypes.c_longlong,
                                                  _VP]),
    "sm_surrogate_destroy": (ctypes.c_int, [_VP]),
    "sm_scatter_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_VP), _VP, _VP, _VP, ctypes.c_int,
                                               ctypes.c_int, ctypes.POINTER(_VP)]),
    "sm_scatter_batch_run": (ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), _VP, ctypes.c_int]),
    "sm_scatter_batch_graph_nodes": (ctypes.c_int, [_VP]),
    "sm_scatter_batch_destroy": (ctypes.c_int, [_VP]),
    "sm_csr_count_device": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_double,
                                           _VP, _VP, ctypes.POINTER(ctypes.c_longlong), _VP]),
    "sm_csr_fill_device": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_double,
                                          _VP, _VP, _VP, _VP, _VP]),
    "sm_pe_compute_csr": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(SmPeOptions), ctypes.c_double,
                                         ctypes.POINTER(_VP)]),
    "sm_csr_info": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                   ctypes.POINTER(ctypes.c_longlong)]),
    "sm_csr_copy": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "sm_csr_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_VP), ctypes.c_int]),
    "sm_csr_destroy": (ctypes.c_int, [_VP]),
    "sm_mlp_param_count": (ctypes.c_longlong, []),
    "sm_mlp_packed_bytes": (ctypes.c_longlong, []),
    "sm_mlp_workspace_bytes": (ctypes.c_longlong, [ctypes.c_int]),
    "sm_mlp_init": (ctypes.c_int, [_VP, ctypes.c_ulonglong, _VP]),
    "sm_mlp_forward": (ctypes.c_int, [_VP, _VP, ctypes.c_int, ctypes.c_int, _VP, ctypes.c_longlong, _VP, _VP]),
    "sm_mlp_backward": (ctypes.c_int, [_VP, _VP, ctypes.c_int, ctypes.c_int, _VP, ctypes.c_longlong, _VP, _VP]),
    "sm_mlp_train_step": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                         ctypes.c_int, _VP, ctypes.c_longlong, _VP, _VP]),
    "sm_mlp_train_run": (ctypes.c_int, [ctypes.c_int, _VP, _VP, _VP, _VP, _VP, _VP, ctypes.c_longlong, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_float, ctypes.c_float, _VP, ctypes.c_int, _VP,
                                        ctypes.c_longlong, _VP, _VP]),
    "sm_mlp_pack": (ctypes.c_int, [_VP, _VP, _VP]),
    "sm_mlp_infer": (ctypes.c_int, [_VP, ctypes.c_int, _VP, ctypes.c_longlong, _VP, _VP]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libsysmat_hip.so and bind every declared entry point (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SysmatError(SM_E_DEVICE, f"{path} is not built; run __graft_entry__.build() (make -C {PKG_DIR})")
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7
    # (same soname).  Importing torch first makes the dynamic loader bind this
    # library to torch's runtime, so device tensors, streams and events are
    # shared; loading ours first would leave torch with a second runtime that
    # sees no GPU.
    if os.environ.get("SYSMAT_NO_TORCH", "0") != "1":
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> int:
    if rc < 0:
        msg = load().sm_last_error().decode(errors="replace")
        raise SysmatError(rc, msg)
    return rc


def ptr(a) -> int:
    """Raw address of a numpy array or torch tensor."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())
