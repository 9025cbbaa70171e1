"""Photoelectric system matrix: Python face of the HIP tracer.

Mirrors the reference's operator interface:
``PESysMatGen(parameter_Collimator, parameter_Detector, parameter_Image, dst, cuda_id)``
(PEGen_RayTracing_CircularHole/PESysMatGen.h:4-9, PESysMatGen.cu:766-833) fills
``dst`` with one Z-slab described in-band by Img[2]/[30]/[31]/[20] and returns
the slab's voxel count; errors raise ``SysmatError`` (the reference has no
error checks and returns ``numImagebin`` unconditionally).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np

from ._lib import SmPeOptions, check, load, ptr
from .params import Geometry, check_tables


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _padded_img(img) -> np.ndarray:
    out = np.zeros(max(100, len(img)), np.float32)
    out[:len(img)] = img
    return out


def PESysMatGen(parameter_Collimator, parameter_Detector, parameter_Image, dst: np.ndarray, cuda_id: int = 0) -> int:
    """Drop-in for the reference PESysMatGen; ``dst`` (float32, >= N*X*Y*Img[2]) is written in place."""
    col, det, img = _f32(parameter_Collimator), _f32(parameter_Detector), _padded_img(parameter_Image)
    check_tables(col, det)
    if dst.dtype != np.float32 or not dst.flags.c_contiguous:
        raise ValueError("dst must be a C-contiguous float32 array")
    N = int(np.floor(det[0] + 1e-6))
    need = N * int(np.floor(img[0])) * int(np.floor(img[1])) * int(np.floor(img[2]))
    if dst.size < need:
        raise ValueError(f"dst holds {dst.size} floats, the slab needs {need}")
    return check(load().sm_PESysMatGen(ptr(col), ptr(det), ptr(img), ptr(dst), int(cuda_id)))


def options(voxel_sub: int = 1, sub: Sequence[int] = (8, 16, 8), det_range: Optional[Tuple[int, int]] = None,
            z_range: Optional[Tuple[int, int]] = None, rotation: int = 0, tile: Optional[Sequence[int]] = None,
            bundle_filter: bool = True, device: int = 0) -> SmPeOptions:
    o = SmPeOptions()
    check(load().sm_pe_default_options(ctypes.byref(o)))
    o.voxel_sub = int(voxel_sub)
    o.sub_x, o.sub_y, o.sub_z = (int(v) for v in sub)
    if det_range is not None:
        o.det_begin, o.det_count = int(det_range[0]), int(det_range[1] - det_range[0])
    if z_range is not None:
        o.z_begin, o.z_count = int(z_range[0]), int(z_range[1] - z_range[0])
    o.rotation = int(rotation)
    if tile is not None:
        o.tile_x, o.tile_y, o.tile_z = (int(v) for v in tile)
    o.bundle_filter = 1 if bundle_filter else 0
    o.device = int(device)
    return o


def pe_compute(geo: Geometry, **kw) -> np.ndarray:
    """PE rows x voxels (det-major float32, host memory) computed on the GPU."""
    o = options(**kw)
    X, Y, Z = geo.shape
    N = geo.n_det
    rows = o.det_count if o.det_count > 0 else N - o.det_begin
    zc = o.z_count if o.z_count > 0 else Z - o.z_begin
    dst = np.zeros((rows, X * Y * zc), np.float32)
    col, det, img = _f32(geo.col), _f32(geo.det), _padded_img(geo.img)
    check(load().sm_pe_compute(ptr(col), ptr(det), ptr(img), ctypes.byref(o), ptr(dst)))
    return dst


def autotune(geo: Geometry, sample_rows: int = 64, cache_path: Optional[str] = None, **kw):
    """Fastest voxel tile for this geometry and sampling (sm_pe_autotune); returns
    (tile, ns per element, from_cache)."""
    o = options(**kw)
    best = SmPeOptions()
    ns = ctypes.c_double()
    col, det, img = _f32(geo.col), _f32(geo.det), _padded_img(geo.img)
    rc = check(load().sm_pe_autotune(ptr(col), ptr(det), ptr(img), ctypes.byref(o), int(sample_rows),
                                     None if cache_path is None else cache_path.encode(), ctypes.byref(best),
                                     ctypes.byref(ns)))
    return (best.tile_x, best.tile_y, best.tile_z), ns.value, rc == 1


class PePlan:
    """Geometry resident in HBM; launches detector blocks into device tensors.

    ``launch`` is asynchronous on the given HIP stream (torch.cuda.Stream or
    raw handle); ``sync`` waits and reports cull-list overflow.
    """

    def __init__(self, geo: Geometry, **kw):
        self._lib = load()
        self.geo = geo
        self.opts = options(**kw)
        self._col, self._det, self._img = _f32(geo.col), _f32(geo.det), _padded_img(geo.img)
        h = ctypes.c_void_p()
        check(self._lib.sm_pe_plan_create(ptr(self._col), ptr(self._det), ptr(self._img), ctypes.byref(self.opts),
                                          ctypes.byref(h)))
        self._h = h

    @property
    def subrays(self) -> int:
        return int(self._lib.sm_pe_plan_subrays(self._h))

    def launch(self, det_begin: int, det_count: int, dst, row_stride: int = 0, stream=None) -> None:
        s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        check(self._lib.sm_pe_plan_launch(self._h, int(det_begin), int(det_count), ptr(dst), int(row_stride),
                                          ctypes.c_void_p(s) if s else None))

    def sync(self) -> None:
        check(self._lib.sm_pe_plan_sync(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sm_pe_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
