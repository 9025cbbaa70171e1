"""Primary-Compton scatter matrices: Python face of the HIP kernels.

Mirrors the reference's ``scatter(...)`` operators
(ScatterGen_RayTracing_CircularHole/scatter.h:4-5, ScatterGen_Collimator/scatter.h:4-5,
ScatterGen_Crystal/scatter.h:4): the combined crystal + collimator scatter of
one rotation (Img[20]) of an N x V PE matrix, with the GeometryRelationShip
bitmaps computed on the GPU or read from files (Phys[8] / Phys[9]).
Errors raise ``SysmatError`` (the reference exits with status 0 on a missing
file, scatter.cu:2145-2157).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np

from ._lib import check, load, ptr
from .params import Geometry, check_tables


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _img(img, rot: Optional[int] = None) -> np.ndarray:
    out = np.zeros(max(100, len(img)), np.float32)
    out[:len(img)] = img
    if rot is not None:
        out[20] = rot
    return out


def n_holes_layer0(geo: Geometry) -> int:
    return int(np.floor(geo.col[10] + 1e-6)) if geo.n_layers > 0 else 0


def bitmap_crystal(geo: Geometry, device: int = 0) -> np.ndarray:
    """GeometryRelationShip_Crystal2Crystal words, computed on the GPU."""
    L = load()
    w = np.zeros(int(L.sm_crystal_bitmap_words(geo.n_det)), np.uint32)
    check(L.sm_bitmap_crystal(ptr(_f32(geo.det)), ptr(w), w.size, device))
    return w


def bitmap_collimator(geo: Geometry, device: int = 0) -> np.ndarray:
    """GeometryRelationShip_Collimator2Crystal words, computed on the GPU."""
    L = load()
    w = np.zeros(max(1, int(L.sm_collimator_bitmap_words(n_holes_layer0(geo), geo.n_det))), np.uint32)
    check(L.sm_bitmap_collimator(ptr(_f32(geo.col)), ptr(_f32(geo.det)), ptr(w), w.size, device))
    return w


def scatter_crystal(geo: Geometry, pe_sysmat: np.ndarray, bits: Optional[np.ndarray] = None, rotation: int = 0,
                    det_range: Optional[Tuple[int, int]] = None, device: int = 0) -> np.ndarray:
    """crystalScatterSysMatCuda rows for the N x V PE matrix (host arrays in and out)."""
    b, e = det_range if det_range is not None else (0, geo.n_det)
    pe = _f32(pe_sysmat).reshape(geo.n_det, geo.n_vox)
    dst = np.zeros((e - b, geo.n_vox), np.float32)
    bw = None if bits is None else np.ascontiguousarray(bits, np.uint32)
    check(load().sm_scatter_crystal(ptr(_f32(geo.det)), ptr(_img(geo.img)), ptr(_f32(geo.phy)), ptr(pe),
                                    None if bw is None else ptr(bw), rotation, b, e - b, ptr(dst), device))
    return dst


def scatter_collimator(geo: Geometry, bits: Optional[np.ndarray] = None, rotation: int = 0, variant: int = 0,
                       det_range: Optional[Tuple[int, int]] = None, device: int = 0) -> np.ndarray:
    """collimatorScatterSysMatCuda rows (variant 0) or the ScatterGen_Collimator kernel (variant 1)."""
    b, e = det_range if det_range is not None else (0, geo.n_det)
    dst = np.zeros((e - b, geo.n_vox), np.float32)
    bw = None if bits is None else np.ascontiguousarray(bits, np.uint32)
    check(load().sm_scatter_collimator(ptr(_f32(geo.col)), ptr(_f32(geo.det)), ptr(_img(geo.img)), ptr(_f32(geo.phy)),
                                       None if bw is None else ptr(bw), rotation, variant, b, e - b, ptr(dst), device))
    return dst


def scatter_cells(geo: Geometry, cell_size: float = 0.5, keep_dist: float = 1.0, rotation: int = 0,
                  det_range: Optional[Tuple[int, int]] = None, device: int = 0) -> np.ndarray:
    """Plate-cell collimator scatter rows (host memory); see ScatterPlan.collimator_cells."""
    b, e = det_range if det_range is not None else (0, geo.n_det)
    dst = np.zeros((e - b, geo.n_vox), np.float32)
    check(load().sm_scatter_cells(ptr(_f32(geo.col)), ptr(_f32(geo.det)), ptr(_img(geo.img)), ptr(_f32(geo.phy)),
                                  rotation, cell_size, keep_dist, b, e - b, ptr(dst), device))
    return dst


def scatter(parameter_Collimator, parameter_Detector, parameter_Image, parameter_Physics, PE_SysMat,
            FnameGeoCrystal: str, FnameGeoCollimator: str, dst: np.ndarray, cuda_id: int = 0) -> int:
    """Drop-in for the combined reference scatter(); ``dst`` (N*V float32) is written in place."""
    col, det = _f32(parameter_Collimator), _f32(parameter_Detector)
    check_tables(col, det)
    img = _img(parameter_Image)
    N = int(np.floor(det[0] + 1e-4))
    V = int(np.floor(img[0] + 1e-3)) * int(np.floor(img[1] + 1e-3)) * int(np.floor(img[2] + 1e-3))
    pe = _f32(PE_SysMat)
    if pe.size < N * V or dst.size < N * V or dst.dtype != np.float32:
        raise ValueError("PE_SysMat and dst must hold N*V float32 values")
    return check(load().sm_scatter(ptr(col), ptr(det), ptr(img), ptr(_f32(parameter_Physics)),
                                   ptr(pe), FnameGeoCrystal.encode(), FnameGeoCollimator.encode(), ptr(dst), cuda_id))


class ScatterPlan:
    """Geometry, bitmaps and tables resident in HBM; row blocks into device tensors."""

    NO_CRYSTAL_BITMAP = 1
    NO_COLLIMATOR_BITMAP = 2

    def __init__(self, geo: Geometry, rotation: int = 0, variant: int = 0, with_collimator: bool = True,
                 crystal_bits: Optional[np.ndarray] = None, collimator_bits: Optional[np.ndarray] = None,
                 device: int = 0, bitmaps: bool = True):
        """``bitmaps=False`` evaluates the GeometryRelationShip criterion inline instead of
        building the bitmaps (same results; needed when N^3/2 bits do not fit)."""
        self._lib = load()
        self.geo = geo
        self._keep = [_f32(geo.col), _f32(geo.det), _img(geo.img), _f32(geo.phy)]
        cb = None if crystal_bits is None else np.ascontiguousarray(crystal_bits, np.uint32)
        hb = None if collimator_bits is None else np.ascontiguousarray(collimator_bits, np.uint32)
        flags = 0 if bitmaps else (self.NO_CRYSTAL_BITMAP | self.NO_COLLIMATOR_BITMAP)
        h = ctypes.c_void_p()
        check(self._lib.sm_scatter_plan_create_ex(ptr(self._keep[0]) if with_collimator else None,
                                                  ptr(self._keep[1]), ptr(self._keep[2]), ptr(self._keep[3]),
                                                  rotation, variant, flags, None if cb is None else ptr(cb),
                                                  None if hb is None else ptr(hb), device, ctypes.byref(h)))
        self._h = h

    def crystal_rows(self, pe_dev, rows, dst_dev, row_stride: int = 0, stream=None) -> None:
        """Crystal scatter of the listed detector rows into dst rows 0..len(rows)-1."""
        r = np.ascontiguousarray(rows, np.int32)
        s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        check(self._lib.sm_scatter_plan_crystal_rows(self._h, ptr(pe_dev), ptr(r), len(r), ptr(dst_dev), row_stride,
                                                     ctypes.c_void_p(s) if s else None))

    def crystal(self, pe_dev, det_begin: int, det_count: int, dst_dev, row_stride: int = 0, stream=None) -> None:
        s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        check(self._lib.sm_scatter_plan_crystal(self._h, ptr(pe_dev), det_begin, det_count, ptr(dst_dev), row_stride,
                                                ctypes.c_void_p(s) if s else None))

    def collimator(self, det_begin: int, det_count: int, dst_dev, row_stride: int = 0, stream=None) -> None:
        s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        check(self._lib.sm_scatter_plan_collimator(self._h, det_begin, det_count, ptr(dst_dev), row_stride,
                                                   ctypes.c_void_p(s) if s else None))

    def collimator_cells(self, det_begin: int, det_count: int, dst_dev, cell_size: float = 0.5,
                         keep_dist: float = 1.0, row_stride: int = 0, cells_per_chunk: int = 0, stream=None) -> None:
        """Collimator scatter with the plates cut into ``cell_size`` mm cells (each a Compton
        site; in-leg and out-leg plate attenuation included).  ``keep_dist`` < 0 keeps every
        material cell, else the cells within ``keep_dist`` mm of a hole wall."""
        s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        check(self._lib.sm_scatter_plan_cells(self._h, cell_size, keep_dist, det_begin, det_count, ptr(dst_dev),
                                              row_stride, cells_per_chunk, ctypes.c_void_p(s) if s else None))

    def cell_table(self, cell_size: float = 0.5, keep_dist: float = 1.0) -> np.ndarray:
        """Kept cells as an (n, 4) float64 array: x, y (FOV2C added), z, mu_c vol / (4 pi)."""
        n = int(check(self._lib.sm_scatter_cell_table(self._h, cell_size, keep_dist, None, 0)))
        out = np.zeros((max(n, 1), 4), np.float64)
        check(self._lib.sm_scatter_cell_table(self._h, cell_size, keep_dist, ptr(out), n))
        return out[:n]

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sm_scatter_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ScatterBatch:
    """Crystal + collimator scatter of several collimator variants (shared detector and image),
    all launches on one stream, optionally captured once into a hipGraph and replayed."""

    def __init__(self, variants: Sequence[Geometry], rotation: int = 0, device: int = 0):
        self._lib = load()
        g0 = variants[0]
        self._cols = [_f32(g.col) for g in variants]
        self._keep = [_f32(g0.det), _img(g0.img), _f32(g0.phy)]
        arr = (ctypes.c_void_p * len(variants))(*[c.ctypes.data for c in self._cols])
        h = ctypes.c_void_p()
        check(self._lib.sm_scatter_batch_create(len(variants), arr, ptr(self._keep[0]), ptr(self._keep[1]),
                                                ptr(self._keep[2]), rotation, device, ctypes.byref(h)))
        self._h = h
        self.n = len(variants)

    def run(self, pe_devs, dst_devs, use_graph: bool = True, stream=None) -> None:
        pe = (ctypes.c_void_p * self.n)(*[ptr(t) for t in pe_devs])
        dst = (ctypes.c_void_p * self.n)(*[ptr(t) for t in dst_devs])
        s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        check(self._lib.sm_scatter_batch_run(self._h, pe, dst, ctypes.c_void_p(s) if s else None,
                                             1 if use_graph else 0))

    @property
    def graph_nodes(self) -> int:
        return int(self._lib.sm_scatter_batch_graph_nodes(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sm_scatter_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
