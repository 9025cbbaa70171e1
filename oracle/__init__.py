"""ctypes front-end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The shipped library
(``gpu-spect-pet-sysmat_amd``) never imports this module.

Functions mirror the reference kernels (file:line cited in the C++ sources):
``pe`` restates photodetectorCudaMe (PESysMatGen.cu:375-764) in FP32 / FP64,
``pe_count`` is its op-counting build.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


class PeOpts(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int), ("cx", ctypes.c_int), ("cy", ctypes.c_int), ("cz", ctypes.c_int),
                ("det_begin", ctypes.c_int), ("det_end", ctypes.c_int), ("threads", ctypes.c_int),
                ("flags", ctypes.c_int), ("jitter", ctypes.c_double), ("jseed", ctypes.c_int)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with make (gcc)."""
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        for name in ("or_pe_f32", "or_pe_f64"):
            fn = getattr(L, name)
            fn.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.POINTER(PeOpts)]
            fn.restype = ctypes.c_int
        L.or_pe_count.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.POINTER(PeOpts), _f64p]
        L.or_pe_count.restype = ctypes.c_int
        for name, rt, at in (("or_chord_box_f32", ctypes.c_float, _f32p), ("or_chord_box_f64", ctypes.c_double, _f64p),
                             ("or_chord_box_open_f32", ctypes.c_float, _f32p),
                             ("or_chord_box_open_f64", ctypes.c_double, _f64p),
                             ("or_chord_hole_f32", ctypes.c_float, _f32p), ("or_chord_hole_f64", ctypes.c_double, _f64p)):
            fn = getattr(L, name)
            fn.argtypes = [at]
            fn.restype = rt
        _init_scatter(L)
        _lib = L
    return _lib


class ScOpts(ctypes.Structure):
    _fields_ = [("det_begin", ctypes.c_int), ("det_end", ctypes.c_int), ("threads", ctypes.c_int),
                ("variant", ctypes.c_int), ("vox", ctypes.c_void_p), ("n_vox", ctypes.c_longlong)]


def _init_scatter(L) -> None:
    ll, i, vp = ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p
    L.or_crystal_bitmap_words.argtypes = [i]
    L.or_crystal_bitmap_words.restype = ll
    L.or_collimator_bitmap_words.argtypes = [i, i]
    L.or_collimator_bitmap_words.restype = ll
    L.or_bitmap_crystal.argtypes = [_f32p, _u32p, ll]
    L.or_bitmap_collimator.argtypes = [_f32p, _f32p, _u32p, ll]
    L.or_crystal_bits_for_pairs.argtypes = [_f32p, vp, vp, i, vp]
    L.or_scatter_crystal.argtypes = [_f32p, _f32p, _f32p, _f32p, vp, _f32p, ctypes.POINTER(ScOpts), i]
    L.or_scatter_collimator.argtypes = [_f32p, _f32p, _f32p, _f32p, vp, _f32p, ctypes.POINTER(ScOpts), i]
    L.or_cells.argtypes = [_f32p, ctypes.c_float, ctypes.c_double, ctypes.c_double, vp, ll]
    L.or_cells.restype = ll
    L.or_plate_path_f64.argtypes = [_f32p, ctypes.c_float, _f64p, _f64p]
    L.or_scatter_cells.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.c_double, ctypes.c_double, _f32p,
                                   ctypes.POINTER(ScOpts), i]
    L.or_kn_riemann_f32.argtypes = [ctypes.c_float] * 3
    L.or_kn_riemann_f32.restype = ctypes.c_float
    L.or_kn_riemann_f64.argtypes = [ctypes.c_double] * 3
    L.or_kn_riemann_f64.restype = ctypes.c_double
    L.or_scatter_energy_f32.argtypes = [ctypes.c_float] * 2
    L.or_scatter_energy_f32.restype = ctypes.c_float
    L.or_window_probability_f32.argtypes = [ctypes.c_float] * 4
    L.or_window_probability_f32.restype = ctypes.c_float
    L.or_cone_angle_f32.argtypes = [_f32p]
    L.or_cone_angle_f32.restype = ctypes.c_float


def bitmap_crystal(geo) -> np.ndarray:
    """GeometryRelationShip_Crystal2Crystal words (geometryRelationShip_Crystal2Crystal, scatter.cu:1440-1540)."""
    n = int(geo.det[0])
    w = np.zeros(lib().or_crystal_bitmap_words(n), np.uint32)
    lib().or_bitmap_crystal(np.ascontiguousarray(geo.det, np.float32), w, w.size)
    return w


def bitmap_collimator(geo) -> np.ndarray:
    """GeometryRelationShip_Collimator2Crystal words (scatter.cu:1939-2009)."""
    n, nh = int(geo.det[0]), int(geo.col[10])
    w = np.zeros(lib().or_collimator_bitmap_words(nh, n), np.uint32)
    lib().or_bitmap_collimator(np.ascontiguousarray(geo.col, np.float32), np.ascontiguousarray(geo.det, np.float32),
                               w, w.size)
    return w


def _sc_img(geo, rot):
    img = np.zeros(100, np.float32)
    img[:len(geo.img)] = geo.img[:100]
    img[20] = rot
    return img


def _vox(vox):
    if vox is None:
        return None, None, 0
    v = np.ascontiguousarray(vox, np.int32)
    return v, v.ctypes.data, int(v.size)


def scatter_crystal(geo, pe_sysmat: np.ndarray, bits: Optional[np.ndarray], det_range=None, precision="f32", rot=0,
                    threads=0, vox=None) -> np.ndarray:
    """crystalScatterSysMatCuda rows (scatter.cu:1077-1437); pe_sysmat is the N x V PE matrix,
    or N x len(vox) columns when the voxel sample ``vox`` is given (rows then hold those voxels).
    bits None: the GeometryRelationShip criterion is evaluated inline (same lists)."""
    b, e = det_range if det_range is not None else (0, geo.n_det)
    keep, vp, nv = _vox(vox)
    V = nv if vox is not None else geo.n_vox
    dst = np.zeros((e - b) * V, np.float32)
    o = ScOpts(b, e, threads, 0, vp, nv)
    bw = None if bits is None else np.ascontiguousarray(bits, np.uint32)
    lib().or_scatter_crystal(np.ascontiguousarray(geo.det, np.float32), _sc_img(geo, rot),
                             np.ascontiguousarray(geo.phy, np.float32),
                             np.ascontiguousarray(pe_sysmat, np.float32).ravel(),
                             None if bw is None else bw.ctypes.data, dst, ctypes.byref(o),
                             1 if precision == "f64" else 0)
    return dst.reshape(e - b, V)


def scatter_collimator(geo, bits: Optional[np.ndarray], det_range=None, precision="f32", variant=0, rot=0,
                       threads=0, vox=None) -> np.ndarray:
    """collimatorScatterSysMatCuda rows (variant 0: scatter.cu:1544-1936; variant 1: the
    ScatterGen_Collimator kernel, ScatterGen_Collimator/scatter.cu:1056-1436); voxel sample
    and bits None as scatter_crystal."""
    b, e = det_range if det_range is not None else (0, geo.n_det)
    keep, vp, nv = _vox(vox)
    V = nv if vox is not None else geo.n_vox
    dst = np.zeros((e - b) * V, np.float32)
    o = ScOpts(b, e, threads, variant, vp, nv)
    bw = None if bits is None else np.ascontiguousarray(bits, np.uint32)
    lib().or_scatter_collimator(np.ascontiguousarray(geo.col, np.float32), np.ascontiguousarray(geo.det, np.float32),
                                _sc_img(geo, rot), np.ascontiguousarray(geo.phy, np.float32),
                                None if bw is None else bw.ctypes.data, dst, ctypes.byref(o),
                                1 if precision == "f64" else 0)
    return dst.reshape(e - b, V)


def crystal_bits_for_pairs(geo, pairs) -> np.ndarray:
    """(len(pairs), N) uint8: bit (pair(i, j), k) of the crystal bitmap for each (i < j)."""
    p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    assert np.all(p[:, 0] < p[:, 1])
    pi, pj = np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1])
    out = np.zeros((len(p), geo.n_det), np.uint8)
    lib().or_crystal_bits_for_pairs(np.ascontiguousarray(geo.det, np.float32), pi.ctypes.data, pj.ctypes.data,
                                    len(p), out.ctypes.data)
    return out


def cells(geo, h: float, keep: float) -> np.ndarray:
    """Collimator plate cells (x, y with FOV2C, z, mu_c vol / 4 pi), (n, 4) float64."""
    col = np.ascontiguousarray(geo.col, np.float32)
    n = lib().or_cells(col, float(geo.img[11]), h, keep, None, 0)
    out = np.zeros((max(n, 1), 4), np.float64)
    lib().or_cells(col, float(geo.img[11]), h, keep, out.ctypes.data, n)
    return out[:n]


def plate_path(geo, a, b) -> np.ndarray:
    """mu_t, mu_pe, mu_c weighted path lengths of segment a -> b through the plates and holes."""
    ab = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64), np.asarray(b, np.float64)]))
    out = np.zeros(3, np.float64)
    lib().or_plate_path_f64(np.ascontiguousarray(geo.col, np.float32), float(geo.img[11]), ab, out)
    return out


def scatter_cells(geo, h: float, keep: float, det_range=None, precision="f64", rot=0, threads=0) -> np.ndarray:
    """Plate-cell collimator scatter rows (the cell model of the shipped library's
    sm_scatter_plan_cells; no reference kernel)."""
    b, e = det_range if det_range is not None else (0, geo.n_det)
    V = geo.n_vox
    dst = np.zeros((e - b) * V, np.float32)
    o = ScOpts(b, e, threads, 0)
    lib().or_scatter_cells(np.ascontiguousarray(geo.col, np.float32), np.ascontiguousarray(geo.det, np.float32),
                           _sc_img(geo, rot), np.ascontiguousarray(geo.phy, np.float32), h, keep, dst,
                           ctypes.byref(o), 1 if precision == "f64" else 0)
    return dst.reshape(e - b, V)


def slab_image(img: np.ndarray, z0: int = 0, slab_z: Optional[int] = None, rot: int = 0) -> np.ndarray:
    """Image array with the PE driver's in-band slab fields filled
    (PEGen_CircularHole.cpp:165,177,189-190): [2] slab Z, [30] full Z, [31] z0, [20] rotation."""
    out = np.array(img, np.float32, copy=True)
    full_z = int(np.floor(img[2] + 0.001))
    out[30] = full_z
    out[2] = full_z - z0 if slab_z is None else slab_z
    out[31] = z0
    out[20] = rot
    return out


def _opts(nv, sub, det_range, n_det, threads, flags=0, jitter=0.0, jseed=0) -> PeOpts:
    b, e = det_range if det_range is not None else (0, n_det)
    return PeOpts(nv, sub[0], sub[1], sub[2], b, e, threads, flags, jitter, jseed)


def pe(geo, nv: int = 1, sub: Tuple[int, int, int] = (8, 16, 8), det_range=None, precision: str = "f32",
       z0: int = 0, slab_z: Optional[int] = None, rot: int = 0, threads: int = 0, flags: int = 0,
       jitter: float = 0.0, jseed: int = 0, full_z: Optional[int] = None) -> np.ndarray:
    """PE sysmat rows [det_range) x slab voxels, det-major, float32.

    full_z overrides the in-band full Z (img[30]) that places the voxels.

    precision: "f32" (reference restatement), "f64" (accuracy oracle) or "mixed"
    (FP32 tracing with an FP64 element accumulator, an FP64 collimator path fed
    by FP64 sub-ray end points, and an FP64 absorption chord: flags 39, the
    numerics of the HIP kernel)."""
    if precision == "mixed":
        precision, flags = "f32", 39
    img = slab_image(geo.img, z0, slab_z, rot)
    if full_z is not None:
        img[30] = full_z
    X, Y = int(np.floor(img[0])), int(np.floor(img[1]))
    bins = X * Y * int(np.floor(img[2]))
    o = _opts(nv, sub, det_range, geo.n_det, threads, flags, jitter, jseed)
    dst = np.zeros((o.det_end - o.det_begin) * bins, np.float32)
    fn = lib().or_pe_f32 if precision == "f32" else lib().or_pe_f64
    fn(np.ascontiguousarray(geo.col, np.float32), np.ascontiguousarray(geo.det, np.float32), img, dst, ctypes.byref(o))
    return dst.reshape(o.det_end - o.det_begin, bins)


def pe_count(geo, nv: int = 1, sub=(8, 16, 8), det_range=None, threads: int = 0, z0: int = 0,
             slab_z: Optional[int] = None, touched: bool = False):
    """(values, per-row stats[rows, 5]) with stats = counted FLOPs, culled crystals, culled holes,
    and (touched=True) culled crystals / holes with a non-zero chord on some sub-ray of their
    element (sums over the row's elements of the slab).  touched=True counts only the work on
    those attenuators (the others add exactly zero); False counts every culled attenuator."""
    img = slab_image(geo.img, z0, slab_z)
    bins = int(np.floor(img[0])) * int(np.floor(img[1])) * int(np.floor(img[2]))
    o = _opts(nv, sub, det_range, geo.n_det, threads, 64 if touched else 0)
    rows = o.det_end - o.det_begin
    dst = np.zeros(rows * bins, np.float32)
    stats = np.zeros(rows * 5, np.float64)
    lib().or_pe_count(np.ascontiguousarray(geo.col, np.float32), np.ascontiguousarray(geo.det, np.float32), img, dst,
                      ctypes.byref(o), stats)
    return dst.reshape(rows, bins), stats.reshape(rows, 5)


def chord_box(args, precision="f32", open_end=False) -> float:
    dt = np.float32 if precision == "f32" else np.float64
    a = np.ascontiguousarray(args, dt)
    name = "or_chord_box" + ("_open" if open_end else "") + ("_f32" if precision == "f32" else "_f64")
    return float(getattr(lib(), name)(a))


def chord_hole(args, precision="f32") -> float:
    dt = np.float32 if precision == "f32" else np.float64
    a = np.ascontiguousarray(args, dt)
    return float((lib().or_chord_hole_f32 if precision == "f32" else lib().or_chord_hole_f64)(a))


def pe_sensitive(geo, ref: np.ndarray, jitter: float = 2e-5, trials: int = 3, tol: float = 2e-6, **kw) -> np.ndarray:
    """Conditioning classifier: True where an element of the FP64 oracle moves by
    more than ``tol`` (relative) when every sub-ray end point is displaced by up
    to ``jitter`` mm -- i.e. some sub-ray sits within FP32 rounding distance of
    one of the reference algorithm's branch thresholds (eps = 0.001 chords,
    Delta <= 1e-5 hole-wall grazing, t <= 1e-4, the 0.1 mm plate-miss test,
    the cull predicate).  Such elements are discontinuous at the rounding level
    and cannot be pinned to 1e-5 by any FP32 evaluation, the reference's own
    included; they are reported separately."""
    ref = np.asarray(ref, np.float64)
    flag = np.zeros(ref.shape, bool)
    for t in range(1, trials + 1):
        alt = pe(geo, precision="f64", jitter=jitter, jseed=t, **kw).astype(np.float64)
        flag |= np.abs(alt - ref) > tol * np.abs(ref)
    return flag


def csr(dense: np.ndarray, rel: float = 1e-9):
    """Row-wise CSR of a dense matrix, keeping v > 0 with float64(v) >= rel * float64(max(0, row max)).

    Returns (row_ptr int64, col int32, val float32).  The reference writes dense
    rows only (PEGen_CircularHole.cpp:201-216); this threshold rule is the sparse
    output stage's definition, pinned by a brute-force loop in the tests."""
    d = np.ascontiguousarray(dense, np.float32)
    rmax = np.maximum(np.nan_to_num(d, nan=0.0).max(axis=1, initial=0.0), 0.0) if d.shape[1] else np.zeros(d.shape[0])
    thr = rel * rmax.astype(np.float64)
    with np.errstate(invalid="ignore"):
        keep = (d > 0) & (d.astype(np.float64) >= thr[:, None])
    counts = keep.sum(axis=1)
    row_ptr = np.zeros(d.shape[0] + 1, np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    rows, cols = np.nonzero(keep)
    return row_ptr, cols.astype(np.int32), d[rows, cols]
