"""Sparse (CSR) output of system-matrix rows.

An element ``v`` of a row is kept when ``v > 0`` and
``float64(v) >= rel * float64(max(0, row max))`` (default ``rel = 1e-9``); the
columns of a row ascend.  The GPU stage (csrc/csr.hip) compacts dense device
rows with wavefront ballots; this module binds it and reads/writes ``.csr``
files:

    64-byte header: b"SMCSR\\0\\0\\x01", int64 rows, int64 cols, int64 nnz,
                    float64 rel, 24 zero bytes
    row_ptr[rows+1] int64, col[nnz] int32, val[nnz] float32   (little endian)

The reference has no sparse output (its only writer is the dense one,
PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:201-216); the dense
``.sysmat`` layout stays the drop-in default.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional

import numpy as np

from ._lib import check, load, ptr
from .params import Geometry
from .pe import _f32, _padded_img, options

MAGIC = b"SMCSR\x00\x00\x01"
HEADER = 64
DEFAULT_REL = 1e-9


@dataclasses.dataclass
class CsrMatrix:
    row_ptr: np.ndarray  # int64, rows + 1
    col: np.ndarray      # int32, nnz
    val: np.ndarray      # float32, nnz
    shape: tuple
    rel: float = DEFAULT_REL

    @property
    def nnz(self) -> int:
        return int(self.col.size)

    def to_dense(self) -> np.ndarray:
        out = np.zeros(self.shape, np.float32)
        rows = np.repeat(np.arange(self.shape[0]), np.diff(self.row_ptr))
        out[rows, self.col] = self.val
        return out

    def to_scipy(self):
        import scipy.sparse
        return scipy.sparse.csr_matrix((self.val, self.col, self.row_ptr), shape=self.shape)


def write_csr(path: str, m: CsrMatrix) -> None:
    head = bytearray(HEADER)
    head[0:8] = MAGIC
    head[8:40] = np.array([m.shape[0], m.shape[1], m.nnz], "<i8").tobytes() + np.float64(m.rel).tobytes()
    with open(path, "wb") as f:
        f.write(bytes(head))
        f.write(np.ascontiguousarray(m.row_ptr, "<i8").tobytes())
        f.write(np.ascontiguousarray(m.col, "<i4").tobytes())
        f.write(np.ascontiguousarray(m.val, "<f4").tobytes())


def read_csr(path: str, mmap: bool = False) -> CsrMatrix:
    """Read a .csr file (``mmap=True`` maps the arrays instead of loading them)."""
    with open(path, "rb") as f:
        head = f.read(HEADER)
    if len(head) != HEADER or head[:8] != MAGIC:
        raise ValueError(f"{path}: not a .csr file")
    rows, cols, nnz = (int(x) for x in np.frombuffer(head[8:32], "<i8"))
    rel = float(np.frombuffer(head[32:40], "<f8")[0])
    mode = "r" if mmap else None
    off = HEADER
    if mmap:
        rp = np.memmap(path, "<i8", mode, off, (rows + 1,))
        off += 8 * (rows + 1)
        col = np.memmap(path, "<i4", mode, off, (nnz,)) if nnz else np.zeros(0, np.int32)
        off += 4 * nnz
        val = np.memmap(path, "<f4", mode, off, (nnz,)) if nnz else np.zeros(0, np.float32)
    else:
        with open(path, "rb") as f:
            f.seek(off)
            rp = np.fromfile(f, "<i8", rows + 1)
            col = np.fromfile(f, "<i4", nnz)
            val = np.fromfile(f, "<f4", nnz)
        if rp.size != rows + 1 or col.size != nnz or val.size != nnz:
            raise ValueError(f"{path}: truncated")
    return CsrMatrix(rp, col, val, (rows, cols), rel)


class _Handle:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        if self.h:
            load().sm_csr_destroy(self.h)
            self.h = None


def _from_handle(h) -> CsrMatrix:
    L = load()
    rows, cols, nnz = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
    check(L.sm_csr_info(h, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz)))
    rp = np.zeros(rows.value + 1, np.int64)
    col = np.zeros(nnz.value, np.int32)
    val = np.zeros(nnz.value, np.float32)
    check(L.sm_csr_copy(h, ptr(rp), ptr(col) if nnz.value else None, ptr(val) if nnz.value else None))
    return CsrMatrix(rp, col, val, (rows.value, cols.value))


def pe_compute_csr(geo: Geometry, rel: float = DEFAULT_REL, path: Optional[str] = None, **kw) -> CsrMatrix:
    """PE rows (options as ``pe.pe_compute``) kept as CSR; with ``path`` also written as a .csr file."""
    L = load()
    o = options(**kw)
    col, det, img = _f32(geo.col), _f32(geo.det), _padded_img(geo.img)
    h = ctypes.c_void_p()
    check(L.sm_pe_compute_csr(ptr(col), ptr(det), ptr(img), ctypes.byref(o), float(rel), ctypes.byref(h)))
    keep = _Handle(h)
    m = _from_handle(h)
    m.rel = float(rel)
    if path is not None:
        arr = (ctypes.c_void_p * 1)(h.value)
        check(L.sm_csr_write(path.encode(), arr, 1))
    del keep
    return m


def csr_from_dense_device(dense, rel: float = DEFAULT_REL, stream=None):
    """GPU compaction of a 2-D float32 torch tensor on the device.

    Returns ``(row_ptr int64, col int32, val float32, row_max float32)`` torch
    tensors on the same device."""
    import torch
    if dense.dtype != torch.float32 or dense.dim() != 2 or not dense.is_cuda:
        raise ValueError("dense must be a 2-D float32 device tensor")
    if dense.stride(1) != 1:
        raise ValueError("dense rows must be contiguous")
    rows, cols = dense.shape
    L = load()
    s = torch.cuda.current_stream(dense.device) if stream is None else stream
    sh = ctypes.c_void_p(int(s.cuda_stream)) if int(s.cuda_stream) else None
    rmax = torch.empty(rows, dtype=torch.float32, device=dense.device)
    rp = torch.empty(rows + 1, dtype=torch.int64, device=dense.device)
    nnz = ctypes.c_longlong()
    check(L.sm_csr_count_device(ptr(dense), rows, cols, dense.stride(0), float(rel), ptr(rmax), ptr(rp),
                                ctypes.byref(nnz), sh))
    col = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dense.device)
    val = torch.empty(max(nnz.value, 1), dtype=torch.float32, device=dense.device)
    if nnz.value:
        check(L.sm_csr_fill_device(ptr(dense), rows, cols, dense.stride(0), float(rel), ptr(rmax), ptr(rp), ptr(col),
                                   ptr(val), sh))
    return rp, col[:nnz.value], val[:nnz.value], rmax
