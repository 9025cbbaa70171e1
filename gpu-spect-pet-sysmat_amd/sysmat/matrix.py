"""Readers for the dense system-matrix files the drivers write.

Layout (PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:201-216): raw
little-endian float32, no header, element ((rot * N + det) * V + z * X * Y +
y * X + x).  The reference's viewers memory-map the file and reshape it
det-major to (N, Z, Y, X) (visualize_sysmat_per_detector.py:77-90,
viz_ppdf_slices_6x3planes.py:103-111); ``order="vox_major"`` reads the
transposed layout those viewers also accept.  Sparse ``.csr`` files:
``sysmat.csr.read_csr``.
"""
from __future__ import annotations

import os
from typing import Sequence, Tuple

import numpy as np

from .params import Geometry


def sysmat_name(prefix: str, geo: Geometry) -> str:
    """``<prefix>_shift_<x>_<y>_<z>.sysmat`` as the drivers name their outputs (%f of the FOV shift)."""
    x, y, z = (float(np.float32(v)) for v in geo.img[8:11])
    return f"{prefix}_shift_{x:f}_{y:f}_{z:f}.sysmat"


def open_sysmat(path: str, n_det: int, shape: Sequence[int], n_rot: int = 1, order: str = "det_major",
                mode: str = "r") -> np.memmap:
    """Memory-map a dense system matrix.

    det-major: returns (n_rot, n_det, Z, Y, X); vox-major: (n_rot, Z, Y, X, n_det).
    ``shape`` is (X, Y, Z).  The file size must match exactly."""
    X, Y, Z = (int(v) for v in shape)
    V = X * Y * Z
    want = int(n_rot) * int(n_det) * V * 4
    have = os.path.getsize(path)
    if have != want:
        raise ValueError(f"{path}: {have} bytes, expected {want} ({n_rot} x {n_det} x {V} float32)")
    mm = np.memmap(path, dtype="<f4", mode=mode)
    if order == "det_major":
        return mm.reshape(n_rot, n_det, Z, Y, X)
    if order == "vox_major":
        return mm.reshape(n_rot, Z, Y, X, n_det)
    raise ValueError("order must be 'det_major' or 'vox_major'")


def open_for(geo: Geometry, directory: str = ".", prefix: str = "PE_SysMat", mode: str = "r") -> np.memmap:
    """The driver output ``prefix`` for ``geo`` in ``directory``: (rotations, N, Z, Y, X)."""
    return open_sysmat(os.path.join(directory, sysmat_name(prefix, geo)), geo.n_det, geo.shape,
                       max(1, geo.n_rot), mode=mode)


def ppdf(mat: np.ndarray, det: int, rot: int = 0, order: str = "det_major") -> np.ndarray:
    """Point-projection density of one detector: its (Z, Y, X) volume."""
    if order == "det_major":
        return np.asarray(mat[rot, det])
    return np.asarray(mat[rot, ..., det])


def sensitivity(mat: np.ndarray, rot: int = 0, order: str = "det_major", block: int = 256) -> np.ndarray:
    """Sum over detectors per voxel (Z, Y, X), streamed in detector blocks."""
    if order != "det_major":
        return np.asarray(mat[rot], np.float64).sum(axis=-1)
    n = mat.shape[1]
    acc = np.zeros(mat.shape[2:], np.float64)
    for d0 in range(0, n, block):
        acc += np.asarray(mat[rot, d0:d0 + block], np.float64).sum(axis=0)
    return acc


def mips(vol: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Maximum-intensity projections of a (Z, Y, X) volume along z, y and x."""
    v = np.asarray(vol)
    return v.max(axis=0), v.max(axis=1), v.max(axis=2)
