"""Analysis of system matrices: beam masks, multiplexing index and PPDF metrics.

Serves the role of the reference's consumer scripts
(PEGen_RayTracing_CircularHole/extract_beam_prop_masks.py, compute_mpxi_3d.py,
ppdf_metrics_sysmat_raw.py), reading the CSR output of this package:

* ``beams_device``: beam masks and MPXI of device-resident CSR rows on the GPU
  (``sm_csr_beams_device``); ``beams_host``: the same from a host CSR (files).
  A detector's beam is the set of voxels with ``v > frac * (row max)``
  (extract_beam_prop_masks.py uses frac = 0.05); the MPXI of a voxel is the
  number of beams that contain it (compute_mpxi_3d.py).
* ``mpxi_summary``: sampled / multiplexed voxel statistics of an MPXI volume.
* ``ppdf_metrics``: per-detector point-projection density: normalised to its
  maximum, support = voxels >= ``support`` of it, 26-connected components of at
  least ``min_voxels``; per component the weighted centroid, the profile through
  it along an axis (trilinear samples every half voxel over +-0.6 of the volume
  diagonal), its FWHM and 25-75 % edge widths (level crossings found from the
  left end up to the peak and from the peak to the right end, linearly
  interpolated), the bounding box and the active volume
  (ppdf_metrics_sysmat_raw.py:31-200 defines the same quantities).

Volumes are indexed (z, y, x) like the reference's viewers; spacings in mm.

Command line (JSON on stdout)::

    python -m sysmat.analysis --csr PE.csr --shape X Y Z --voxel-mm X Y Z --rows 0 17 --frac 0.05
"""
from __future__ import annotations

import argparse
import json
import sys
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import check, load, ptr

BEAM_FRACTION = 0.05


# ---- beams and MPXI ------------------------------------------------------
def beams_device(row_ptr, col, val, n_cols: int, frac: float = BEAM_FRACTION, mpxi=None, stream=None):
    """Beam sizes (int32 per row) and MPXI (int32 per column, accumulated into ``mpxi`` when
    given) of device CSR rows (torch tensors: row_ptr int64, col int32, val float32)."""
    import ctypes
    import torch
    rows = int(row_ptr.numel()) - 1
    dev = row_ptr.device
    if mpxi is None:
        mpxi = torch.zeros(n_cols, dtype=torch.int32, device=dev)
    size = torch.empty(max(rows, 1), dtype=torch.int32, device=dev)
    s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
    check(load().sm_csr_beams_device(ptr(row_ptr), ptr(col), ptr(val), rows, float(frac), ptr(mpxi), ptr(size),
                                     ctypes.c_void_p(s) if s else None))
    return size[:rows], mpxi


def beams_host(m, frac: float = BEAM_FRACTION, rows: Optional[Sequence[int]] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Beam sizes and MPXI of a host CSR (``sysmat.csr.CsrMatrix``), optionally of a row subset."""
    rp = np.asarray(m.row_ptr, np.int64)
    n_rows, n_cols = m.shape
    rows = np.arange(n_rows) if rows is None else np.asarray(rows, np.int64)
    mpxi = np.zeros(n_cols, np.int64)
    sizes = np.zeros(len(rows), np.int64)
    for i, r in enumerate(rows):
        v = np.asarray(m.val[rp[r]:rp[r + 1]], np.float64)
        if v.size == 0:
            continue
        mx = v.max()
        if not mx > 0:
            continue
        sel = v > frac * mx
        sizes[i] = int(sel.sum())
        np.add.at(mpxi, np.asarray(m.col[rp[r]:rp[r + 1]])[sel], 1)
    return sizes, mpxi


def mpxi_summary(mpxi) -> Dict:
    """Voxel statistics of an MPXI volume (any shape)."""
    v = np.asarray(mpxi).ravel().astype(np.int64)
    nz = v[v > 0]
    vals, counts = np.unique(v, return_counts=True)
    return {
        "voxels": int(v.size),
        "sampled": int(nz.size),
        "unsampled": int(v.size - nz.size),
        "single": int(np.sum(v == 1)),
        "multiplexed": int(np.sum(v > 1)),
        "max": int(v.max()) if v.size else 0,
        "mean_all": float(v.mean()) if v.size else 0.0,
        "mean_sampled": float(nz.mean()) if nz.size else 0.0,
        "median_sampled": float(np.median(nz)) if nz.size else 0.0,
        "histogram": [[int(a), int(b)] for a, b in zip(vals, counts)],
    }


# ---- PPDF metrics ---------------------------------------------------------
def _crossing(x: np.ndarray, y: np.ndarray, level: float, rising: bool) -> Optional[float]:
    """First position along x where y crosses `level` (upward if rising, else downward), linear."""
    a, b = y[:-1], y[1:]
    hit = ((a <= level) & (level <= b)) if rising else ((a >= level) & (level >= b))
    idx = np.flatnonzero(hit)
    if idx.size == 0:
        return None
    i = int(idx[0])
    if b[i] == a[i]:
        return float(x[i])
    return float(x[i] + (level - a[i]) / (b[i] - a[i]) * (x[i + 1] - x[i]))


def profile_widths(s_mm: np.ndarray, prof: np.ndarray) -> Dict[str, float]:
    """FWHM, 25-75 % edge widths and peak position of a 1-D profile."""
    prof = np.asarray(prof, np.float64)
    s_mm = np.asarray(s_mm, np.float64)
    out = {"fwhm_mm": float("nan"), "edge_left_mm": float("nan"), "edge_right_mm": float("nan"),
           "peak_mm": float("nan")}
    if prof.size == 0 or not prof.max() > 0:
        return out
    p = prof / prof.max()
    k = int(np.argmax(p))
    out["peak_mm"] = float(s_mm[k])
    lx, lp, rx, rp = s_mm[:k + 1], p[:k + 1], s_mm[k:], p[k:]
    l50, r50 = _crossing(lx, lp, 0.5, True), _crossing(rx, rp, 0.5, False)
    if l50 is not None and r50 is not None:
        out["fwhm_mm"] = r50 - l50
    l25, l75 = _crossing(lx, lp, 0.25, True), _crossing(lx, lp, 0.75, True)
    if l25 is not None and l75 is not None:
        out["edge_left_mm"] = abs(l75 - l25)
    r75, r25 = _crossing(rx, rp, 0.75, False), _crossing(rx, rp, 0.25, False)
    if r75 is not None and r25 is not None:
        out["edge_right_mm"] = abs(r25 - r75)
    return out


def line_profile(vol: np.ndarray, centre_zyx, axis_zyx, spacing_zyx, half_mm: Optional[float] = None,
                 step_mm: Optional[float] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Trilinear samples of `vol` along the unit mm-space direction `axis_zyx` through `centre_zyx`
    (voxel indices): +-half_mm (default 0.6 x volume diagonal) every step_mm (default half the
    smallest voxel side)."""
    from scipy.ndimage import map_coordinates
    sp = np.asarray(spacing_zyx, np.float64)
    ax = np.asarray(axis_zyx, np.float64)
    ax = ax / np.linalg.norm(ax)
    if half_mm is None:
        half_mm = 0.6 * float(np.linalg.norm(np.asarray(vol.shape) * sp))
    if step_mm is None:
        step_mm = 0.5 * float(sp.min())
    s = np.arange(-half_mm, half_mm + step_mm, step_mm)
    pts = np.asarray(centre_zyx, np.float64)[:, None] + (ax / sp)[:, None] * s[None, :]
    return s, map_coordinates(np.asarray(vol, np.float64), pts, order=1, mode="constant", cval=0.0)


def ppdf_metrics(vol, spacing_zyx=(1.0, 1.0, 1.0), axis_zyx=(0.0, 0.0, 1.0), support: float = 0.10,
                 min_voxels: int = 5) -> List[Dict]:
    """Per-component metrics of one detector's PPDF volume (z, y, x)."""
    from scipy.ndimage import center_of_mass, label
    v = np.maximum(np.asarray(vol, np.float64), 0.0)
    mx = v.max() if v.size else 0.0
    if not mx > 0:
        return []
    v = v / mx
    lab, n = label(v >= support, structure=np.ones((3, 3, 3), np.int32))
    sp = np.asarray(spacing_zyx, np.float64)
    out = []
    for c in range(1, n + 1):
        mask = lab == c
        cnt = int(mask.sum())
        if cnt < min_voxels:
            continue
        w = v * mask
        centre = np.asarray(center_of_mass(w)) if w.sum() > 0 else np.argwhere(mask).mean(axis=0)
        s, prof = line_profile(w, centre, axis_zyx, sp)
        m = profile_widths(s, prof)
        idx = np.argwhere(mask)
        span = (idx.max(axis=0) - idx.min(axis=0) + 1) * sp
        m.update({"voxels": cnt, "active_volume_mm3": float(cnt * np.prod(sp)),
                  "bbox_mm_zyx": [float(x) for x in span],
                  "centroid_zyx": [float(x) for x in centre]})
        out.append(m)
    return out


def row_volume(m, row: int, shape_xyz: Sequence[int]) -> np.ndarray:
    """One CSR row as a dense (z, y, x) volume."""
    X, Y, Z = (int(a) for a in shape_xyz)
    out = np.zeros(X * Y * Z, np.float32)
    rp = m.row_ptr
    out[np.asarray(m.col[rp[row]:rp[row + 1]])] = m.val[rp[row]:rp[row + 1]]
    return out.reshape(Z, Y, X)


def main(argv=None) -> int:
    from .csr import read_csr
    ap = argparse.ArgumentParser(description="Beams, MPXI and PPDF metrics of a .csr system matrix")
    ap.add_argument("--csr", required=True)
    ap.add_argument("--shape", nargs=3, type=int, required=True, metavar=("X", "Y", "Z"))
    ap.add_argument("--voxel-mm", nargs=3, type=float, default=(1.0, 1.0, 1.0), metavar=("X", "Y", "Z"))
    ap.add_argument("--rows", nargs="*", type=int, default=[0])
    ap.add_argument("--frac", type=float, default=BEAM_FRACTION)
    ap.add_argument("--axis-zyx", nargs=3, type=float, default=(0.0, 0.0, 1.0))
    ap.add_argument("--support", type=float, default=0.10)
    ap.add_argument("--min-voxels", type=int, default=5)
    a = ap.parse_args(argv)
    m = read_csr(a.csr, mmap=True)
    X, Y, Z = a.shape
    if m.shape[1] != X * Y * Z:
        print(f"shape {X}x{Y}x{Z} does not match the {m.shape[1]} columns", file=sys.stderr)
        return 2
    sizes, mpxi = beams_host(m, a.frac)
    sp = (a.voxel_mm[2], a.voxel_mm[1], a.voxel_mm[0])
    rec = {"rows": int(m.shape[0]), "beam_fraction": a.frac,
           "beam_size_mean": float(sizes.mean()) if sizes.size else 0.0,
           "mpxi": mpxi_summary(mpxi),
           "ppdf": {str(r): ppdf_metrics(row_volume(m, r, a.shape), sp, a.axis_zyx, a.support, a.min_voxels)
                    for r in a.rows}}
    print(json.dumps(rec))
    return 0


if __name__ == "__main__":
    sys.exit(main())
