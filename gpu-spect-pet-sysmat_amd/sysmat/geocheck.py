"""Geometry sanity report for a Params_*.dat set.

Serves the role of the reference's parameter checker
(PEGen_RayTracing_CircularHole/check_geometry_sanity.py: image/FOV box, plate and
hole summary, detector extents and layer planes, FOV-versus-plate position)
and adds the checks that matter for the kernels of this package:

* hole faces vs plate box: the tracer takes hole faces as absolute
  ``[y1, y2] + FOV2C`` (PESysMatGen.cu:559-561) while the plate box is
  ``FOV2C + offset +- t/2`` (PESysMatGen.cu:650-651); holes that do not span
  their plate leave plate material in the hole (or subtract nothing);
* mu_pe + mu_c vs mu_total of plates, holes and crystals;
* crystals with mu_total <= 0.01 (their rows are zero, PESysMatGen.cu:504, 598);
* crystals inside a plate's y-range or in front of the collimator;
* detector record count vs file length, and N > 6666 (the reference drivers'
  fixed 80000-float buffers hold at most 6666 crystals; this package reads the
  counts from the headers).

Coordinates are reported in the collimator frame: layer-0 plate centred on
y = 0, image centre at y = shift_y - FOV2C, crystals at their stored y.

Command line::

    python -m sysmat.geocheck --param_collimator Params_Collimator.dat \\
        --param_detector Params_Detector.dat --param_image Params_Image.dat
"""
from __future__ import annotations

import argparse
import dataclasses
import sys
from typing import List, Optional

import numpy as np

from .params import DET_STRIDE, HOLE_BASE, HOLE_STRIDE, MAX_LAYERS, Geometry

REFERENCE_MAX_CRYSTALS = (80000 - 1) // DET_STRIDE


@dataclasses.dataclass
class Report:
    lines: List[str] = dataclasses.field(default_factory=list)
    warnings: List[str] = dataclasses.field(default_factory=list)
    errors: List[str] = dataclasses.field(default_factory=list)

    def text(self) -> str:
        out = list(self.lines)
        out += [f"WARNING: {w}" for w in self.warnings]
        out += [f"ERROR: {e}" for e in self.errors]
        out.append("OK" if not (self.warnings or self.errors) else
                   f"{len(self.errors)} error(s), {len(self.warnings)} warning(s)")
        return "\n".join(out)


def _mu_split(tag: str, mt: float, mpe: float, mc: float, rep: Report) -> None:
    if mt > 0 and abs((mpe + mc) - mt) > 0.01 * mt + 1e-6:
        rep.warnings.append(f"{tag}: mu_pe + mu_c = {mpe + mc:g} differs from mu_total = {mt:g}")


def check(geo: Geometry, det_file_floats: Optional[int] = None, max_holes_print: int = 5) -> Report:
    """Build the report for ``geo``; ``det_file_floats`` = length of the detector file
    when it was read from disk (checked against the crystal count)."""
    rep = Report()
    img = geo.img.astype(np.float64)
    X, Y, Z = geo.shape
    vox = img[3:6]
    fov2c = img[11]
    size = np.array([X, Y, Z]) * vox
    centre = np.array([img[8], img[9] - fov2c, img[10]])
    lo, hi = centre - size / 2, centre + size / 2
    rep.lines.append("== image ==")
    rep.lines.append(f"voxels {X} x {Y} x {Z} of {vox[0]:g} x {vox[1]:g} x {vox[2]:g} mm, FOV {size.tolist()} mm")
    rep.lines.append(f"FOV centre {centre.tolist()} (FOV2C {fov2c:g}), box {lo.tolist()} .. {hi.tolist()}")
    rep.lines.append(f"rotations {geo.n_rot} x {img[7]:g} rad")
    if X <= 0 or Y <= 0 or Z <= 0 or np.any(vox <= 0):
        rep.errors.append("image voxel counts and sizes must be positive")

    L = geo.n_layers
    rep.lines.append("== collimator ==")
    rep.lines.append(f"layers {L}")
    if L < 0 or L > MAX_LAYERS:
        rep.errors.append(f"layer count {L} outside 0..{MAX_LAYERS}")
        L = 0
    plates = []
    h0 = 0
    col = geo.col.astype(np.float64)
    for m in range(L):
        rec = col[(m + 1) * 10:(m + 1) * 10 + 8]
        nh = int(np.floor(rec[0]))
        w, t, hgt, off, mt, mpe, mc = rec[1:8]
        y1, y2 = off - t / 2, off + t / 2
        plates.append((y1, y2, w, hgt))
        rep.lines.append(f"layer {m}: plate {w:g} x {t:g} x {hgt:g} mm at y {y1:g} .. {y2:g}, "
                         f"mu {mt:g}/{mpe:g}/{mc:g}, {nh} holes")
        if w <= 0 or t <= 0 or hgt <= 0:
            rep.errors.append(f"layer {m}: plate size must be positive")
        _mu_split(f"layer {m} plate", mt, mpe, mc, rep)
        if nh < 0:
            rep.errors.append(f"layer {m}: negative hole count")
            continue
        end = HOLE_BASE + HOLE_STRIDE * (h0 + nh)
        if end > col.size:
            rep.errors.append(f"layer {m}: hole table runs past the collimator record ({end} > {col.size} floats)")
            break
        H = col[HOLE_BASE + HOLE_STRIDE * h0:end].reshape(nh, HOLE_STRIDE)
        h0 += nh
        if nh == 0:
            continue
        rep.lines.append(f"  hole radius {H[:, 4].min():g} .. {H[:, 4].max():g} (mean {H[:, 4].mean():.4g}), "
                         f"x {H[:, 0].min():g} .. {H[:, 0].max():g}, z {H[:, 3].min():g} .. {H[:, 3].max():g}")
        for i in range(min(max_holes_print, nh)):
            rep.lines.append(f"  hole {i}: x {H[i, 0]:g} y {H[i, 1]:g}..{H[i, 2]:g} z {H[i, 3]:g} R {H[i, 4]:g}")
        hy1, hy2 = np.minimum(H[:, 1], H[:, 2]), np.maximum(H[:, 1], H[:, 2])
        if np.any(hy2 - hy1 < 1e-3):
            rep.warnings.append(f"layer {m}: {int(np.sum(hy2 - hy1 < 1e-3))} hole(s) with y1 == y2 "
                                f"(zero-length cylinders: no hole chord)")
        span = (hy1 <= y1 + 1e-4) & (hy2 >= y2 - 1e-4)
        if not np.all(span):
            rep.warnings.append(f"layer {m}: {int(np.sum(~span))} hole(s) do not span the plate y-range "
                                f"{y1:g} .. {y2:g} (hole faces are absolute y, the plate is offset +- t/2)")
        outside = (np.abs(H[:, 0]) + H[:, 4] > w / 2) | (np.abs(H[:, 3]) + H[:, 4] > hgt / 2)
        if np.any(outside):
            rep.warnings.append(f"layer {m}: {int(outside.sum())} hole(s) reach past the plate edge")
        if np.any(H[:, 4] <= 0):
            rep.errors.append(f"layer {m}: hole radius must be positive")
        for q in range(nh):
            _mu_split(f"layer {m} hole {q}", H[q, 5], H[q, 6], H[q, 7], rep)

    N = geo.n_det
    rep.lines.append("== detector ==")
    rep.lines.append(f"crystals {N}")
    if N <= 0:
        rep.errors.append("detector count must be positive")
        return rep
    need = 1 + DET_STRIDE * N
    have = det_file_floats if det_file_floats is not None else geo.det.size
    if have < need:
        rep.errors.append(f"detector record holds {have} floats, {N} crystals need {need}")
        return rep
    if N > REFERENCE_MAX_CRYSTALS:
        rep.lines.append(f"note: {N} crystals exceed the reference drivers' fixed buffers "
                         f"({REFERENCE_MAX_CRYSTALS}); this package reads the header count")
    C = geo.crystals.astype(np.float64)
    xyz, sz = C[:, 0:3], C[:, 3:6]
    rep.lines.append(f"centres {xyz.min(0).tolist()} .. {xyz.max(0).tolist()}")
    us = np.unique(np.round(sz, 6), axis=0)
    rep.lines.append(f"{len(us)} distinct size(s) (width, thickness, height): "
                     + "; ".join(str(u.tolist()) for u in us[:8]))
    ys, cnt = np.unique(np.round(xyz[:, 1], 2), return_counts=True)
    order = np.argsort(-cnt)[:10]
    rep.lines.append("y planes: " + ", ".join(f"{ys[i]:g} ({cnt[i]})" for i in order))
    if np.any(sz <= 0):
        rep.errors.append(f"{int(np.any(sz <= 0, axis=1).sum())} crystal(s) with non-positive size")
    dead = C[:, 6] <= 0.01
    if np.any(dead):
        rep.warnings.append(f"{int(dead.sum())} crystal(s) with mu_total <= 0.01 (their rows are zero)")
    bad_mu = (C[:, 6] > 0) & (np.abs(C[:, 7] + C[:, 8] - C[:, 6]) > 0.01 * C[:, 6] + 1e-6)
    if np.any(bad_mu):
        rep.warnings.append(f"{int(bad_mu.sum())} crystal(s) with mu_pe + mu_c != mu_total")
    cy1, cy2 = xyz[:, 1] - sz[:, 1] / 2, xyz[:, 1] + sz[:, 1] / 2
    for m, (y1, y2, w, hgt) in enumerate(plates):
        inside = (cy2 > y1) & (cy1 < y2) & (np.abs(xyz[:, 0]) < w / 2) & (np.abs(xyz[:, 2]) < hgt / 2)
        if np.any(inside):
            rep.warnings.append(f"{int(inside.sum())} crystal(s) overlap plate {m} (y {y1:g} .. {y2:g})")
    if plates:
        front = min(p[0] for p in plates)
        before = cy2 <= front
        if np.any(before):
            rep.warnings.append(f"{int(before.sum())} crystal(s) lie in front of the collimator (y < {front:g})")
        if hi[1] > front:
            rep.warnings.append(f"FOV reaches y = {hi[1]:g}, past the collimator front face y = {front:g}")
        else:
            rep.lines.append(f"FOV entirely in front of the collimator (max y {hi[1]:g} <= {front:g})")
    return rep


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--param_collimator", required=True)
    ap.add_argument("--param_detector", required=True)
    ap.add_argument("--param_image", required=True)
    ap.add_argument("--max_holes_print", type=int, default=5)
    a = ap.parse_args(argv)

    def rd(path, minlen):
        v = np.fromfile(path, dtype="<f4")
        out = np.zeros(max(minlen, v.size), np.float32)
        out[:v.size] = v
        return out, v.size

    col, _ = rd(a.param_collimator, 100)
    det, n_det_floats = rd(a.param_detector, 13)
    img, _ = rd(a.param_image, 100)
    geo = Geometry(col, det, img, np.zeros(100, np.float32))
    rep = check(geo, det_file_floats=n_det_floats, max_holes_print=a.max_holes_print)
    print(rep.text())
    return 1 if rep.errors else 0


if __name__ == "__main__":
    sys.exit(main())
