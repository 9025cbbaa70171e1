"""Params_*.dat reader/writer and seeded geometry generators.

File formats (raw little-endian float32, no header), as read by the reference
drivers (PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:39-62) and laid
out in PEGen_RayTracing_CircularHole/ReadMe.txt:1-61:

* Collimator: ``[0]`` layer count; layer ``m`` at ``(m+1)*10 + k`` with
  k = 0 holes, 1 width X, 2 thickness Y, 3 height Z, 4 offset from layer 0,
  5 mu_total, 6 mu_pe, 7 mu_compton; hole ``h`` (holes numbered layer by
  layer) at ``100 + 9*h + k`` with k = 0 x, 1 y1, 2 y2, 3 z, 4 R, 5 mu_total,
  6 mu_pe, 7 mu_compton, 8 flag.
* Detector: ``[0]`` crystal count N; crystal ``d`` at ``12*d + k`` with
  k = 1 x, 2 y, 3 z, 4 width, 5 thickness, 6 height, 7 mu_total, 8 mu_pe,
  9 mu_compton, 10 energy resolution, 11 rotation about y, 12 flag.
* Image: 0-2 voxel counts, 3-5 voxel widths, 6 rotations, 7 angle per
  rotation, 8-10 FOV shift, 11 FOV-centre-to-collimator distance.  Slots 2,
  20, 30, 31 are also used in-band by the PE driver (slab Z, rotation index,
  full Z, slab start; PEGen_CircularHole.cpp:165,177,189-190).
* Physics: 0 use Compton, 1-3 save flags, 4 same-window flag, 5/6 energy
  window, 7 photopeak energy (keV), 8/9 compute-vs-load geometry bitmaps.

The reference reads fixed 80000/80000/100/100 floats, which caps N at 6666
crystals.  This reader takes the counts from the headers instead, so larger
detectors work; the writer pads to at least the reference's fixed sizes so
the files stay readable by the reference drivers.
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional, Sequence

import numpy as np

COL_MIN = 80000
DET_MIN = 80000
IMG_LEN = 100
PHY_LEN = 100
HOLE_BASE = 100
HOLE_STRIDE = 9
DET_STRIDE = 12
MAX_LAYERS = 9  # layer records at (m+1)*10 must stay below the hole table at 100


def count(x: float, eps: float = 0.001) -> int:
    """Integer stored as float, read as floor(x + eps) like the reference drivers."""
    return int(np.floor(np.float32(x) + np.float32(eps)))


def check_tables(col: np.ndarray, det: np.ndarray) -> None:
    """Raise ValueError unless the collimator and detector arrays hold every record their
    headers claim (at most MAX_LAYERS layers, 9 floats per hole after the 100-float layer
    block, 12 per crystal): the library reads the records by the headers alone."""
    if len(col) < HOLE_BASE:
        raise ValueError(f"collimator array holds {len(col)} floats, the layer block needs {HOLE_BASE}")
    n_layers = count(col[0], 1e-6)
    if not 0 <= n_layers <= MAX_LAYERS:
        raise ValueError(f"collimator layer count {n_layers} outside 0..{MAX_LAYERS}")
    holes = sum(max(0, int(np.floor(col[(m + 1) * 10]))) for m in range(n_layers))
    if len(col) < HOLE_BASE + HOLE_STRIDE * holes:
        raise ValueError(f"collimator array holds {len(col)} floats, {holes} holes need "
                         f"{HOLE_BASE + HOLE_STRIDE * holes}")
    n = count(det[0], 1e-6) if len(det) else 0
    if n <= 0 or len(det) < 1 + DET_STRIDE * n:
        raise ValueError(f"detector array holds {len(det)} floats, its header claims {n} crystals")


@dataclasses.dataclass
class Geometry:
    """The four reference parameter arrays, as float32 numpy arrays."""

    col: np.ndarray
    det: np.ndarray
    img: np.ndarray
    phy: np.ndarray

    # ---- derived views -------------------------------------------------
    @property
    def n_layers(self) -> int:
        return count(self.col[0])

    @property
    def n_det(self) -> int:
        return count(self.det[0], 0.0001)

    @property
    def layers(self) -> np.ndarray:
        """(L, 8) float32 layer records."""
        L = self.n_layers
        return np.stack([self.col[(m + 1) * 10:(m + 1) * 10 + 8] for m in range(L)]) if L else np.zeros((0, 8), np.float32)

    @property
    def holes_per_layer(self) -> List[int]:
        return [int(np.floor(self.col[(m + 1) * 10])) for m in range(self.n_layers)]

    @property
    def n_holes(self) -> int:
        return int(sum(self.holes_per_layer))

    @property
    def holes(self) -> np.ndarray:
        H = self.n_holes
        return self.col[HOLE_BASE:HOLE_BASE + HOLE_STRIDE * H].reshape(H, HOLE_STRIDE)

    @property
    def crystals(self) -> np.ndarray:
        """(N, 12) float32 crystal records (x, y, z, w, t, h, mut, mupe, muc, res, rot, flag)."""
        N = self.n_det
        return self.det[1:1 + DET_STRIDE * N].reshape(N, DET_STRIDE)

    @property
    def shape(self):
        return tuple(count(self.img[i]) for i in range(3))

    @property
    def n_vox(self) -> int:
        x, y, z = self.shape
        return x * y * z

    @property
    def n_rot(self) -> int:
        return count(self.img[6])

    def shift(self):
        return float(self.img[8]), float(self.img[9]), float(self.img[10])

    def copy(self) -> "Geometry":
        return Geometry(self.col.copy(), self.det.copy(), self.img.copy(), self.phy.copy())

    # ---- file I/O ------------------------------------------------------
    def write(self, directory: str, prefix: str = "Params_") -> None:
        os.makedirs(directory, exist_ok=True)
        for name, arr in (("Collimator", self.col), ("Detector", self.det), ("Image", self.img), ("Physics", self.phy)):
            np.asarray(arr, dtype="<f4").tofile(os.path.join(directory, f"{prefix}{name}.dat"))

    @staticmethod
    def read(directory: str, prefix: str = "Params_") -> "Geometry":
        def rd(name: str, minlen: int) -> np.ndarray:
            path = os.path.join(directory, f"{prefix}{name}.dat")
            a = np.fromfile(path, dtype="<f4")
            out = np.zeros(max(minlen, a.size), np.float32)
            out[:a.size] = a  # short files read zero-filled tails, as fread into new float[]()
            return out

        col, det = rd("Collimator", COL_MIN), rd("Detector", DET_MIN)
        img, phy = rd("Image", IMG_LEN), rd("Physics", PHY_LEN)
        return Geometry(col, det, img, phy)

    # ---- pretty --------------------------------------------------------
    def describe(self) -> str:
        X, Y, Z = self.shape
        return (f"image {X}x{Y}x{Z} vox {self.img[3]:g}x{self.img[4]:g}x{self.img[5]:g} mm, "
                f"{self.n_layers} plate(s), {self.n_holes} holes, {self.n_det} crystals, "
                f"{self.n_rot} rotation(s)")


def build(layers: Sequence[dict], holes: np.ndarray, crystals: np.ndarray, image: dict, physics: dict) -> Geometry:
    """Pack plain records into the four reference arrays.

    ``layers``: dicts with width, thickness, height, offset, mut, mupe, muc
    (the hole count is filled from ``holes``' layer column).
    ``holes``: (H, 10) array x, y1, y2, z, R, mut, mupe, muc, flag, layer.
    ``crystals``: (N, 12) array in Detector record order.
    """
    L = len(layers)
    if L > MAX_LAYERS:
        raise ValueError(f"at most {MAX_LAYERS} collimator layers fit the record layout")
    holes = np.asarray(holes, np.float64).reshape(-1, 10)
    order = np.argsort(holes[:, 9], kind="stable")
    holes = holes[order]
    H = holes.shape[0]
    col = np.zeros(max(COL_MIN, HOLE_BASE + HOLE_STRIDE * H), np.float32)
    col[0] = L
    for m, lay in enumerate(layers):
        base = (m + 1) * 10
        col[base + 0] = int(np.sum(holes[:, 9] == m))
        col[base + 1] = lay["width"]
        col[base + 2] = lay["thickness"]
        col[base + 3] = lay["height"]
        col[base + 4] = lay["offset"]
        col[base + 5] = lay["mut"]
        col[base + 6] = lay["mupe"]
        col[base + 7] = lay["muc"]
    if H:
        col[HOLE_BASE:HOLE_BASE + HOLE_STRIDE * H] = holes[:, :9].astype(np.float32).ravel()
    crystals = np.asarray(crystals, np.float32).reshape(-1, DET_STRIDE)
    N = crystals.shape[0]
    det = np.zeros(max(DET_MIN, 1 + DET_STRIDE * N), np.float32)
    det[0] = N
    det[1:1 + DET_STRIDE * N] = crystals.ravel()
    img = np.zeros(IMG_LEN, np.float32)
    img[0:3] = image["shape"]
    img[3:6] = image["voxel"]
    img[6] = image.get("n_rot", 1)
    img[7] = image.get("angle", 0.0)
    img[8:11] = image.get("shift", (0.0, 0.0, 0.0))
    img[11] = image["fov2col"]
    phy = np.zeros(PHY_LEN, np.float32)
    phy[0] = physics.get("use_compton", 1)
    phy[1] = physics.get("save_pe", 1)
    phy[2] = physics.get("save_compton", 1)
    phy[3] = physics.get("save_total", 1)
    phy[4] = physics.get("same_window", 1)
    phy[5] = physics["window"][0]
    phy[6] = physics["window"][1]
    phy[7] = physics["energy"]
    phy[8] = physics.get("calc_crystal_geo", 1)
    phy[9] = physics.get("calc_collimator_geo", 1)
    return Geometry(col, det, img, phy)


def hole_grid(n: int, pitch: float, diam, y1: float, y2: float, layer: int, mu=(0.0, 0.0, 0.0)) -> np.ndarray:
    """n x n holes centred on the axis; ``diam`` is a scalar or an (n*n,) array."""
    c = (np.arange(n) - (n - 1) / 2.0) * pitch
    xs, zs = np.meshgrid(c, c, indexing="ij")
    d = np.broadcast_to(np.asarray(diam, np.float64), (n * n,))
    out = np.zeros((n * n, 10))
    out[:, 0] = xs.ravel()
    out[:, 1] = y1
    out[:, 2] = y2
    out[:, 3] = zs.ravel()
    out[:, 4] = d / 2.0
    out[:, 5:8] = mu
    out[:, 8] = 1.0
    out[:, 9] = layer
    return out


def crystal_layers(n: int, pitch: float, thickness: float, y_front: float, n_layers: int,
                   mu=(0.35, 0.28, 0.07), res: float = 0.10, width: Optional[float] = None) -> np.ndarray:
    """n x n crystals per layer, layers stacked behind each other starting at ``y_front``."""
    w = pitch if width is None else width
    c = (np.arange(n) - (n - 1) / 2.0) * pitch
    out = []
    for l in range(n_layers):
        yc = y_front + thickness * (l + 0.5)
        for zi in range(n):
            for xi in range(n):
                out.append([c[xi], yc, c[zi], w, thickness, w, mu[0], mu[1], mu[2], res, 0.0, 1.0])
    return np.asarray(out, np.float32)


W_MU = (3.9, 3.78, 0.12)  # tungsten at ~140 keV: total, photoelectric, Compton (1/mm)
PHYS_140 = dict(energy=140.5, window=(126.45, 154.55), same_window=1)


def config1(seed: int = 0) -> Geometry:
    """Plain PE, small: 16^3 x 1 mm image, one 64x5x64 W plate with 5x5 holes of 1 mm
    at 4 mm pitch, 32x32 CdTe 1.6 mm pitch 3 mm thick, 10 mm behind the plate."""
    t = 5.0
    layers = [dict(width=64.0, thickness=t, height=64.0, offset=0.0, mut=W_MU[0], mupe=W_MU[1], muc=W_MU[2])]
    holes = hole_grid(5, 4.0, 1.0, -t / 2, t / 2, 0)
    cry = crystal_layers(32, 1.6, 3.0, t / 2 + 10.0, 1)
    return build(layers, holes, cry, dict(shape=(16, 16, 16), voxel=(1.0, 1.0, 1.0), fov2col=50.0), PHYS_140)


def config2(seed: int = 0) -> Geometry:
    """PE realistic: 64^3 x 0.5 mm image, two 5 mm W plates 20 mm apart (offset 25 mm),
    8x8 holes of U(0.8,1.2) mm at 6 mm pitch, 4 layers of 64x64 crystals 0.8 mm pitch
    2 mm thick (N = 16384), first layer 10 mm behind the second plate."""
    rng = np.random.default_rng(seed)
    t, off = 5.0, 25.0
    layers = [dict(width=64.0, thickness=t, height=64.0, offset=o, mut=W_MU[0], mupe=W_MU[1], muc=W_MU[2])
              for o in (0.0, off)]
    holes = np.concatenate([hole_grid(8, 6.0, rng.uniform(0.8, 1.2, 64), o - t / 2, o + t / 2, m)
                            for m, o in enumerate((0.0, off))])
    cry = crystal_layers(64, 0.8, 2.0, off + t / 2 + 10.0, 4)
    return build(layers, holes, cry, dict(shape=(64, 64, 64), voxel=(0.5, 0.5, 0.5), fov2col=50.0), PHYS_140)


def config3(seed: int = 0) -> Geometry:
    """Collimator Compton scatter: 32^3 x 1 mm image, one 5 mm W plate with 6x6 holes of
    1 mm at 4 mm pitch, 2 layers of 32x32 CdTe 1.6 mm pitch 3 mm thick."""
    t = 5.0
    layers = [dict(width=64.0, thickness=t, height=64.0, offset=0.0, mut=W_MU[0], mupe=W_MU[1], muc=W_MU[2])]
    holes = hole_grid(6, 4.0, 1.0, -t / 2, t / 2, 0)
    cry = crystal_layers(32, 1.6, 3.0, t / 2 + 10.0, 2)
    return build(layers, holes, cry, dict(shape=(32, 32, 32), voxel=(1.0, 1.0, 1.0), fov2col=50.0), PHYS_140)


def config4_variants(seed: int = 0, n_variants: int = 6, image=(48, 48, 48), voxel: float = 0.75,
                     det_side: int = 48, n_layers: int = 3) -> List[Geometry]:
    """Crystal + collimator scatter over collimator variants drawn from ``seed``:
    1-3 plates, hole diameter U(0.5, 2.0), 4x4..12x12 grids; shared 48^3 x 0.75 mm
    image and shared detector, 3 layers of 48x48 crystals (1.2 mm pitch, 2 mm thick,
    mu_c 0.07), placed 10 mm behind the deepest (3-plate) collimator so every
    variant sees the same detector."""
    rng = np.random.default_rng(seed)
    out = []
    t = 4.0
    gap = 12.0
    max_plates = 3
    for _ in range(n_variants):
        n_pl = int(rng.integers(1, max_plates + 1))
        offs = [i * gap for i in range(n_pl)]
        layers = [dict(width=72.0, thickness=t, height=72.0, offset=o, mut=W_MU[0], mupe=W_MU[1], muc=W_MU[2])
                  for o in offs]
        g = int(rng.integers(4, 13))
        d = float(rng.uniform(0.5, 2.0))
        pitch = max(2.0 * d, 60.0 / g)
        holes = np.concatenate([hole_grid(g, pitch, d, o - t / 2, o + t / 2, m) for m, o in enumerate(offs)])
        cry = crystal_layers(det_side, 1.2, 2.0, (max_plates - 1) * gap + t / 2 + 10.0, n_layers)
        out.append(build(layers, holes, cry, dict(shape=tuple(image), voxel=(voxel,) * 3, fov2col=50.0),
                         PHYS_140))
    return out


def config5(seed: int = 0) -> Geometry:
    """Surrogate target: 128^3 x 0.4 mm image, C3-like collimator, 4 layers of 64x64 crystals."""
    t = 5.0
    layers = [dict(width=64.0, thickness=t, height=64.0, offset=0.0, mut=W_MU[0], mupe=W_MU[1], muc=W_MU[2])]
    holes = hole_grid(6, 4.0, 1.0, -t / 2, t / 2, 0)
    cry = crystal_layers(64, 0.8, 2.0, t / 2 + 10.0, 4)
    return build(layers, holes, cry, dict(shape=(128, 128, 128), voxel=(0.4, 0.4, 0.4), fov2col=50.0), PHYS_140)


CONFIGS = {"c1": config1, "c2": config2, "c3": config3, "c5": config5}


def tiny(seed: int = 0, n_det_side: int = 4, n_layers: int = 1, image=(4, 4, 4), holes_side: int = 3,
         n_plates: int = 1, rot_angle: float = 0.0) -> Geometry:
    """A seeded small geometry for exhaustive parity tests (a few thousand elements)."""
    rng = np.random.default_rng(seed)
    t = 3.0
    offs = [i * 8.0 for i in range(n_plates)]
    layers = [dict(width=30.0, thickness=t, height=30.0, offset=o, mut=3.9, mupe=3.78, muc=0.12) for o in offs]
    hl = []
    for m, o in enumerate(offs):
        h = hole_grid(holes_side, 3.0, rng.uniform(0.8, 1.4, holes_side * holes_side), o - t / 2, o + t / 2, m)
        h[:, 0] += rng.uniform(-0.3, 0.3, h.shape[0])
        h[:, 3] += rng.uniform(-0.3, 0.3, h.shape[0])
        hl.append(h)
    holes = np.concatenate(hl)
    cry = crystal_layers(n_det_side, 1.6, 3.0, offs[-1] + t / 2 + 6.0, n_layers)
    return build(layers, holes, cry, dict(shape=image, voxel=(1.0, 1.0, 1.0), fov2col=25.0, angle=rot_angle),
                 PHYS_140)
