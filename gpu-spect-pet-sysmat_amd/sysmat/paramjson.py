"""Annotated geometry JSON -> the four Param_*.dat float32 arrays.

Same JSON schema and the same output files as the reference's converter
(PEGen_RayTracing_CircularHole/json_to_dat.py:166-223): keys starting with
``_comment`` are ignored; Param_Image is 12 floats (voxel counts, voxel sizes,
rotations, angle, FOV shift, FOV-to-collimator distance), Param_Physics 10
floats, Param_Detector 1 + 12 * bins floats, Param_Collimator holds the layer
count at [0], 10-float layer records from ``layer_base`` (default 10) and
9-float hole records (holes of all layers in order) from ``holes_base``
(default 100), zero elsewhere.  Little-endian float32.

Missing keys, non-numeric values, hole counts that do not match the hole
lists and bases that overlap raise ``ValueError`` naming the field.
"""
from __future__ import annotations

import json
import os
from typing import Any, Dict, Tuple

import numpy as np

F32 = np.dtype("<f4")

IMAGE_KEYS = ("num_voxel_x", "num_voxel_y", "num_voxel_z", "voxel_size_x_mm", "voxel_size_y_mm", "voxel_size_z_mm",
              "num_rotation", "angle_per_rotation_rad", "shift_fov_x_mm", "shift_fov_y_mm", "shift_fov_z_mm",
              "fov_to_collimator0_mm")
IMAGE_INT = {0, 1, 2, 6}
PHYSICS_KEYS = ("use_compton", "save_pe_sysmat", "save_compton_sysmat", "save_pecompton_sysmat",
                "use_same_energy_window", "energy_window_lower", "energy_window_upper", "target_pe_energy",
                "calc_crystal_geometry_relationship", "calc_collimator_geometry_relationship")
BIN_KEYS = ("cx", "cy", "cz", "width_mm", "thickness_mm", "height_mm", "mu_total", "mu_pe", "mu_compton",
            "energy_resolution_at_target", "rotation_about_y_rad", "flag")
LAYER_KEYS = ("width_mm", "thickness_mm", "height_mm", "offset_from_layer0_mm", "mu_total", "mu_pe", "mu_compton")
HOLE_KEYS = ("cx", "cy1", "cy2", "cz", "radius", "mu_total", "mu_pe", "mu_compton", "flag")


def _field(d: Dict[str, Any], key: str, where: str, integer: bool = False) -> float:
    if not isinstance(d, dict) or key not in d:
        raise ValueError(f"{where}: missing '{key}'")
    v = d[key]
    try:
        return float(int(v)) if integer else float(v)
    except (TypeError, ValueError):
        raise ValueError(f"{where}.{key}: expected a number, got {v!r}") from None


def image_array(image: Dict[str, Any]) -> np.ndarray:
    return np.array([_field(image, k, "image", i in IMAGE_INT) for i, k in enumerate(IMAGE_KEYS)], F32)


def physics_array(phys: Dict[str, Any]) -> np.ndarray:
    return np.array([_field(phys, k, "physics") for k in PHYSICS_KEYS], F32)


def detector_array(det: Dict[str, Any]) -> np.ndarray:
    bins = det.get("bins") if isinstance(det, dict) else None
    if not isinstance(bins, list) or not bins:
        raise ValueError("detector.bins must be a non-empty list")
    rows = [[_field(b, k, f"detector.bins[{i}]") for k in BIN_KEYS] for i, b in enumerate(bins)]
    return np.concatenate([[float(len(bins))], np.ravel(rows)]).astype(F32)


def collimator_array(col: Dict[str, Any], layer_base: int = 10, holes_base: int = 100) -> np.ndarray:
    layers = col.get("layers") if isinstance(col, dict) else None
    if not isinstance(layers, list) or not layers:
        raise ValueError("collimator.layers must be a non-empty list")
    if layer_base < 0 or holes_base < layer_base:
        raise ValueError("need 0 <= layer_base <= holes_base")
    if layer_base + 10 * len(layers) > holes_base and any(l.get("holes") for l in layers if isinstance(l, dict)):
        raise ValueError(f"{len(layers)} layer records from {layer_base} overlap the hole table at {holes_base}")
    holes = []
    records = []
    for li, layer in enumerate(layers):
        where = f"collimator.layers[{li}]"
        n = int(_field(layer, "hole_count", where, integer=True))
        hl = layer.get("holes")
        if not isinstance(hl, list) or len(hl) != n:
            raise ValueError(f"{where}: hole_count {n} does not match its holes list")
        records.append([float(n)] + [_field(layer, k, where) for k in LAYER_KEYS])
        holes += [[_field(h, k, f"{where}.holes[{hi}]") for k in HOLE_KEYS] for hi, h in enumerate(hl)]
    size = max(1, layer_base + 10 * len(layers), holes_base + 9 * len(holes))
    out = np.zeros(size, F32)
    out[0] = len(layers)
    for li, r in enumerate(records):
        out[layer_base + 10 * li:layer_base + 10 * li + len(r)] = r
    for hi, h in enumerate(holes):
        out[holes_base + 9 * hi:holes_base + 9 * hi + 9] = h
    return out


def param_arrays(doc: Dict[str, Any], layer_base: int = 10, holes_base: int = 100
                 ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """(collimator, detector, image, physics) float32 arrays of an annotated geometry document."""
    return (collimator_array(doc.get("collimator"), layer_base, holes_base), detector_array(doc.get("detector")),
            image_array(doc.get("image")), physics_array(doc.get("physics")))


def write_param_files(doc_or_path, outdir: str, prefix: str = "Param_", layer_base: int = 10,
                      holes_base: int = 100) -> Dict[str, str]:
    """Write <prefix>{Collimator,Detector,Image,Physics}.dat; returns the paths."""
    doc = doc_or_path
    if isinstance(doc_or_path, (str, os.PathLike)):
        with open(doc_or_path) as f:
            doc = json.load(f)
    arrays = param_arrays(doc, layer_base, holes_base)
    os.makedirs(outdir, exist_ok=True)
    paths = {}
    for name, arr in zip(("Collimator", "Detector", "Image", "Physics"), arrays):
        p = os.path.join(outdir, f"{prefix}{name}.dat")
        np.asarray(arr, F32).tofile(p)
        paths[name] = p
    return paths


def geometry_from_json(doc_or_path):
    """A Geometry (arrays zero-padded to the sizes the drivers read) from a JSON document."""
    from .params import COL_MIN, DET_MIN, IMG_LEN, PHY_LEN, Geometry
    doc = doc_or_path
    if isinstance(doc_or_path, (str, os.PathLike)):
        with open(doc_or_path) as f:
            doc = json.load(f)
    col, det, img, phy = param_arrays(doc)

    def pad(a, n):
        out = np.zeros(max(n, a.size), np.float32)
        out[:a.size] = a
        return out
    return Geometry(pad(col, COL_MIN), pad(det, DET_MIN), pad(img, IMG_LEN), pad(phy, PHY_LEN))
