"""GPU parity of the Compton-scatter kernels against the CPU oracle.

Bitmaps (GeometryRelationShip_*): bit-exact against the FP32 restatement of the
reference builders (scatter.cu:1440-1540, 1939-2009).

Scatter matrices: every element at or above 1e-9 of its row maximum within
1e-5 relative of the FP64 oracle (crystalScatterSysMatCuda /
collimatorScatterSysMatCuda restated in double, same bitmaps and PE input);
rows that are all zero in the oracle are all zero on the GPU.
"""
import os

import numpy as np
import pytest

import oracle
from sysmat import params
from sysmat import scatter as sc

pytestmark = pytest.mark.gpu

REL = 1e-5


def assert_parity(gpu, ref, rel=REL, what=""):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    assert gpu.shape == ref.shape
    rowmax = ref.max(axis=1, keepdims=True)
    zero_rows = rowmax[:, 0] <= 0
    assert np.all(gpu[zero_rows] == 0), f"{what}: rows zero in the oracle are not zero on the GPU"
    sig = (ref >= rowmax * 1e-9) & (rowmax > 0)
    rel_err = np.where(sig, np.abs(gpu - ref) / np.maximum(ref, 1e-300), 0.0)
    worst = float(rel_err.max()) if rel_err.size else 0.0
    print(f"[{what}] max rel err = {worst:.3e}, significant = {sig.mean():.3f}, "
          f"total sum rel = {abs(gpu.sum() - ref.sum()) / max(ref.sum(), 1e-300):.2e}", flush=True)
    assert worst <= rel, f"{what}: max relative error {worst:.3e} > {rel:.0e}"
    small = ~sig & ~zero_rows[:, None]
    if small.any():
        assert np.all(np.abs(gpu - ref)[small] <= (rowmax * 1e-14).repeat(ref.shape[1], 1)[small])


def geo_small(seed=1, side=4, layers=2, image=(4, 4, 4), holes=3, plates=1):
    return params.tiny(seed=seed, n_det_side=side, n_layers=layers, image=image, holes_side=holes, n_plates=plates)


def test_bitmaps_bit_exact(lib):
    for seed, side, layers in ((1, 4, 2), (2, 5, 3), (3, 6, 2)):
        g = geo_small(seed, side, layers, holes=3)
        assert np.array_equal(sc.bitmap_crystal(g), oracle.bitmap_crystal(g)), f"crystal bitmap seed {seed}"
        assert np.array_equal(sc.bitmap_collimator(g), oracle.bitmap_collimator(g)), f"collimator bitmap seed {seed}"


def test_crystal_scatter_parity(lib):
    for seed in (1, 2):
        g = geo_small(seed)
        pe = oracle.pe(g, precision="f64")
        bits = oracle.bitmap_crystal(g)
        gpu = sc.scatter_crystal(g, pe, bits)
        assert_parity(gpu, oracle.scatter_crystal(g, pe, bits, precision="f64"), what=f"crystal seed {seed}")
        f32 = oracle.scatter_crystal(g, pe, bits, precision="f32")
        assert_parity(gpu, f32, rel=1e-4, what=f"crystal vs f32 seed {seed}")


@pytest.mark.parametrize("variant", [0, 1])
def test_collimator_scatter_parity(lib, variant):
    for seed in (1, 2):
        g = geo_small(seed)
        bits = oracle.bitmap_collimator(g)
        gpu = sc.scatter_collimator(g, bits, variant=variant)
        ref = oracle.scatter_collimator(g, bits, precision="f64", variant=variant)
        assert_parity(gpu, ref, what=f"collimator v{variant} seed {seed}")


@pytest.mark.parametrize("side,layers,image", [(1, 1, (1, 1, 1)), (1, 2, (1, 2, 3))])
def test_smallest_tables_and_images(lib, side, layers, image):
    """One crystal (no scatter site but itself: crystal scatter all zero) or two; one voxel or a
    ragged 1 x 2 x 3 image; one hole."""
    g = geo_small(10, side, layers, image=image, holes=1)
    pe = oracle.pe(g, precision="f64")
    bits, cbits = oracle.bitmap_crystal(g), oracle.bitmap_collimator(g)
    assert np.array_equal(sc.bitmap_crystal(g), bits) and np.array_equal(sc.bitmap_collimator(g), cbits)
    gpu = sc.scatter_crystal(g, pe, bits)
    assert_parity(gpu, oracle.scatter_crystal(g, pe, bits, precision="f64"), what=f"crystal {g.n_det} det {image}")
    gpu = sc.scatter_collimator(g, cbits)
    ref = oracle.scatter_collimator(g, cbits, precision="f64")
    assert ref.min() > 0
    assert_parity(gpu, ref, what=f"collimator {g.n_det} det {image}")


def test_rotation_and_window_variants(lib):
    g = geo_small(4, side=3, layers=2, image=(3, 4, 3))
    g.img[7] = 0.3
    g.phy[4] = 0  # per-crystal window (1 +- res/2) E0
    g.crystals[:, 10] = np.linspace(-0.2, 0.3, g.n_det)  # rotated crystals
    pe = oracle.pe(g, rot=1, precision="f64", sub=(2, 2, 2))
    cb, hb = oracle.bitmap_crystal(g), oracle.bitmap_collimator(g)
    assert_parity(sc.scatter_crystal(g, pe, cb, rotation=1),
                  oracle.scatter_crystal(g, pe, cb, precision="f64", rot=1), what="crystal rot")
    assert_parity(sc.scatter_collimator(g, hb, rotation=1),
                  oracle.scatter_collimator(g, hb, precision="f64", rot=1), what="collimator rot")


@pytest.mark.parametrize("energy", [140.5, 364.0])
def test_photon_energy_both_kn_logarithms(lib, energy):
    """The scatter kernel takes the KN logarithm by its series up to 236 keV and from the
    library above (two kernel instances): 140.5 keV (Tc-99m) and 364 keV (I-131), per-crystal
    windows (1 +- res/2) E0, both scatter kinds against the FP64 oracle."""
    g = geo_small(6, side=3, layers=2, image=(3, 3, 4))
    g.phy[7] = energy
    g.phy[4] = 0
    pe = oracle.pe(g, precision="f64", sub=(2, 2, 2))
    cb, hb = oracle.bitmap_crystal(g), oracle.bitmap_collimator(g)
    want = oracle.scatter_crystal(g, pe, cb, precision="f64")
    assert want.max() > 0
    assert_parity(sc.scatter_crystal(g, pe, cb), want, what=f"crystal {energy} keV")
    assert_parity(sc.scatter_collimator(g, hb), oracle.scatter_collimator(g, hb, precision="f64"),
                  what=f"collimator {energy} keV")


def test_config3_rows(lib):
    """C3 detector (2 x 32 x 32 crystals, 6x6 holes) on a reduced 8^3 image, GPU bitmaps."""
    g = params.config3()
    g.img[0:3] = 8
    g.img[3:6] = 4.0
    rows = (1024 + 16 * 32 + 14, 1024 + 16 * 32 + 18)  # back layer, centre
    pe = oracle.pe(g, precision="f64", sub=(2, 2, 2))
    cb = sc.bitmap_crystal(g)
    hb = sc.bitmap_collimator(g)
    gpu_c = sc.scatter_crystal(g, pe, cb, det_range=rows)
    assert_parity(gpu_c, oracle.scatter_crystal(g, pe, cb, det_range=rows, precision="f64"), what="c3 crystal")
    gpu_h = sc.scatter_collimator(g, hb, det_range=rows)
    assert_parity(gpu_h, oracle.scatter_collimator(g, hb, det_range=rows, precision="f64"), what="c3 collimator")


def test_dropin_scatter_files(lib, tmp_path):
    """scatter() with Phys[8]/[9] = 1 writes the bitmap files; with 0 it reads them back
    (scatter.cu:2116-2164, 2246-2296); dst = collimator + crystal."""
    g = geo_small(5)
    pe = oracle.pe(g, precision="f64")
    fc, fh = str(tmp_path / "GeometryRelationShip_Crystal2Crystal"), str(tmp_path / "GeometryRelationShip_Collimator2Crystal")
    phy = g.phy.copy()
    phy[8] = phy[9] = 1
    dst = np.zeros(g.n_det * g.n_vox, np.float32)
    assert sc.scatter(g.col, g.det, g.img, phy, pe, fc, fh, dst) == g.n_vox
    assert np.array_equal(np.fromfile(fc, np.uint32), oracle.bitmap_crystal(g))
    assert np.array_equal(np.fromfile(fh, np.uint32), oracle.bitmap_collimator(g))
    phy[8] = phy[9] = 0
    dst2 = np.zeros_like(dst)
    sc.scatter(g.col, g.det, g.img, phy, pe, fc, fh, dst2)
    assert np.array_equal(dst, dst2)
    both = sc.scatter_collimator(g) + sc.scatter_crystal(g, pe)
    assert np.array_equal(dst.reshape(g.n_det, -1), both)


def test_missing_bitmap_file_is_an_error(lib, tmp_path):
    g = geo_small(6)
    pe = np.zeros((g.n_det, g.n_vox), np.float32)
    phy = g.phy.copy()
    phy[8] = phy[9] = 0
    dst = np.zeros(g.n_det * g.n_vox, np.float32)
    with pytest.raises(Exception):
        sc.scatter(g.col, g.det, g.img, phy, pe, str(tmp_path / "nope"), str(tmp_path / "nope2"), dst)


# ---- plate-cell collimator model (no reference kernel; FP64 oracle restatement) ----
def _cells_gpu(g, h, keep, rotation=0, chunk=0, rows=None):
    import torch
    plan = sc.ScatterPlan(g, rotation=rotation, bitmaps=False)
    b, e = rows if rows is not None else (0, g.n_det)
    out = torch.full((e - b, g.n_vox), -1.0, dtype=torch.float32, device="cuda")
    plan.collimator_cells(b, e - b, out, cell_size=h, keep_dist=keep, cells_per_chunk=chunk)
    torch.cuda.synchronize()
    res = out.cpu().numpy()
    plan.close()
    return res


def test_cell_table_matches_oracle(lib):
    for g, h, keep in ((geo_small(1), 1.0, -1.0), (geo_small(2, plates=2), 1.0, 0.75), (params.config3(), 0.5, 1.0)):
        plan = sc.ScatterPlan(g, bitmaps=False)
        got = plan.cell_table(h, keep)
        plan.close()
        ref = oracle.cells(g, h, keep)
        print(f"cells h={h} keep={keep}: {len(ref)}", flush=True)
        assert len(ref) > 0
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("case", [(1, 1, -1.0, 0, 0), (2, 2, 0.75, 0, 0), (3, 1, 1.5, 0, 37), (4, 2, -1.0, 1, 500)])
def test_cells_scatter_parity(lib, case):
    """Plate cells of 1 mm: every material cell or those near hole walls, one or two
    plates, the image rotated, and small cell chunks (accumulate path)."""
    seed, plates, keep, rot, chunk = case
    g = geo_small(seed, plates=plates)
    if rot:
        g.img[7] = 0.4
    print(f"cells parity seed={seed} plates={plates} keep={keep} rot={rot} chunk={chunk}", flush=True)
    gpu = _cells_gpu(g, 1.0, keep, rotation=rot, chunk=chunk)
    ref = oracle.scatter_cells(g, 1.0, keep, precision="f64", rot=rot)
    assert ref.max() > 0
    assert_parity(gpu, ref, what=f"cells seed {seed}")


def test_cells_config3_rows(lib):
    """C3 collimator (64 x 5 x 64 mm plate, 0.5 mm cells within 1 mm of a hole wall)
    on a reduced 8^3 image, two centre detectors of the back layer."""
    g = params.config3()
    g.img[0:3] = 8
    g.img[3:6] = 4.0
    r0 = 1024 + 16 * 32 + 14
    print("cells c3 rows", flush=True)
    gpu = _cells_gpu(g, 0.5, 1.0, rows=(r0, r0 + 2))
    ref = oracle.scatter_cells(g, 0.5, 1.0, det_range=(r0, r0 + 2), precision="f64")
    assert ref.max() > 0
    assert_parity(gpu, ref, what="cells c3")
