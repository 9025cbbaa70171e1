"""GPU parity of the HIP PE tracer against the CPU oracle.

Bar (written here, from the design's precision plan): every element at or
above 1e-9 of its row maximum within 1e-5 relative of the FP64 oracle (the
reference algorithm restated in double); elements below that within 1e-14 of
the row maximum absolutely.  Against the FP32 restatement of the reference the
difference stays inside the FP32 algorithm's own rounding envelope (checked
<= 2e-4, measured ~5e-5; the reference's 1000 mm-extended absorption chord and
hole-wall differences carry that much FP32 noise, see test_oracle_pe.py).
"""
import numpy as np
import pytest

import oracle
from sysmat import params, pe
from sysmat.pe import PESysMatGen, pe_compute

pytestmark = pytest.mark.gpu

REL = 1e-5


def assert_parity(gpu, ref, rel=REL, what=""):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    assert gpu.shape == ref.shape
    rowmax = np.maximum(ref.max(axis=1, keepdims=True), 1e-300)
    sig = ref >= rowmax * 1e-9
    err = np.abs(gpu - ref)
    rel_err = np.where(sig, err / np.maximum(ref, 1e-300), 0.0)
    worst = float(rel_err.max()) if rel_err.size else 0.0
    print(f"[{what}] max rel err (significant) = {worst:.3e}, significant = {sig.mean():.3f}", flush=True)
    assert worst <= rel, f"{what}: max relative error {worst:.3e} > {rel:.0e}"
    small = ~sig
    if small.any():
        assert np.all(err[small] <= rowmax.repeat(ref.shape[1], 1)[small] * 1e-14)


def test_tiny_reference_sampling(lib):
    g = params.tiny(seed=1, n_det_side=4, n_layers=2, image=(4, 4, 4), holes_side=3, n_plates=2)
    gpu = pe_compute(g)
    assert_parity(gpu, oracle.pe(g, precision="f64"), what="tiny f64")
    assert_parity(gpu, oracle.pe(g, precision="mixed"), rel=5e-6, what="tiny mixed")
    f32 = oracle.pe(g, precision="f32")
    assert_parity(gpu, f32, rel=2e-4, what="tiny f32")


def test_tiny_config_sampling(lib):
    g = params.tiny(seed=2, n_det_side=3, n_layers=2, image=(3, 4, 5), holes_side=2, n_plates=1)
    gpu = pe_compute(g, voxel_sub=3, sub=(2, 2, 2))
    assert_parity(gpu, oracle.pe(g, nv=3, sub=(2, 2, 2), precision="f64"), what="tiny cfg f64")


def test_config1_centre_rows(lib):
    g = params.config1()
    rows = (520, 528)
    gpu = pe_compute(g, det_range=rows)
    assert_parity(gpu, oracle.pe(g, det_range=rows, precision="f64"), what="c1 rows f64")


def test_config1_config_sampling(lib):
    g = params.config1()
    rows = (16 * 32 + 14, 16 * 32 + 18)
    gpu = pe_compute(g, voxel_sub=4, sub=(2, 2, 2), det_range=rows)
    assert_parity(gpu, oracle.pe(g, nv=4, sub=(2, 2, 2), det_range=rows, precision="f64"), what="c1 cfg f64")


@pytest.mark.parametrize("row,z,nv", [(2080, 20, 1), (12288 + 2080, 37, 1), (12288 + 2081, 31, 3)])
def test_config2_row_planes_against_fp64_oracle(lib, row, z, nv):
    """16384 crystals in 4 layers, two plates with 8 x 8 holes: one front- or back-layer row on
    one z plane of the 64^3 image (every filter level active on the real geometry) against the
    FP64 oracle; significance is relative to the plane's own maximum."""
    g = params.config2(seed=0)
    gpu = pe_compute(g, voxel_sub=nv, sub=(2, 2, 2), det_range=(row, row + 1), z_range=(z, z + 1))
    ref = oracle.pe(g, nv=nv, sub=(2, 2, 2), det_range=(row, row + 1), precision="f64", z0=z, slab_z=1)
    assert gpu.max() > 0
    assert_parity(gpu, ref, what=f"c2 row {row} z {z} nv {nv}")


def test_rotated_image_and_crystals(lib):
    g = params.tiny(seed=3, n_det_side=3, n_layers=1, image=(4, 3, 2), holes_side=2, rot_angle=0.35)
    g.img[8:11] = (0.7, -0.4, 0.3)
    cr = g.crystals
    cr[:, 10] = np.linspace(-0.3, 0.4, cr.shape[0])  # rotation about y of every crystal
    for rot in (0, 1):
        gpu = pe_compute(g, rotation=rot, sub=(4, 4, 4))
        ref = oracle.pe(g, rot=rot, sub=(4, 4, 4), precision="f64")
        assert_parity(gpu, ref, what=f"rotated r{rot}")


def test_tiles_and_determinism(lib):
    g = params.tiny(seed=4, n_det_side=4, n_layers=2, image=(6, 5, 4), holes_side=3, n_plates=2)
    a = pe_compute(g, tile=(4, 4, 4))
    b = pe_compute(g, tile=(1, 1, 1))
    c = pe_compute(g, tile=(8, 2, 3))
    d = pe_compute(g, tile=(4, 4, 4))
    assert np.array_equal(a, b) and np.array_equal(a, c) and np.array_equal(a, d)


def test_overflowing_tile_groups_retry_bit_identically(lib):
    """16384 crystals (4 layers of 64 x 64): with 16^3-voxel tiles (an 8 mm cube) the
    candidate list of a back-layer detector exceeds the 4096 entries kept per workgroup
    (about 5000 by the sphere bound), so the launch retries with halved tiles; the rows
    must equal the 4^3-tile rows bit for bit."""
    g = params.config2(seed=0)
    rows = (12288 + 2080, 12288 + 2084)
    a = pe_compute(g, sub=(1, 1, 1), det_range=rows, tile=(16, 16, 16))
    b = pe_compute(g, sub=(1, 1, 1), det_range=rows, tile=(4, 4, 4))
    assert np.array_equal(a, b) and a.max() > 0


def test_tile_groups_bit_identical_to_single_tiles(lib):
    """16384 crystals, 4^3 tiles: four rows in one launch (4 x 512 workgroups) run 2 x 2 x 2
    tile groups that share one cull; one row per launch (512 workgroups) runs one tile per
    workgroup.  Every element's candidate list, in index order, is the same either way."""
    torch = pytest.importorskip("torch")
    g = params.config2(seed=0)
    plan = pe.PePlan(g, sub=(1, 1, 1), tile=(4, 4, 4))
    r0, n = 12288 + 2080, 4
    a = torch.empty((n, g.n_vox), dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    plan.launch(r0, n, a)
    for i in range(n):
        plan.launch(r0 + i, 1, b[i:])
    plan.sync()
    assert torch.equal(a, b) and float(a.max()) > 0.0
    plan.close()


def test_tile_group_runs_bit_identical_to_single_tiles(lib):
    """16384 crystals with a 52^3 image: 4^3 tiles give 7 x 7 x 7 = 343 tile groups, so a
    400-row launch (>= 16384 workgroups) runs 8 consecutive groups per workgroup with a
    partial last run of 7; one row per launch (343 workgroups) runs one tile per
    workgroup.  A row whose crystal is below the mu threshold takes the zero path over
    its whole run.  Every element must be the same bit for bit."""
    torch = pytest.importorskip("torch")
    g = params.config5()
    g.img[0:3] = 52
    g.img[3:6] = 0.8
    r0, n = 8000, 400
    g.crystals[r0 + 17, 6] = 0.005
    plan = pe.PePlan(g, sub=(2, 2, 2), tile=(4, 4, 4))
    a = torch.empty((n, g.n_vox), dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(-1.0)
    plan.launch(r0, n, a)
    for i in range(n):
        plan.launch(r0 + i, 1, b[i:])
    plan.sync()
    assert torch.equal(a, b) and float(a.max()) > 0.0
    assert float(a[17].abs().max()) == 0.0 and float(a.min()) >= 0.0
    plan.close()


def test_bundle_filter_changes_nothing(lib):
    g = params.tiny(seed=5, n_det_side=4, n_layers=2, image=(4, 4, 4), holes_side=3, n_plates=2)
    a = pe_compute(g, bundle_filter=True)
    b = pe_compute(g, bundle_filter=False)
    assert_parity(a, b, rel=5e-6, what="filter on/off")


def test_bundle_filters_change_nothing_on_16384_crystals(lib):
    """The bundle bounds (per detector, tile group, tile and element) only drop crystals and
    holes with a zero chord on every sub-ray; back-layer rows of the 16384-crystal geometry,
    three crystal layers in front, with and without them.  Elements whose every candidate hole
    drops out take the FP32 plate path instead of the FP64 one, hence the tolerance."""
    g = params.config2(seed=0)
    rows = (12288 + 2080, 12288 + 2084)
    a = pe_compute(g, sub=(2, 2, 2), det_range=rows, tile=(4, 4, 4), bundle_filter=True)
    b = pe_compute(g, sub=(2, 2, 2), det_range=rows, tile=(4, 4, 4), bundle_filter=False)
    assert a.max() > 0
    assert_parity(a, b, rel=5e-6, what="c2 filter on/off")


def test_dropin_slabs_match_full_volume(lib):
    """PESysMatGen called per Z-slab with the in-band Img[2]/[30]/[31] fields
    (PEGen_CircularHole.cpp:165-216) reproduces the whole-volume result."""
    g = params.tiny(seed=6, n_det_side=3, n_layers=1, image=(4, 4, 6), holes_side=2)
    full = pe_compute(g)
    X, Y, Z = g.shape
    N = g.n_det
    img = np.zeros(100, np.float32)
    img[:len(g.img)] = g.img
    img[30] = Z
    got = np.zeros((N, X * Y * Z), np.float32)
    for z0 in range(0, Z, 4):
        sz = min(4, Z - z0)
        img[2], img[31] = sz, z0
        slab = np.zeros(N * X * Y * sz, np.float32)
        q = PESysMatGen(g.col, g.det, img, slab, 0)
        assert q == X * Y * sz
        got[:, z0 * X * Y:(z0 + sz) * X * Y] = slab.reshape(N, X * Y * sz)
    assert np.array_equal(got, full)


def test_dropin_unset_full_z_places_voxels_like_the_reference(lib):
    """Img[30] = 0 (unset): the reference places plane z0 + iz at (z0 + iz + 0.5) w + shift
    (PESysMatGen.cu:438-444 with Img[30] = 0), and so must the candidate lists."""
    g = params.tiny(seed=6, n_det_side=3, n_layers=1, image=(4, 4, 6), holes_side=2)
    X, Y, Z = g.shape
    N = g.n_det
    img = np.zeros(100, np.float32)
    img[:len(g.img)] = g.img
    z0, sz = 2, 3
    img[2], img[30], img[31] = sz, 0, z0
    slab = np.zeros(N * X * Y * sz, np.float32)
    assert PESysMatGen(g.col, g.det, img, slab, 0) == X * Y * sz
    ref = oracle.pe(g, precision="f64", z0=z0, slab_z=sz, full_z=0)
    assert slab.max() > 0
    assert_parity(slab.reshape(N, -1), ref, what="unset full Z")


@pytest.mark.parametrize("side,layers,image", [(1, 1, (1, 1, 1)), (1, 2, (1, 2, 3))])
def test_smallest_tables_and_images(lib, side, layers, image):
    """One crystal (no other crystal on any path), or one behind another; one voxel or a
    ragged 1 x 2 x 3 image: empty candidate lists and tiles clamped to the image."""
    g = params.tiny(seed=10, n_det_side=side, n_layers=layers, image=image, holes_side=1, n_plates=1)
    gpu = pe_compute(g, tile=(4, 4, 4))
    assert gpu.shape == (g.n_det, g.n_vox) and gpu.min() > 0
    assert_parity(gpu, oracle.pe(g, precision="f64"), what=f"{g.n_det} det {image}")


def test_zero_mu_detector_rows_are_zero(lib):
    g = params.tiny(seed=7, n_det_side=2, n_layers=1, image=(3, 3, 3), holes_side=2)
    g.crystals[1, 6] = 0.005
    gpu = pe_compute(g)
    assert np.all(gpu[1] == 0.0)
    assert np.any(gpu[0] > 0.0)


def test_plan_into_device_tensor(lib):
    torch = pytest.importorskip("torch")
    g = params.tiny(seed=8, n_det_side=4, n_layers=1, image=(4, 4, 4), holes_side=3)
    host = pe_compute(g)
    plan = pe.PePlan(g)
    out = torch.full((g.n_det, g.n_vox), -1.0, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    plan.launch(0, 8, out, stream=stream)
    plan.launch(8, g.n_det - 8, out[8:], stream=stream)
    plan.sync()
    assert np.array_equal(out.cpu().numpy(), host)
    plan.close()


def test_bad_arguments_raise(lib):
    g = params.tiny(seed=9)
    with pytest.raises(pe.SysmatError if hasattr(pe, "SysmatError") else Exception):
        pe_compute(g, det_range=(0, g.n_det + 5))


def test_autotune_picks_a_tile_and_caches(lib, tmp_path):
    """The tuner returns a candidate tile, a second call answers from the cache, and the
    tuned tile gives the same matrix bit for bit (every element's work is tile-independent)."""
    from sysmat import pe as pe_mod
    g = params.tiny(seed=8, n_det_side=4, n_layers=2, image=(8, 8, 8), holes_side=3)
    cache = str(tmp_path / "tiles.txt")
    tile, ns, cached = pe_mod.autotune(g, sample_rows=8, cache_path=cache, sub=(2, 2, 2))
    assert not cached and ns > 0 and all(1 <= t <= 16 for t in tile)
    tile2, ns2, cached2 = pe_mod.autotune(g, sample_rows=8, cache_path=cache, sub=(2, 2, 2))
    assert cached2 and tile2 == tile and abs(ns2 - ns) < 1e-3
    a = pe_mod.pe_compute(g, sub=(2, 2, 2), tile=tile)
    b = pe_mod.pe_compute(g, sub=(2, 2, 2), tile=(4, 4, 4))
    assert np.array_equal(a, b)
