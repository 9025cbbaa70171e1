"""The command-line drivers keep the reference's command lines, file names and
.sysmat layout (PEGen_CircularHole.cpp:37-237, ScatterGen_CircularHole.cpp:38-226,
ScatterGen_Collimator.cpp:34-219, ScatterGen_Crystal.cpp:34-215).

CPU tests: usage text and argument errors (no GPU call is made on those paths).
GPU tests: the files the drivers write equal, bit for bit, what the library
entry points return for the same parameter files, and the PE file matches the
FP64 oracle within 1e-5."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from sysmat import params, pe as pe_mod, scatter as sc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gpu-spect-pet-sysmat_amd", "bin")
DRIVERS = ("PEGen_CircularHole", "ScatterGen_CircularHole", "ScatterGen_Collimator", "ScatterGen_Crystal")


def run(name, args, cwd, timeout=300):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} not built (make -C gpu-spect-pet-sysmat_amd)"
    print(f"running {name} {' '.join(args)} in {cwd}", flush=True)
    return subprocess.run([exe] + list(args), cwd=cwd, capture_output=True, text=True, timeout=timeout)


def shift_name(prefix, g):
    x, y, z = (float(np.float32(v)) for v in g.img[8:11])
    return f"{prefix}_shift_{x:f}_{y:f}_{z:f}.sysmat"


@pytest.mark.parametrize("name", DRIVERS)
def test_usage_and_unknown_argument(name, tmp_path):
    params.tiny(seed=1).write(str(tmp_path))
    r = run(name, ["-h"], tmp_path)
    assert r.returncode == 0 and "Usage:" in r.stdout
    r = run(name, ["-bogus"], tmp_path)
    assert r.returncode != 0 and "Unknown parameter" in r.stderr


@pytest.mark.parametrize("name", DRIVERS)
def test_missing_params_file_is_an_error(name, tmp_path):
    r = run(name, [], tmp_path)
    assert r.returncode != 0 and "Params_" in r.stderr


@pytest.mark.parametrize("name", DRIVERS)
def test_inconsistent_collimator_headers_are_errors(name, tmp_path):
    """Layer headers claiming more holes than Params_Collimator.dat holds (beyond the
    80000-float padding), or more than 9 layers, stop the driver before any read past the data."""
    g = params.tiny(seed=1)
    g.write(str(tmp_path))
    col = np.fromfile(tmp_path / "Params_Collimator.dat", np.float32)
    bad = col.copy()
    bad[10] = 10000  # 100 + 9 * 10000 floats needed
    bad.tofile(tmp_path / "Params_Collimator.dat")
    r = run(name, [], tmp_path)
    assert r.returncode != 0 and "fewer holes" in r.stderr
    bad = col.copy()
    bad[0] = 12
    bad.tofile(tmp_path / "Params_Collimator.dat")
    r = run(name, [], tmp_path)
    assert r.returncode != 0 and "layer count" in r.stderr


def small_geometry():
    g = params.tiny(seed=7, n_det_side=4, n_layers=2, image=(4, 4, 4), holes_side=3)
    g.img[6] = 2  # two rotations
    g.img[7] = 0.25
    g.img[8:11] = (0.5, -0.25, 0.0)
    g.phy[2] = g.phy[3] = 1
    g.phy[8] = g.phy[9] = 1
    return g


@pytest.mark.gpu
def test_pe_driver_file(lib, tmp_path):
    g = small_geometry()
    g.write(str(tmp_path))
    r = run("PEGen_CircularHole", ["-cuda", "0", "-sub", "1", "4", "4", "4"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(tmp_path / shift_name("PE_SysMat", g), np.float32)
    N, V = g.n_det, g.n_vox
    assert out.size == 2 * N * V
    for rot in range(2):
        lib_rows = pe_mod.pe_compute(g, sub=(4, 4, 4), rotation=rot)
        assert np.array_equal(out[rot * N * V:(rot + 1) * N * V].reshape(N, V), lib_rows), f"rotation {rot}"
        ref = oracle.pe(g, sub=(4, 4, 4), rot=rot, precision="f64")
        sig = ref > 0
        assert np.all(out[rot * N * V:(rot + 1) * N * V].reshape(N, V)[~sig] == 0)
        err = np.abs(lib_rows[sig] - ref[sig]) / ref[sig]
        assert err.max() <= 1e-5, f"rotation {rot}: {err.max():.3e}"


@pytest.mark.gpu
def test_scatter_drivers_files(lib, tmp_path):
    g = small_geometry()
    g.write(str(tmp_path))
    N, V = g.n_det, g.n_vox
    pe = np.stack([oracle.pe(g, sub=(2, 2, 2), rot=r, precision="f64") for r in range(2)]).astype(np.float32)
    pe.tofile(tmp_path / "pe.sysmat")
    r = run("ScatterGen_CircularHole", ["-PE", "pe.sysmat", "-cuda", "0"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(np.fromfile(tmp_path / "GeometryRelationShip_Crystal2Crystal", np.uint32),
                          oracle.bitmap_crystal(g))
    assert np.array_equal(np.fromfile(tmp_path / "GeometryRelationShip_Collimator2Crystal", np.uint32),
                          oracle.bitmap_collimator(g))
    scat = np.fromfile(tmp_path / shift_name("Scatter_SysMat", g), np.float32).reshape(2, N, V)
    full = np.fromfile(tmp_path / shift_name("SysMat_withScatter", g), np.float32).reshape(2, N, V)
    col0 = [sc.scatter_collimator(g, rotation=rot) for rot in range(2)]
    for rot in range(2):
        want = col0[rot] + sc.scatter_crystal(g, pe[rot], rotation=rot)
        assert np.array_equal(scat[rot], want), f"combined rotation {rot}"
        assert np.array_equal(full[rot], pe[rot] + scat[rot])
    # -pe-rotation0: every rotation reads the rotation-0 PE block, as the reference drivers do
    r = run("ScatterGen_CircularHole", ["-PE", "pe.sysmat", "-pe-rotation0"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    scat0 = np.fromfile(tmp_path / shift_name("Scatter_SysMat", g), np.float32).reshape(2, N, V)
    assert np.array_equal(scat0[0], scat[0])
    assert np.array_equal(scat0[1], col0[1] + sc.scatter_crystal(g, pe[0], rotation=1))
    # the bitmap files written above are read back when named on the command line
    r = run("ScatterGen_Collimator", ["-PE", "pe.sysmat", "-GeoCollimator", "GeometryRelationShip_Collimator2Crystal"],
            tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    col = np.fromfile(tmp_path / shift_name("Scatter_SysMat", g), np.float32).reshape(2, N, V)
    for rot in range(2):
        assert np.array_equal(col[rot], sc.scatter_collimator(g, rotation=rot, variant=1)), f"collimator {rot}"
    r = run("ScatterGen_Crystal", ["-PE", "pe.sysmat", "-GeoCrystal", "GeometryRelationShip_Crystal2Crystal"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    cry = np.fromfile(tmp_path / shift_name("Scatter_Crystal_SysMat", g), np.float32).reshape(2, N, V)
    for rot in range(2):
        assert np.array_equal(cry[rot], sc.scatter_crystal(g, pe[rot], rotation=rot)), f"crystal {rot}"
    tot = np.fromfile(tmp_path / shift_name("SysMat", g), np.float32).reshape(2, N, V)
    assert np.array_equal(tot, pe + cry)


@pytest.mark.gpu
def test_scatter_drivers_cells(lib, tmp_path):
    """-cells SIZE KEEP: the collimator term comes from the plate-cell model."""
    g = small_geometry()
    g.write(str(tmp_path))
    N, V = g.n_det, g.n_vox
    pe = np.stack([oracle.pe(g, sub=(2, 2, 2), rot=r, precision="f64") for r in range(2)]).astype(np.float32)
    pe.tofile(tmp_path / "pe.sysmat")
    r = run("ScatterGen_Collimator", ["-PE", "pe.sysmat", "-cells", "1.0", "1.5"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    col = np.fromfile(tmp_path / shift_name("Scatter_SysMat", g), np.float32).reshape(2, N, V)
    cells = [sc.scatter_cells(g, 1.0, 1.5, rotation=rot) for rot in range(2)]
    for rot in range(2):
        assert cells[rot].max() > 0
        assert np.array_equal(col[rot], cells[rot]), f"cells rotation {rot}"
    r = run("ScatterGen_CircularHole", ["-PE", "pe.sysmat", "-cells", "1.0", "1.5"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    scat = np.fromfile(tmp_path / shift_name("Scatter_SysMat", g), np.float32).reshape(2, N, V)
    for rot in range(2):
        assert np.array_equal(scat[rot], cells[rot] + sc.scatter_crystal(g, pe[rot], rotation=rot)), f"rotation {rot}"
    r = run("ScatterGen_Collimator", ["-PE", "pe.sysmat", "-cells", "0", "1"], tmp_path)
    assert r.returncode != 0
