"""The reference-signature C++ drop-ins (include/sysmat_ref.h), called by a C++
harness the way the reference drivers call them (tests/refabi/harness.cpp:
PESysMatGen per rotation and Z-slab of 4 as in PEGen_CircularHole.cpp:139-216;
the three scatter() overloads of scatter.h:4-5 (x2) and scatter.h:4 once each).

The harness output must equal, bit for bit, the C ABI entry points on the same
parameter files (sm_pe_compute for whole rotations, sm_scatter /
sm_scatter_collimator / sm_scatter_crystal for the scatter terms); the bitmap
files it makes the library write must equal the oracle's bitmaps; and the
PE matrix stays within 1e-5 of the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from sysmat import params, pe as pe_mod, scatter as sc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "refabi", "_build", "harness")


def geometry():
    g = params.tiny(seed=5, n_det_side=4, n_layers=2, image=(4, 4, 6), holes_side=3, n_plates=1)
    g.img[6] = 2  # two rotations
    g.img[7] = 0.3
    return g


def test_cpp_dropins_match_c_abi(lib, tmp_path):
    assert os.path.exists(HARNESS), "harness not built (__graft_entry__.build())"
    g = geometry()
    g.write(str(tmp_path))
    print(f"running {HARNESS} {tmp_path}", flush=True)
    r = subprocess.run([HARNESS, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PESysMatGen rotation") == 4  # 2 rotations x slabs z0 = 0, 4
    N, V = g.n_det, g.n_vox
    pe = np.fromfile(tmp_path / "harness_PE.sysmat", np.float32).reshape(2, N, V)
    for rot in range(2):
        want = pe_mod.pe_compute(g, rotation=rot)
        assert np.array_equal(pe[rot], want), f"PESysMatGen slabs vs sm_pe_compute, rotation {rot}"
        ref = oracle.pe(g, rot=rot, precision="f64").astype(np.float64)
        sig = ref >= ref.max(axis=1, keepdims=True) * 1e-9
        assert pe[rot].max() > 0
        assert (np.abs(pe[rot][sig] - ref[sig]) / ref[sig]).max() <= 1e-5
    pe0 = pe[0]
    col = sc.scatter_collimator(g, rotation=0, variant=0)
    col1 = sc.scatter_collimator(g, rotation=0, variant=1)
    cry = sc.scatter_crystal(g, pe0, rotation=0)
    assert cry.max() > 0 and col.max() > 0
    for tag in ("", "_reread"):
        comb = np.fromfile(tmp_path / f"harness_scatter_combined{tag}.bin", np.float32).reshape(N, V)
        assert np.array_equal(comb, col + cry), f"scatter(combined){tag}"
        hcol = np.fromfile(tmp_path / f"harness_scatter_collimator{tag}.bin", np.float32).reshape(N, V)
        assert np.array_equal(hcol, col1), f"scatter(collimator){tag}"
        hcry = np.fromfile(tmp_path / f"harness_scatter_crystal{tag}.bin", np.float32).reshape(N, V)
        assert np.array_equal(hcry, cry), f"scatter(crystal){tag}"
    assert np.array_equal(np.fromfile(tmp_path / "harness_GeoCrystal", np.uint32), oracle.bitmap_crystal(g))
    assert np.array_equal(np.fromfile(tmp_path / "harness_GeoCollimator", np.uint32), oracle.bitmap_collimator(g))
