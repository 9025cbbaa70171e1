"""CPU-side checks of the C ABI library and the host logic (no compute calls)."""
import os
import re

import numpy as np
import pytest

from sysmat import _lib, params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "sysmat.h")).read()
    return sorted(set(re.findall(r"\b(sm_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libsysmat_hip.so not built"
    L = _lib.load()
    names = header_functions()
    assert names, "no sm_* declarations found"
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/sysmat.h but not exported"
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"
    # and the C++-linkage reference drop-ins declared in include/sysmat_ref.h:
    # PESysMatGen.h:4-9 and the three scatter.h variants
    for mangled in ("_Z11PESysMatGenPfS_S_S_i", "_Z7scatterPfS_S_S_S_PKcS1_S_i", "_Z7scatterPfS_S_S_S_PKcS_i",
                    "_Z7scatterPfS_S_S_PKcS_i"):
        assert hasattr(L, mangled), f"{mangled} not exported"


def test_version_and_errors_without_gpu():
    L = _lib.load()
    assert L.sm_version() >= 100
    assert L.sm_device_count() >= 0


def test_null_arguments_are_rejected_before_any_device_call():
    import ctypes
    L = _lib.load()
    o = _lib.SmPeOptions()
    assert L.sm_pe_default_options(ctypes.byref(o)) == 0
    buf = np.zeros(256, np.float32)
    p = buf.ctypes.data
    assert L.sm_pe_compute(p, None, p, ctypes.byref(o), p) == _lib.SM_E_ARG
    assert L.sm_scatter_crystal(p, p, p, None, None, 0, 0, 0, p, 0) == _lib.SM_E_ARG
    assert L.sm_scatter_collimator(p, p, p, p, None, 0, 0, 0, 0, None, 0) == _lib.SM_E_ARG
    assert L.sm_scatter_cells(None, p, p, p, 0, 0.5, 1.0, 0, 0, p, 0) == _lib.SM_E_ARG
    assert b"null" in L.sm_last_error()


def test_params_roundtrip(tmp_path):
    g = params.config2(seed=3)
    g.write(str(tmp_path))
    h = params.Geometry.read(str(tmp_path))
    for a, b in ((g.col, h.col), (g.det, h.det), (g.img, h.img), (g.phy, h.phy)):
        assert np.array_equal(a, b)
    assert h.n_det == 16384 and h.n_layers == 2 and h.n_holes == 128
    assert h.shape == (64, 64, 64)


def test_short_param_files_read_zero_filled(tmp_path):
    g = params.config1()
    g.write(str(tmp_path))
    np.asarray(g.img[:12], "<f4").tofile(str(tmp_path / "Params_Image.dat"))
    h = params.Geometry.read(str(tmp_path))
    assert h.img.size == 100 and np.all(h.img[12:] == 0)


def test_config_shapes():
    c1 = params.config1()
    assert c1.n_det == 1024 and c1.shape == (16, 16, 16) and c1.n_holes == 25
    c3 = params.config3()
    assert c3.n_det == 2048 and c3.n_holes == 36
    vs = params.config4_variants(seed=0)
    assert len(vs) == 6 and all(v.n_det == 6912 for v in vs)
    assert 1 <= min(v.n_layers for v in vs) and max(v.n_layers for v in vs) <= 3


def test_dropin_arrays_shorter_than_their_headers_are_rejected():
    """The ctypes drop-ins check that the parameter arrays hold every record their headers
    claim before the library reads them (no GPU call on these paths)."""
    from sysmat import pe, scatter
    g = params.tiny(seed=1)
    dst = np.zeros(1, np.float32)
    with pytest.raises(ValueError, match="holes need"):
        pe.PESysMatGen(g.col[:100], g.det, g.img, dst)
    with pytest.raises(ValueError, match="claims"):
        pe.PESysMatGen(g.col, g.det[:13], g.img, dst)
    bad = g.col.copy()
    bad[0] = 10
    with pytest.raises(ValueError, match="layer count"):
        scatter.scatter(bad, g.det, g.img, g.phy, dst, "a", "b", dst)
    params.check_tables(g.col, g.det)  # the generator's arrays are consistent
    params.check_tables(params.config2().col, params.config2().det)


def test_pesysmatgen_rejects_non_integral_voxel_counts():
    """A count such as 63.9995 floors to 63 in the caller's buffer arithmetic (the reference
    driver and PESysMatGen.cu:396-405) but to 64 with the +0.001 rounding of the tables: the
    drop-in refuses it before any device call instead of writing rows longer than dst."""
    L = _lib.load()
    g = params.tiny(seed=1)
    col, det = np.ascontiguousarray(g.col, np.float32), np.ascontiguousarray(g.det, np.float32)
    for k in (0, 1, 2):
        img = np.zeros(100, np.float32)
        img[:len(g.img)] = g.img
        img[30] = img[2]
        img[k] = np.float32(img[k]) - np.float32(0.0005)
        dst = np.zeros(1, np.float32)
        rc = L.sm_PESysMatGen(col.ctypes.data, det.ctypes.data, img.ctypes.data, dst.ctypes.data, 0)
        assert rc == _lib.SM_E_ARG
        assert b"not integral" in L.sm_last_error()
