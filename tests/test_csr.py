"""Sparse output stage: keep v > 0 with float64(v) >= rel * float64(row max).

CPU: the numpy CSR oracle against a brute-force loop (ties at the threshold,
zero and negative rows, NaN), and the .csr file format round trip.
GPU: the ballot/mbcnt compaction kernels are bit-exact against the oracle on
the same dense rows (index work: exact equality), on both the float4 and the
scalar paths, padded row strides, multi-chunk rows and empty rows; the PE CSR
entry point and the PE driver's -csr file equal the CSR of the dense PE rows."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from sysmat import csr as csr_mod
from sysmat import params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def brute_csr(d, rel):
    rp, cols, vals = [0], [], []
    for row in d:
        m = 0.0
        for v in row:
            if v == v and float(v) > m:
                m = float(v)
        for j, v in enumerate(row):
            if v > 0 and float(v) >= rel * m:
                cols.append(j)
                vals.append(v)
        rp.append(len(cols))
    return np.array(rp, np.int64), np.array(cols, np.int32), np.array(vals, np.float32)


def tricky_rows():
    rng = np.random.default_rng(5)
    d = rng.random((9, 37), dtype=np.float32) * np.float32(1e-3)
    d[0] = 0.0                                     # all-zero row: nothing kept
    d[1] = -rng.random(37, dtype=np.float32)       # negative row: max clamps to 0, nothing kept
    d[2, :] = 0.0
    d[2, 5] = 1.0
    d[2, 6] = 0.5                                  # exactly at rel*max for rel = 0.5
    d[2, 7] = np.nextafter(np.float32(0.5), np.float32(0))  # one ulp below: dropped
    d[3, 3] = np.nan                               # NaN neither kept nor counted in the max
    d[4, ::2] = 1e-12                              # far below the row max
    return d


@pytest.mark.parametrize("rel", [1e-9, 0.5, 0.0])
def test_oracle_csr_matches_brute_force(rel):
    d = tricky_rows()
    rp, c, v = oracle.csr(d, rel)
    brp, bc, bv = brute_csr(d, rel)
    assert np.array_equal(rp, brp) and np.array_equal(c, bc) and np.array_equal(v, bv)
    if rel == 0.5:
        assert 6 in c[rp[2]:rp[2 + 1]] and 7 not in c[rp[2]:rp[2 + 1]]
    assert rp[1] == 0 and rp[2] == rp[1]


def test_csr_file_roundtrip(tmp_path):
    d = tricky_rows()
    rp, c, v = oracle.csr(d, 1e-9)
    m = csr_mod.CsrMatrix(rp, c, v, d.shape, 1e-9)
    p = str(tmp_path / "x.csr")
    csr_mod.write_csr(p, m)
    for mm in (False, True):
        r = csr_mod.read_csr(p, mmap=mm)
        assert r.shape == d.shape and r.rel == 1e-9
        assert np.array_equal(r.row_ptr, rp) and np.array_equal(r.col, c) and np.array_equal(r.val, v)
    dense = m.to_dense()
    assert np.array_equal(dense, np.where((d > 0) & np.isfinite(d), d, 0) * (dense != 0))
    assert np.array_equal(m.to_scipy().toarray(), dense)
    with open(p, "r+b") as f:
        f.write(b"XXXX")
    with pytest.raises(ValueError):
        csr_mod.read_csr(p)


def _dense_cases():
    rng = np.random.default_rng(11)
    # (rows, cols, column offset into a wider buffer): offset 1 / odd cols force the scalar path
    yield "vec", rng.random((17, 4096), dtype=np.float32) ** 8, 0
    yield "odd-cols", rng.random((5, 1027), dtype=np.float32) ** 8, 0
    yield "offset", rng.random((6, 2053), dtype=np.float32) ** 8, 1
    a = np.zeros((3, 1 << 18), np.float32)
    a[1, ::997] = rng.random(a[1, ::997].size, dtype=np.float32)
    a[2, -1] = 1.0
    yield "multichunk", a, 0
    yield "tricky", tricky_rows(), 0


@pytest.mark.gpu
def test_csr_device_bit_exact(lib):
    import torch
    for name, d, off in _dense_cases():
        for rel in (1e-9, 1e-3, 0.5):
            buf = torch.zeros((d.shape[0], d.shape[1] + 8), dtype=torch.float32, device="cuda")
            buf[:, off:off + d.shape[1]] = torch.from_numpy(d).cuda()
            view = buf[:, off:off + d.shape[1]]
            print(f"csr case {name} {d.shape} rel {rel}", flush=True)
            rp, c, v, rmax = csr_mod.csr_from_dense_device(view, rel)
            torch.cuda.synchronize()
            orp, oc, ov = oracle.csr(d, rel)
            assert np.array_equal(rp.cpu().numpy(), orp), name
            assert np.array_equal(c.cpu().numpy(), oc), name
            assert np.array_equal(v.cpu().numpy(), ov), name
            want_max = np.maximum(np.nan_to_num(d, nan=0.0).max(axis=1), 0)
            assert np.array_equal(rmax.cpu().numpy(), want_max), name


@pytest.mark.gpu
def test_pe_csr_equals_csr_of_dense(lib, tmp_path):
    from sysmat import pe
    g = params.tiny(seed=3, n_det_side=4, n_layers=2, image=(5, 4, 3), holes_side=3)
    dense = pe.pe_compute(g, sub=(2, 2, 2))
    ref = oracle.pe(g, sub=(2, 2, 2), precision="f64")
    for rel in (1e-9, 1e-2):
        m = csr_mod.pe_compute_csr(g, rel=rel, sub=(2, 2, 2), path=str(tmp_path / "pe.csr"))
        orp, oc, ov = oracle.csr(dense, rel)
        assert np.array_equal(m.row_ptr, orp) and np.array_equal(m.col, oc) and np.array_equal(m.val, ov)
        r = csr_mod.read_csr(str(tmp_path / "pe.csr"))
        assert np.array_equal(r.row_ptr, orp) and np.array_equal(r.col, oc) and np.array_equal(r.val, ov)
        # kept values carry the PE parity bar against the FP64 oracle
        rows = np.repeat(np.arange(g.n_det), np.diff(m.row_ptr))
        exact = ref[rows, m.col]
        assert np.max(np.abs(m.val - exact) / exact) <= 1e-5
    # a detector sub-range
    m = csr_mod.pe_compute_csr(g, rel=1e-9, sub=(2, 2, 2), det_range=(3, 11))
    orp, oc, ov = oracle.csr(dense[3:11], 1e-9)
    assert np.array_equal(m.row_ptr, orp) and np.array_equal(m.col, oc) and np.array_equal(m.val, ov)


@pytest.mark.gpu
def test_pe_driver_csr_file(lib, tmp_path):
    from sysmat import pe
    g = params.tiny(seed=4, n_det_side=3, n_layers=2, image=(4, 4, 4), holes_side=3)
    g.img[6] = 2
    g.img[7] = 0.5
    g.write(str(tmp_path))
    exe = os.path.join(ROOT, "gpu-spect-pet-sysmat_amd", "bin", "PEGen_CircularHole")
    print("running PEGen_CircularHole -csr", flush=True)
    r = subprocess.run([exe, "-sub", "1", "2", "2", "2", "-csr", "1e-9"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    x, y, z = (float(np.float32(v)) for v in g.img[8:11])
    m = csr_mod.read_csr(str(tmp_path / f"PE_SysMat_shift_{x:f}_{y:f}_{z:f}.csr"))
    dense = np.concatenate([pe.pe_compute(g, sub=(2, 2, 2), rotation=k) for k in range(2)])
    orp, oc, ov = oracle.csr(dense, 1e-9)
    assert m.shape == dense.shape
    assert np.array_equal(m.row_ptr, orp) and np.array_equal(m.col, oc) and np.array_equal(m.val, ov)
