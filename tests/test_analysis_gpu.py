"""GPU beam masks / MPXI of device CSR rows (sm_csr_beams_device) against the host
restatement (exact integer equality), on random CSR rows with threshold ties and on
the CSR of a PE matrix produced by the tracer."""
import numpy as np
import pytest

from sysmat import analysis, csr, params
from sysmat import pe as pe_mod

pytestmark = pytest.mark.gpu


def to_dev(m):
    import torch
    return (torch.as_tensor(np.asarray(m.row_ptr, np.int64), device="cuda"),
            torch.as_tensor(np.asarray(m.col, np.int32), device="cuda"),
            torch.as_tensor(np.asarray(m.val, np.float32), device="cuda"))


@pytest.mark.parametrize("frac", [0.05, 0.5])
def test_beams_device_random(lib, frac):
    from test_analysis import random_csr
    import torch
    rng = np.random.default_rng(3)
    dense, m = random_csr(rng, 40, 3000)
    rp, cl, vl = to_dev(m)
    sizes, mpxi = analysis.beams_device(rp, cl, vl, m.shape[1], frac)
    torch.cuda.synchronize()
    hs, hm = analysis.beams_host(m, frac)
    np.testing.assert_array_equal(sizes.cpu().numpy(), hs)
    np.testing.assert_array_equal(mpxi.cpu().numpy(), hm)
    # accumulation into an existing MPXI buffer
    analysis.beams_device(rp, cl, vl, m.shape[1], frac, mpxi=mpxi)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(mpxi.cpu().numpy(), 2 * hm)


def test_beams_device_pe_csr(lib):
    import torch
    g = params.tiny(seed=5, n_det_side=4, n_layers=2, image=(8, 8, 8), holes_side=3)
    dense = torch.as_tensor(pe_mod.pe_compute(g, sub=(2, 2, 2)), device="cuda")
    rp, cl, vl, _ = csr.csr_from_dense_device(dense, 1e-9)
    sizes, mpxi = analysis.beams_device(rp, cl, vl, g.n_vox, 0.05)
    torch.cuda.synchronize()
    host = csr.CsrMatrix(rp.cpu().numpy(), cl.cpu().numpy(), vl.cpu().numpy(), (g.n_det, g.n_vox))
    hs, hm = analysis.beams_host(host, 0.05)
    np.testing.assert_array_equal(sizes.cpu().numpy(), hs)
    np.testing.assert_array_equal(mpxi.cpu().numpy(), hm)
    # same beams from the dense rows directly
    d = dense.cpu().numpy().astype(np.float64)
    mx = d.max(axis=1, keepdims=True)
    mask = (d > 0.05 * mx) & (mx > 0)
    np.testing.assert_array_equal(hm, mask.sum(axis=0))
    assert hm.max() > 1
