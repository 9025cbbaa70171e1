"""The allocation-free PE entry point (sm_pe_compute_device_noalloc) against
sm_pe_compute_device, through the C ABI.

Both run the same tracer kernels; the allocation-free form keeps its geometry
tables, candidate lists and status word in the library's file-scope device
array instead of hipMalloc'd buffers, so its output must equal the allocating
entry point's bit for bit.  Two calls in a row, the second on a smaller input,
show that contents the first call left in the shared array do not leak into the
second.  A geometry whose tables exceed the array is refused before any kernel
runs."""
import ctypes

import numpy as np
import pytest

from sysmat import _lib, params
from sysmat.pe import _f32, _padded_img, options

pytestmark = pytest.mark.gpu


def run_both(lib, geo, **kw):
    import torch
    o = options(**kw)
    X, Y, Z = geo.shape
    rows = o.det_count if o.det_count > 0 else geo.n_det - o.det_begin
    zc = o.z_count if o.z_count > 0 else Z - o.z_begin
    col, det, img = _f32(geo.col), _f32(geo.det), _padded_img(geo.img)
    outs = []
    for fn in (lib.sm_pe_compute_device, lib.sm_pe_compute_device_noalloc):
        dst = torch.full((rows, X * Y * zc), float("nan"), dtype=torch.float32, device="cuda:0")
        print(f"calling {fn.__name__} rows={rows} voxels={X * Y * zc}", flush=True)
        rc = fn(_lib.ptr(col), _lib.ptr(det), _lib.ptr(img), ctypes.byref(o), _lib.ptr(dst), 0, None)
        _lib.check(rc)
        torch.cuda.synchronize()
        outs.append(dst.cpu().numpy())
    return outs


def test_noalloc_equals_allocating_entry_two_calls(lib):
    g1 = params.config1()
    a, b = run_both(lib, g1, det_range=(500, 532), voxel_sub=1, sub=(8, 16, 8))
    assert np.isfinite(a).all() and a.max() > 0
    assert np.array_equal(a, b)
    # second call: a smaller geometry and fewer rows, other sampling
    g2 = params.tiny(seed=3, n_det_side=3, n_layers=2, image=(4, 3, 5), holes_side=2, n_plates=2)
    a, b = run_both(lib, g2, voxel_sub=2, sub=(2, 2, 2), det_range=(2, 9))
    assert np.isfinite(a).all() and a.max() > 0
    assert np.array_equal(a, b)


def test_noalloc_refuses_tables_beyond_the_scratch(lib):
    """16384 crystals stacked one behind the other on the image axis: every crystal lies in
    front of every detector behind it, so the per-detector candidate lists hold ~N^2/2 =
    1.3e8 16-bit indices (268 MB), beyond the 64 MiB array.  The count pass runs, the
    list carve fails, and no tracer kernel is launched."""
    import torch
    cap = int(lib.sm_pe_noalloc_scratch_bytes())
    assert cap == 64 << 20
    g = params.tiny(seed=1)
    n = 16384
    det = np.zeros(1 + 12 * n, np.float32)
    det[0] = n
    rec = det[1:].reshape(n, 12)
    rec[:] = g.det[1:13]
    rec[:, 0] = 0.0
    rec[:, 2] = 0.0
    rec[:, 1] = 20.0 + 0.5 * np.arange(n, dtype=np.float32)  # y, 0.5 mm apart
    rec[:, 3:6] = (1.0, 0.4, 1.0)  # width, thickness, height
    rec[:, 10] = 0.0
    assert rec[0, 6] > 0.01
    o = options(det_range=(0, 1), z_range=(0, 1))
    col, img = _f32(g.col), _padded_img(g.img)
    X, Y, _ = g.shape
    dst = torch.zeros(X * Y, dtype=torch.float32, device="cuda:0")  # one row of one plane: in bounds either way
    print(f"calling sm_pe_compute_device_noalloc with {n} stacked crystals", flush=True)
    rc = lib.sm_pe_compute_device_noalloc(_lib.ptr(col), _lib.ptr(det), _lib.ptr(img), ctypes.byref(o),
                                          _lib.ptr(dst), 0, None)
    msg = lib.sm_last_error()
    assert rc == _lib.SM_E_ARG, msg
    assert b"allocation-free scratch" in msg
