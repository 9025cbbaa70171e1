"""Pins for the PE oracle (restatement of photodetectorCudaMe, PESysMatGen.cu:375-764).

The reference ships no expected outputs, so the oracle is pinned by what the
physics fixes:
* a far isolated crystal: the sub-cell sum telescopes to
  Omega * (mu_pe/mu_t) * (1 - exp(-mu_t * T)) (solid angle, self attenuation
  and absorption terms; a dropped term or wrong sign fails it);
* a plate without holes multiplies it by exp(-mu * T / cos) (PESysMatGen.cu:661-685);
* a hole on the line of sight removes the plate again;
* a crystal in front multiplies it by exp(-mu_k * t_k) (PESysMatGen.cu:687-725);
* a zero-attenuation detector (mu_t <= 0.01) gives 0 (PESysMatGen.cu:504, 598);
* mirror symmetry x -> -x and z -> -z of a symmetric geometry;
* the FP32 restatement stays within its documented rounding envelope of FP64.
"""
import numpy as np
import pytest

import oracle
from sysmat import params

D_FAR = 1.0e5


def far_geometry(layers=(), holes=None, blockers=(), mu=(0.35, 0.28), fov2col=0.0, det_y=D_FAR):
    cry = [[0.0, det_y, 0.0, 1.6, 3.0, 1.6, mu[0], mu[1], 0.07, 0.1, 0.0, 1.0]]
    for (y, t, w, m) in blockers:
        cry.append([0.0, y, 0.0, w, t, w, m, m * 0.8, m * 0.2, 0.1, 0.0, 1.0])
    h = np.zeros((0, 10)) if holes is None else holes
    return params.build(list(layers), h, np.array(cry, np.float32),
                        dict(shape=(1, 1, 1), voxel=(1.0, 1.0, 1.0), fov2col=fov2col), params.PHYS_140)


def closed_form(mu_t=0.35, mu_pe=0.28, w=1.6, h=1.6, t=3.0, d=D_FAR):
    return w * h / (4 * np.pi * d * d) * (mu_pe / mu_t) * (1 - np.exp(-mu_t * t))


@pytest.mark.parametrize("sub", [(8, 16, 8), (2, 2, 2)])
def test_far_crystal_closed_form(sub):
    g = far_geometry()
    v = oracle.pe(g, sub=sub, det_range=(0, 1), precision="f64")[0, 0]
    assert v == pytest.approx(closed_form(), rel=2e-4)


def test_plate_attenuation_closed_form():
    T, mu = 5.0, 0.3
    lay = [dict(width=100.0, thickness=T, height=100.0, offset=0.0, mut=mu, mupe=0.2, muc=0.1)]
    g = far_geometry(layers=lay, fov2col=D_FAR / 2, det_y=D_FAR / 2)
    v = oracle.pe(g, det_range=(0, 1), precision="f64")[0, 0]
    assert v == pytest.approx(closed_form() * np.exp(-mu * T), rel=2e-4)


def test_hole_on_axis_removes_plate():
    T = 5.0
    lay = [dict(width=100.0, thickness=T, height=100.0, offset=0.0, mut=3.9, mupe=3.78, muc=0.12)]
    holes = params.hole_grid(1, 1.0, 4.0, -T / 2, T / 2, 0)
    g = far_geometry(layers=lay, holes=holes, fov2col=D_FAR / 2, det_y=D_FAR / 2)
    v = oracle.pe(g, det_range=(0, 1), precision="f64")[0, 0]
    assert v == pytest.approx(closed_form(), rel=2e-4)


def test_plate_miss_blocks():
    # the line of sight passes beside a narrow plate: length_box_ray < 0.1 -> +1e5 (PESysMatGen.cu:656-658)
    lay = [dict(width=1.0, thickness=5.0, height=1.0, offset=0.0, mut=0.0, mupe=0.0, muc=0.0)]
    g = far_geometry(layers=lay, fov2col=D_FAR / 2, det_y=D_FAR / 2)
    g.col[11 + 3] = 1.0
    g.img[8] = 50.0  # shift the FOV sideways so the ray misses the plate box
    v = oracle.pe(g, det_range=(0, 1), precision="f64")[0, 0]
    assert v == 0.0


def test_blocking_crystal_closed_form():
    mu_k, t_k = 0.5, 2.0
    g = far_geometry(blockers=[(D_FAR - 20.0, t_k, 6.0, mu_k)])
    v = oracle.pe(g, det_range=(0, 1), precision="f64")[0, 0]
    assert v == pytest.approx(closed_form() * np.exp(-mu_k * t_k), rel=2e-4)


def test_zero_attenuation_detector_is_zero():
    g = far_geometry(mu=(0.01, 0.005))
    assert oracle.pe(g, det_range=(0, 1), precision="f64")[0, 0] == 0.0


def symmetric_geometry():
    t = 3.0
    lay = [dict(width=30.0, thickness=t, height=30.0, offset=0.0, mut=3.9, mupe=3.78, muc=0.12)]
    holes = params.hole_grid(3, 3.0, 1.2, -t / 2, t / 2, 0)
    cry = params.crystal_layers(4, 1.6, 3.0, t / 2 + 6.0, 2)
    return params.build(lay, holes, cry, dict(shape=(4, 4, 2), voxel=(1.0, 1.0, 1.0), fov2col=25.0), params.PHYS_140)


def test_mirror_symmetry_fp64():
    g = symmetric_geometry()
    a = oracle.pe(g, sub=(2, 4, 2), precision="f64")  # (32 det, 32 vox)
    n = 4
    X, Y, Z = g.shape
    for d in range(g.n_det):
        layer, rem = divmod(d, n * n)
        zi, xi = divmod(rem, n)
        dm = layer * n * n + zi * n + (n - 1 - xi)  # x mirror
        dz = layer * n * n + (n - 1 - zi) * n + xi  # z mirror
        for v in range(X * Y * Z):
            z, r = divmod(v, X * Y)
            y, x = divmod(r, X)
            vm = z * X * Y + y * X + (X - 1 - x)
            vz = (Z - 1 - z) * X * Y + y * X + x
            assert a[d, v] == pytest.approx(a[dm, vm], rel=1e-9, abs=1e-30)
            assert a[d, v] == pytest.approx(a[dz, vz], rel=1e-9, abs=1e-30)


def test_monotone_in_plate_mu():
    g = symmetric_geometry()
    base = oracle.pe(g, sub=(2, 4, 2), det_range=(0, 8), precision="f64")
    g2 = g.copy()
    g2.col[15] = 5.0
    more = oracle.pe(g2, sub=(2, 4, 2), det_range=(0, 8), precision="f64")
    assert np.all(more <= base * (1 + 1e-12))
    assert np.any(more < base)


def test_fp32_envelope_and_mixed_precision():
    """FP32 restatement vs FP64: rounding envelope of the reference algorithm.
    The mixed build (FP64 accumulator + collimator path + absorption chord,
    the numerics of the HIP kernel) is within 1e-5 of FP64."""
    g = params.config1()
    rows = (528, 530)
    f64 = oracle.pe(g, det_range=rows, precision="f64")
    f32 = oracle.pe(g, det_range=rows, precision="f32")
    mix = oracle.pe(g, det_range=rows, precision="mixed")
    sig = f64 >= f64.max(axis=1, keepdims=True) * 1e-9
    r32 = np.abs(f32 - f64)[sig] / f64[sig]
    rmx = np.abs(mix - f64)[sig] / f64[sig]
    assert r32.max() < 2e-4
    assert rmx.max() < 1e-5
