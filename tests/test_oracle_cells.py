"""Pins of the CPU oracle's plate-cell collimator model (no reference kernel).

The cell model reuses the hole-site physics after the site (pinned with the
reference-mode collimator tests); what is new is pinned here:
  * the plate path of a segment (mu-weighted length through plates minus holes):
    closed forms for axis-parallel, oblique, hole-axis and cross-hole segments,
    segments ending inside a plate, and dense midpoint sampling of random segments;
  * the cell list: counts and membership against a brute-force numpy recount.
"""
import numpy as np
import pytest

import oracle
from sysmat import params

MU = params.W_MU  # tungsten total / photoelectric / Compton (1/mm)


def c3():
    return params.config3()


def test_plate_path_along_y_through_material():
    g = c3()
    f = float(g.img[11])
    # x = 10 is a hole column, z = 1 is 1 mm from the nearest hole axis (R = 0.5)
    p = oracle.plate_path(g, [10.0, f - 10.0, 1.0], [10.0, f + 10.0, 1.0])
    np.testing.assert_allclose(p, np.array(MU, np.float32).astype(np.float64) * 5.0, rtol=1e-12)


def test_plate_path_along_hole_axis_is_free():
    g = c3()
    f = float(g.img[11])
    p = oracle.plate_path(g, [2.0, f - 10.0, 2.0], [2.0, f + 10.0, 2.0])
    np.testing.assert_allclose(p, 0.0, atol=1e-12)


def test_plate_path_oblique_and_partial():
    g = c3()
    f = float(g.img[11])
    a = np.array([0.0, f - 10.0, 0.0])           # (0, 0) lies between holes (nearest axis 2.83 mm away)
    d = np.array([0.1, 1.0, 0.05])
    b = a + 20.0 * d
    mu = np.array(MU, np.float32).astype(np.float64)
    L = 5.0 * np.linalg.norm(d) / d[1]            # thickness / cos
    # the path must stay clear of holes: check the lateral excursion inside the plate
    ts = np.linspace((f - 2.5 - a[1]) / d[1], (f + 2.5 - a[1]) / d[1], 50)
    assert np.all(np.hypot(a[0] + ts * d[0] - 2.0, a[2] + ts * d[2] - 2.0) > 0.5)
    np.testing.assert_allclose(oracle.plate_path(g, a, b), mu * L, rtol=1e-12)
    # ending 1.5 mm into the plate
    e = a + d * ((f - 2.5 + 1.5) - a[1]) / d[1]
    np.testing.assert_allclose(oracle.plate_path(g, a, e), mu * 1.5 * np.linalg.norm(d) / d[1], rtol=1e-12)
    # reversed direction gives the same sums
    np.testing.assert_allclose(oracle.plate_path(g, b, a), oracle.plate_path(g, a, b), rtol=1e-12)


def test_plate_path_across_a_hole():
    g = c3()
    f = float(g.img[11])
    # along x through the middle of the plate at the z of a hole row: the row's 6 holes (R 0.5) are air
    p = oracle.plate_path(g, [-40.0, f, 2.0], [40.0, f, 2.0])
    mu = np.array(MU, np.float32).astype(np.float64)
    np.testing.assert_allclose(p, mu * (64.0 - 6 * 1.0), rtol=1e-12)


def _sampled_path(g, a, b, n=200001):
    """Midpoint-rule integral of mu along a -> b by point classification."""
    cells_col = g.col
    f = float(g.img[11])
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    t = (np.arange(n) + 0.5) / n
    pts = a[None, :] + t[:, None] * (b - a)[None, :]
    mu = np.zeros((n, 3))
    nl = int(cells_col[0])
    h0 = 0
    for m in range(nl):
        L = cells_col[(m + 1) * 10:(m + 1) * 10 + 8].astype(np.float64)
        nh = int(L[0])
        inbox = ((np.abs(pts[:, 0]) <= L[1] / 2) & (np.abs(pts[:, 1] - (f + L[4])) <= L[2] / 2)
                 & (np.abs(pts[:, 2]) <= L[3] / 2))
        air = np.zeros(n, bool)
        for q in range(h0, h0 + nh):
            H = cells_col[100 + 9 * q:100 + 9 * q + 9].astype(np.float64)
            ya, yb = sorted((H[1] + f, H[2] + f))
            air |= ((pts[:, 0] - H[0]) ** 2 + (pts[:, 2] - H[3]) ** 2 < H[4] ** 2) & (pts[:, 1] >= ya) & (pts[:, 1] <= yb)
        mu[inbox & ~air] += L[5:8]
        h0 += nh
    return mu.sum(axis=0) * np.linalg.norm(b - a) / n


@pytest.mark.parametrize("which", ["c3", "tiny2"])
def test_plate_path_matches_sampling(which):
    g = c3() if which == "c3" else params.tiny(seed=3, n_det_side=3, n_layers=1, holes_side=3, n_plates=2)
    f = float(g.img[11])
    rng = np.random.default_rng(7)
    for _ in range(12):
        a = np.array([rng.uniform(-12, 12), f - rng.uniform(6, 20), rng.uniform(-12, 12)])
        b = np.array([rng.uniform(-12, 12), f + rng.uniform(0, 20), rng.uniform(-12, 12)])
        # steer half the segments through a hole axis region
        if rng.random() < 0.5:
            q = rng.integers(0, int(g.col[10]))
            H = g.col[100 + 9 * q:100 + 9 * q + 9]
            b[0], b[2] = 2 * H[0] - a[0] + rng.normal(0, 0.3), 2 * H[3] - a[2] + rng.normal(0, 0.3)
            b[1] = 2 * f - a[1]
        exact = oracle.plate_path(g, a, b)
        sampled = _sampled_path(g, a, b)
        np.testing.assert_allclose(exact, sampled, rtol=2e-3, atol=2e-3)


def _brute_cells(g, h, keep):
    f = float(g.img[11])
    out = []
    h0 = 0
    for m in range(int(g.col[0])):
        L = g.col[(m + 1) * 10:(m + 1) * 10 + 8].astype(np.float64)
        nh = int(L[0])
        H = g.col[100 + 9 * h0:100 + 9 * (h0 + nh)].reshape(nh, 9).astype(np.float64)
        h0 += nh
        n = [max(1, int(round(L[k] / h))) for k in (1, 2, 3)]
        s = [L[k] / n[i] for i, k in enumerate((1, 2, 3))]
        x = -L[1] / 2 + (np.arange(n[0]) + 0.5) * s[0]
        y = f + L[4] - L[2] / 2 + (np.arange(n[1]) + 0.5) * s[1]
        z = -L[3] / 2 + (np.arange(n[2]) + 0.5) * s[2]
        Y, Z, X = np.meshgrid(y, z, x, indexing="ij")
        rho = np.sqrt((X[..., None] - H[:, 0]) ** 2 + (Z[..., None] - H[:, 3]) ** 2)
        ya = np.minimum(H[:, 1], H[:, 2]) + f
        yb = np.maximum(H[:, 1], H[:, 2]) + f
        air = ((rho < H[:, 4]) & (Y[..., None] >= ya) & (Y[..., None] <= yb)).any(-1)
        near = np.ones_like(air) if keep < 0 else (rho - H[:, 4] <= keep).any(-1)
        sel = ~air & near
        w = L[7] * s[0] * s[1] * s[2] / (4 * np.pi)
        out.append(np.stack([X[sel], Y[sel], Z[sel], np.full(sel.sum(), w)], 1))
    return np.concatenate(out)


@pytest.mark.parametrize("keep", [-1.0, 0.5, 1.0, 2.0])
def test_cell_list_brute_force(keep):
    g = c3()
    ref = _brute_cells(g, 0.5, keep)
    got = oracle.cells(g, 0.5, keep)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def test_cell_count_closed_form():
    # 64 x 5 x 64 mm plate in 0.5 mm cells; each 1 mm hole (R 0.5 centred on a cell corner)
    # holds the centres of 4 cells per cell layer: 36 holes x 4 x 10 layers are air
    assert len(oracle.cells(c3(), 0.5, -1.0)) == 128 * 10 * 128 - 36 * 4 * 10
    w = oracle.cells(c3(), 0.5, -1.0)[:, 3]
    np.testing.assert_allclose(w, float(np.float32(MU[2])) * 0.125 / (4 * np.pi), rtol=1e-12)


def test_cells_scatter_zero_without_compton_plate():
    g = params.tiny(seed=1, n_det_side=3, n_layers=1, image=(2, 2, 2), holes_side=2)
    g.col[17] = 0.0  # plate mu_c = 0: no cells
    assert len(oracle.cells(g, 1.0, -1.0)) == 0
    out = oracle.scatter_cells(g, 1.0, -1.0)
    assert np.all(out == 0)
