"""Pins for the oracle's ray/solid primitives (CPU).

chord_box       <- length_box_ray        PESysMatGen.cu:19-151
chord_box_open  <- length_box_ray_inside PESysMatGen.cu:153-307
chord_hole      <- length_cylinder_ray   PESysMatGen.cu:309-373

Each function is checked against closed forms and against brute-force
sampling of the segment (independent of the slab method), in FP32 and FP64.
"""
import numpy as np
import pytest

import oracle


def brute_box(p, q, lo, hi, n=400001):
    t = np.linspace(0.0, 1.0, n)
    pts = p[None, :] + (q - p)[None, :] * t[:, None]
    inside = np.all((pts >= lo) & (pts <= hi), axis=1)
    return inside.mean() * np.linalg.norm(q - p)


def brute_cyl(p, q, xc, y1, y2, zc, r, n=400001):
    t = np.linspace(0.0, 1.0, n)
    pts = p[None, :] + (q - p)[None, :] * t[:, None]
    inside = ((pts[:, 0] - xc) ** 2 + (pts[:, 2] - zc) ** 2 <= r * r) & (pts[:, 1] >= min(y1, y2)) & (
        pts[:, 1] <= max(y1, y2))
    return inside.mean() * np.linalg.norm(q - p)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_box_axis_and_diagonal(prec):
    # straight through a 2x3x4 box along y (two parallel axes): chord = thickness
    assert oracle.chord_box([0.5, -5, 1, 0.5, 9, 1, 0, 0, 0, 2, 3, 4], prec) == pytest.approx(3.0, rel=1e-6)
    # along z, start inside => 0 (segment must enter and leave)
    assert oracle.chord_box([0.5, 1, 2, 0.5, 1, 9, 0, 0, 0, 2, 3, 4], prec) == 0.0
    # corner-to-corner diagonal of a unit cube from outside: sqrt(3)
    args = [-1, -1, -1, 2, 2, 2, 0, 0, 0, 1, 1, 1]
    assert oracle.chord_box(args, prec) == pytest.approx(np.sqrt(3.0), rel=1e-5)
    # miss
    assert oracle.chord_box([5, -5, 5, 5, 9, 5.5, 0, 0, 0, 2, 3, 4], prec) == 0.0


@pytest.mark.parametrize("seed", range(12))
def test_box_oblique_brute_force(seed):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(-1, 0, 3)
    hi = lo + rng.uniform(0.5, 2.0, 3)
    c = (lo + hi) / 2
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    off = rng.normal(size=3) * 0.3
    p = c + off - 6 * d
    q = c + off + 6 * d
    ref = brute_box(p, q, lo, hi)
    args = np.concatenate([p, q, lo, hi])
    assert oracle.chord_box(args, "f64") == pytest.approx(ref, abs=1e-4)
    assert oracle.chord_box(args, "f32") == pytest.approx(ref, abs=1e-4)


@pytest.mark.parametrize("seed", range(8))
def test_box_open_end_inside(seed):
    rng = np.random.default_rng(100 + seed)
    lo = np.array([-1.0, -1.5, -0.8])
    hi = -lo
    q = rng.uniform(lo * 0.8, hi * 0.8)  # end point inside
    d = rng.normal(size=3)
    d[1] = abs(d[1]) + 0.3  # |dy| well above the 0.001 cut
    d /= np.linalg.norm(d)
    p = q - 20 * d
    ref = brute_box(p, q, lo, hi)
    args = np.concatenate([p, q, lo, hi])
    assert oracle.chord_box(args, "f64", open_end=True) == pytest.approx(ref, abs=1e-4)
    # the "through" version rejects it: the segment ends inside
    assert oracle.chord_box(args, "f64") == 0.0


def test_box_open_rules():
    # |dy| < 0.001 -> 0 regardless of geometry
    assert oracle.chord_box([0, 0.5, 0, 0.5, 0.5004, 0.5, -1, 0, -1, 1, 1, 1], "f64", open_end=True) == 0.0
    # along y, end inside: y_out - y1
    assert oracle.chord_box([0.2, -5, 0.1, 0.2, 0.25, 0.1, -1, -1, -1, 1, 1, 1], "f64",
                            open_end=True) == pytest.approx(1.25)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_hole_closed_forms(prec):
    # on axis: chord = y2 - y1
    assert oracle.chord_hole([0, 0, 0, 0, 50, 0, 0, 20, 25, 0, 0.5], prec) == pytest.approx(5.0, rel=1e-6)
    # oblique through both caps: (y2 - y1) / |k_y|
    p, q = np.array([0.0, 0, 0]), np.array([1.0, 50, 0.5])
    k = (q - p) / np.linalg.norm(q - p)
    xc, zc = p[0] + k[0] / k[1] * 22.5, p[2] + k[2] / k[1] * 22.5
    got = oracle.chord_hole(np.concatenate([p, q, [xc, 20, 25, zc, 0.5]]), prec)
    assert got == pytest.approx(5.0 / k[1], rel=1e-5)
    # miss
    assert oracle.chord_hole([0, 0, 0, 0, 50, 0, 3, 20, 25, 0, 0.5], prec) == 0.0
    # degenerate faces (|y1 - y2| < 0.001) -> 0
    assert oracle.chord_hole([0, 0, 0, 0, 50, 0, 0, 20, 20.0005, 0, 0.5], prec) == 0.0


@pytest.mark.parametrize("seed", range(16))
def test_hole_wall_entries_brute_force(seed):
    rng = np.random.default_rng(200 + seed)
    xc, zc, r, y1, y2 = 0.1, -0.2, 0.5, 20.0, 25.0
    # rays from the image side aimed near the hole so wall hits occur
    p = np.array([rng.uniform(-8, 8), 0.0, rng.uniform(-8, 8)])
    aim = np.array([xc + rng.uniform(-0.7, 0.7), rng.uniform(y1, y2), zc + rng.uniform(-0.7, 0.7)])
    d = aim - p
    q = p + d / d[1] * 60.0
    ref = brute_cyl(p, q, xc, y1, y2, zc, r)
    args = np.concatenate([p, q, [xc, y1, y2, zc, r]])
    assert oracle.chord_hole(args, "f64") == pytest.approx(ref, abs=2e-4)


def test_fp32_fp64_agree_on_primitives():
    rng = np.random.default_rng(7)
    for _ in range(200):
        lo = rng.uniform(-2, 0, 3)
        hi = lo + rng.uniform(0.3, 3, 3)
        p = rng.uniform(-10, 10, 3)
        q = rng.uniform(-10, 10, 3)
        a = np.concatenate([p, q, lo, hi])
        assert oracle.chord_box(a, "f32") == pytest.approx(oracle.chord_box(a, "f64"), abs=5e-5)
