"""Pins the composition of the scatter oracle (oracle/scatter_oracle.cpp: site_term,
crystal_run, collimator_run and the two bitmap builders) against results the
physics fixes independently of that code.

The reference (ScatterGen_RayTracing_CircularHole/scatter.cu) evaluates, for
detector r, image point I and scatter site S (crystal k, scatter.cu:1232-1430, or
layer-0 hole h, :1728-1929):

    source weight x Range_phi / 2 pi x KN(theta_min..theta_max) / KN(0..pi) x p_E
        x exp(-sum mu'(E') chord) x (1 - exp(-mu_tot' l_abs)) mu_pe' / mu_tot'

With one site and one detector and nothing in between, each factor has a closed
form that this file builds from geometry and textbook physics (numpy, float64):
the scatter angle from the two direction vectors; E' from the Compton formula;
the Klein-Nishina shape in its textbook form P^2 (P + 1/P - sin^2 theta), P = E'/E
(the reference writes it differently, scatter.cu:961-976); the window fraction
from math.erf; the absorption chord of a ray through a box centre as
2 min(half extent / |direction|); mu_pe scaled by (E0/E')^3 and mu_c by E0/E'.
Dropping a factor, swapping theta_min and theta_max, using E^-1 where E^-3
belongs, or taking mu_c/mu_pe from the scatter crystal instead of the detecting
one (the reference's choice, scatter.cu:1218-1219) fails these tests.

Further pins: switches that must give exactly zero (Compton off, mu_c = 0 in the
detecting crystal, a window that excludes E'); a crystal on the site -> detector
line multiplies the element by exp(-mu'(E') t) and the same crystal off the line
changes nothing; mirror symmetry x -> -x of a symmetric geometry; the bitmap
builders (scatter.cu:1440-1540, 1939-2009) against a brute-force recount of their
criteria and the triangular pair index against enumeration."""
import math

import numpy as np
import pytest

import oracle
from sysmat import params

F2C = 50.0      # Img[11]: FOV centre to collimator
E0 = 140.5
H32 = float(np.float32(0.01))       # the reference's 0.01f KN step, widened
PI32 = float(np.float32(math.pi))   # (float)M_PI, the upper end of the normalisation sum
REL = 1e-6      # the oracle writes float32


def crystal(x, y, z, w=1.6, t=3.0, h=1.6, mu=(0.35, 0.28, 0.07), res=0.1, rot=0.0):
    return [x, y, z, w, t, h, mu[0], mu[1], mu[2], res, rot, 1.0]


def make_geo(crystals, holes=(), plate=(64.0, 5.0, 64.0, 0.0, 3.9, 3.78, 0.12), window=(126.45, 154.55),
             use_compton=1.0):
    n = len(crystals)
    det = np.zeros(1 + 12 * n, np.float32)
    det[0] = n
    det[1:] = np.asarray(crystals, np.float32).ravel()
    col = np.zeros(100 + 9 * max(1, len(holes)), np.float32)
    col[0] = 1
    col[10] = len(holes)
    col[11:18] = plate
    for i, hrec in enumerate(holes):
        col[100 + 9 * i:100 + 9 * i + 9] = hrec
    img = np.zeros(100, np.float32)
    img[0:3] = 1
    img[3:6] = 1.0
    img[6] = 1
    img[11] = F2C
    phy = np.zeros(100, np.float32)
    phy[0] = use_compton
    phy[4] = 1
    phy[5], phy[6] = window
    phy[7] = E0
    return params.Geometry(col, det, img, phy)


def world(c):
    """Crystal centre in the frame the kernels use (y + Img[11])."""
    return np.array([c[0], c[1] + F2C, c[2]], np.float64)


def kn_textbook(theta):
    p = 1.0 / (1.0 + E0 / 511.0 * (1.0 - np.cos(theta)))
    return p * p * (p + 1.0 / p - np.sin(theta) ** 2)


def kn_sum(lo, hi):
    n = int((hi - lo) / H32)
    th = lo + np.arange(n + 1, dtype=np.float64) * H32
    return float(np.sum(kn_textbook(th) * np.sin(th) * H32))


def angle(a, b):
    return math.acos(max(-1.0, min(1.0, float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))))))


def site_factor(I, S, r, window=(126.45, 154.55)):
    """Everything after the source weight for site S and detecting crystal r (unrotated,
    nothing attenuating), from closed forms."""
    Q = world(r)
    w, t, h = r[3], r[4], r[5]
    mu_pe, mu_c, res = r[7], r[8], r[9]
    theta = angle(S - I, Q - S)
    e = E0 / (1.0 + E0 / 511.0 * (1.0 - math.cos(theta)))
    res_e = res * math.sqrt(e / E0)
    lo, hi = window
    if e * (1 + 2 * res_e / 2.35482) <= lo or e * (1 - 2 * res_e / 2.35482) >= hi:
        return 0.0
    sig = res_e / 2.35482 * e
    p_e = 0.5 * (math.erf((hi - e) / (math.sqrt(2) * sig)) - math.erf((lo - e) / (math.sqrt(2) * sig)))
    L = float(np.linalg.norm(Q - S))
    rdet = math.sqrt(w * w + h * h + t * t) / 2
    s = math.sin(theta)
    phi = 2 * math.pi if 2 * L * s <= rdet else 4 * math.asin(min(rdet / 2 / L / s, 1.0))
    corners = [Q + np.array([sx * w / 2, sy * t / 2, sz * h / 2]) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
    th = [angle(S - I, c - S) for c in corners]
    tmin, tmax = min(th), max(th)
    ratio = kn_sum(tmin, tmax) / kn_sum(0.0, PI32)
    d = (Q - S) / L
    half = np.array([w / 2, t / 2, h / 2])
    l_abs = 2 * min(half[i] / abs(d[i]) for i in range(3) if d[i] != 0)
    pe_e = mu_pe * (E0 / e) ** 3
    c_e = mu_c * (E0 / e)
    tot = pe_e + c_e
    absorb = (1 - math.exp(-l_abs * tot)) * pe_e / tot
    return phi / (2 * math.pi) * ratio * p_e * absorb, dict(theta=theta, e=e, chord_dir=d)


# Site k at (0, 10, 0) + F2C, straight above the voxel at the origin; detector r 20 mm
# away at a 30 degree scatter angle.
K = crystal(0.0, 10.0, 0.0, mu=(0.5, 0.3, 0.2))
R30 = crystal(20.0 * math.sin(math.pi / 6), 10.0 + 20.0 * math.cos(math.pi / 6), 0.0)
PE_K = 2.0e-3


def crystal_element(g, pe_rows, r, precision="f64"):
    bits = oracle.bitmap_crystal(g)
    pe = np.asarray(pe_rows, np.float32).reshape(g.n_det, g.n_vox)
    return float(oracle.scatter_crystal(g, pe, bits, det_range=(r, r + 1), precision=precision)[0, 0])


def test_single_site_equals_closed_form():
    g = make_geo([K, R30])
    got = crystal_element(g, [PE_K, 7e-4], 1)
    f, info = site_factor(np.zeros(3), world(K), R30)
    assert abs(info["theta"] - math.pi / 6) < 1e-6
    want = PE_K * R30[8] / R30[7] * f  # mu_c / mu_pe of the DETECTING crystal (scatter.cu:1218-1219)
    assert want > 0
    assert abs(got - want) <= REL * want, (got, want)
    got32 = crystal_element(g, [PE_K, 7e-4], 1, precision="f32")
    assert abs(got32 - want) <= 2e-4 * want
    # the scatter crystal's own ratio would be a different number: the test tells them apart
    assert abs(PE_K * K[8] / K[7] * f - want) > 0.1 * want


@pytest.mark.parametrize("deg", [12.0, 25.0, 40.0])
def test_single_site_other_angles_and_distances(deg):
    a = math.radians(deg)
    for dist in (8.0, 35.0):
        r = crystal(dist * math.sin(a), 10.0 + dist * math.cos(a), 1.0, w=1.2, t=2.0, h=0.9,
                    mu=(0.42, 0.3, 0.12))
        g = make_geo([K, r])
        got = crystal_element(g, [PE_K, 1e-3], 1)
        f, _ = site_factor(np.zeros(3), world(K), r)
        want = PE_K * r[8] / r[7] * f
        assert want > 0
        assert abs(got - want) <= REL * want, (deg, dist, got, want)


def test_switches_that_give_zero():
    g = make_geo([K, R30], use_compton=0.0)
    assert crystal_element(g, [PE_K, 1e-3], 1) == 0.0
    r = list(R30)
    r[8] = 0.0  # mu_c of the detecting crystal
    g = make_geo([K, r])
    assert crystal_element(g, [PE_K, 1e-3], 1) == 0.0
    # E' = 135.5 keV, 2 sigma band up to ~146.8 keV: a window from 150 keV excludes it
    g = make_geo([K, R30], window=(150.0, 170.0))
    assert site_factor(np.zeros(3), world(K), R30, window=(150.0, 170.0)) == 0.0
    assert crystal_element(g, [PE_K, 1e-3], 1) == 0.0
    # a window that only clips the photopeak gives the erf fraction, not zero
    g = make_geo([K, R30], window=(136.0, 170.0))
    f, _ = site_factor(np.zeros(3), world(K), R30, window=(136.0, 170.0))
    got = crystal_element(g, [PE_K, 1e-3], 1)
    assert 0 < got and abs(got - PE_K * R30[8] / R30[7] * f) <= REL * got


def test_crystal_on_the_line_attenuates_by_its_chord():
    base = crystal_element(make_geo([K, R30]), [PE_K, 1e-3], 1)
    S, Q = world(K), world(R30)
    M = 0.5 * (S + Q)
    mu = (0.6, 0.45, 0.15)
    a = crystal(M[0], M[1] - F2C, M[2], w=1.0, t=1.2, h=1.0, mu=mu)
    g = make_geo([K, R30, a])
    got = crystal_element(g, [PE_K, 1e-3, 0.0], 1)  # the attenuator is no source (PE 0)
    _, info = site_factor(np.zeros(3), S, R30)
    d = info["chord_dir"]
    half = np.array([a[3] / 2, a[4] / 2, a[5] / 2])
    t = 2 * min(half[i] / abs(d[i]) for i in range(3) if d[i] != 0)
    e = info["e"]
    mu_e = mu[1] * (E0 / e) ** 3 + mu[2] * (E0 / e)  # E^-3 photoelectric, E^-1 Compton
    want = base * math.exp(-mu_e * t)
    assert got < base
    assert abs(got - want) <= 2 * REL * want, (got, want, base)
    # E^-1 for both would be a different number
    assert abs(base * math.exp(-(mu[1] + mu[2]) * (E0 / e) * t) - want) > 1e-3 * want
    # the same crystal 4 mm off the line (outside the ray, inside the bounding-sphere
    # criterion's reach or not) changes nothing
    off = list(a)
    off[2] = 4.0
    got_off = crystal_element(make_geo([K, R30, off]), [PE_K, 1e-3, 0.0], 1)
    assert got_off == base


def test_hole_site_weight_closed_form():
    """collimatorScatterSysMatCuda (scatter.cu:1544-1936, virtual_r 0.1 projected disc) and the
    ScatterGen_Collimator variant (ScatterGen_Collimator/scatter.cu:1212-1215, virtual_r 0.2)."""
    y1, y2, rh = -2.5, 2.5, 0.5
    hole = [3.0, y1, y2, -2.0, rh, 0.0, 0.0, 0.0, 1.0]
    S = np.array([hole[0], (y1 + y2) / 2 + F2C, hole[3]])
    I = np.zeros(3)
    a = math.radians(28.0)
    u = (S - I) / np.linalg.norm(S - I)
    # detector at 28 degrees from the incoming direction, in the x-y plane through S
    perp = np.array([u[1], -u[0], 0.0])
    perp /= np.linalg.norm(perp)
    Qw = S + 15.0 * (math.cos(a) * u + math.sin(a) * perp)
    r = crystal(Qw[0], Qw[1] - F2C, Qw[2])
    g = make_geo([r], holes=[hole])
    f, info = site_factor(I, S, r)
    assert abs(info["theta"] - a) < 1e-5
    L = float(np.linalg.norm(S - I))
    mu_t, mu_c = g.col[15], g.col[17]
    solid0 = (math.pi * 0.1 ** 2 + (y2 - y1) * 0.1) / (4 * math.pi * L * L) * abs(S[1] - I[1]) / L
    solid1 = (0.2 ** 2 + (y2 - y1) * 0.2) / (4 * L * L)
    bits = oracle.bitmap_collimator(g)
    for variant, solid in ((0, solid0), (1, solid1)):
        got = float(oracle.scatter_collimator(g, bits, det_range=(0, 1), precision="f64", variant=variant)[0, 0])
        want = solid * float(mu_c) / float(mu_t) * f
        assert want > 0
        assert abs(got - want) <= REL * want, (variant, got, want)


def mirror_geometry():
    """Two layers of 3 x 2 crystals and a 3 x 3 image, symmetric under x -> -x."""
    cr = []
    for ly, y in enumerate((12.0, 16.0)):
        for z in (-1.0, 1.0):
            for x in (-2.0, 0.0, 2.0):
                cr.append(crystal(x, y, z, w=1.8, t=2.5, h=1.8))
    g = make_geo(cr)
    g.img[0:3] = (3, 3, 2)
    g.img[3:6] = (1.5, 1.0, 1.0)
    return g


def test_mirror_symmetry():
    g = mirror_geometry()
    N, V = g.n_det, g.n_vox
    X, Y, Z = g.shape
    c = g.crystals
    mir_det = np.array([int(np.nonzero((np.abs(c[:, 0] + c[k, 0]) < 1e-6) & (c[:, 1] == c[k, 1]) &
                                       (c[:, 2] == c[k, 2]))[0][0]) for k in range(N)])
    vz, vy, vx = np.unravel_index(np.arange(V), (Z, Y, X))
    mir_vox = np.ravel_multi_index((vz, vy, X - 1 - vx), (Z, Y, X))
    rng = np.random.default_rng(4)
    p = rng.uniform(1e-4, 1e-3, (N, V))
    p = 0.5 * (p + p[mir_det][:, mir_vox])  # a mirror-symmetric PE matrix
    bits = oracle.bitmap_crystal(g)
    m = oracle.scatter_crystal(g, p.astype(np.float32), bits, precision="f64").astype(np.float64)
    assert m.max() > 0
    mm = m[mir_det][:, mir_vox]
    assert np.allclose(m, mm, rtol=1e-6, atol=0), np.abs(m - mm).max() / m.max()
    # a mistake that breaks the symmetry (e.g. a sign of x in the frame transform) would show:
    # the matrix is not symmetric under the identity-on-detectors map
    assert not np.allclose(m, m[:, mir_vox], rtol=1e-3)


def brute_crystal_bits(det):
    n = int(det[0])
    c = np.asarray(det[1:1 + 12 * n], np.float64).reshape(n, 12)
    P = c[:, 0:3]
    R = np.sqrt(c[:, 3] ** 2 + c[:, 4] ** 2 + c[:, 5] ** 2) / 2
    bits, margin = {}, {}
    for i in range(n):
        for j in range(i + 1, n):
            Lij = np.linalg.norm(P[i] - P[j])
            uji, uij = (P[j] - P[i]) / Lij, (P[i] - P[j]) / Lij
            for k in range(n):
                if k in (i, j):
                    continue
                Lik, Ljk = np.linalg.norm(P[k] - P[i]), np.linalg.norm(P[k] - P[j])
                d1 = np.linalg.norm((P[k] - P[i]) / Lik - uji) - (R[k] / Lik + R[j] / Lij)
                d2 = np.linalg.norm((P[k] - P[j]) / Ljk - uij) - (R[k] / Ljk + R[i] / Lij)
                bits[(i, j, k)] = d1 <= 0 or d2 <= 0
                margin[(i, j, k)] = min(abs(d1), abs(d2))
    return bits, margin


def test_triangular_pair_index_enumerates_the_upper_triangle():
    n = 9
    idx = [i * n - i * (i + 1) // 2 + j for i in range(n) for j in range(i, n)]
    assert idx == list(range(n * (n + 1) // 2))
    assert int(oracle.lib().or_crystal_bitmap_words(n)) == (n * (n + 1) // 2 * n + 31) // 32


def test_crystal_bitmap_against_brute_force():
    rng = np.random.default_rng(11)
    cr = [crystal(float(x), float(y), float(z), w=1.2, t=2.0, h=1.2)
          for x, y, z in zip(rng.uniform(-8, 8, 14), rng.choice([10.0, 14.0, 18.0], 14), rng.uniform(-8, 8, 14))]
    # three collinear crystals and one far off the line
    cr += [crystal(0.0, 30.0, 0.0), crystal(0.0, 36.0, 0.0), crystal(0.0, 42.0, 0.0), crystal(60.0, 36.0, 0.0)]
    g = make_geo(cr)
    n = g.n_det
    words = oracle.bitmap_crystal(g)

    def bit(i, j, k):
        b = (i * n - i * (i + 1) // 2 + j) * n + k
        return bool((int(words[b >> 5]) >> (b & 31)) & 1)
    want, margin = brute_crystal_bits(g.det)
    near = 0
    for (i, j, k), v in want.items():
        if margin[(i, j, k)] < 1e-5:
            near += 1
            continue
        assert bit(i, j, k) == v, (i, j, k)
    assert near <= 2
    a, b, c, far = n - 4, n - 3, n - 2, n - 1
    assert bit(a, c, b)            # the middle crystal lies between the outer two
    assert not bit(a, c, far)      # the far crystal does not
    for i in range(n):             # no bit for k = i or k = j, nor for the diagonal pairs
        for j in range(i, n):
            assert not bit(i, j, i) and not bit(i, j, j)
    assert sum(want.values()) > 0 and sum(not v for v in want.values()) > 0


def test_collimator_bitmap_against_brute_force():
    rng = np.random.default_rng(12)
    holes = [[float(x), -2.5, 2.5, float(z), 0.5, 0.0, 0.0, 0.0, 1.0] for x, z in rng.uniform(-10, 10, (5, 2))]
    cr = [crystal(float(x), float(y), float(z), w=1.2, t=2.0, h=1.2)
          for x, y, z in zip(rng.uniform(-8, 8, 12), rng.choice([10.0, 15.0], 12), rng.uniform(-8, 8, 12))]
    g = make_geo(cr, holes=holes)
    n, nh = g.n_det, len(holes)
    words = oracle.bitmap_collimator(g)
    c = g.crystals.astype(np.float64)
    Pk = c[:, 0:3]
    R = np.sqrt(c[:, 3] ** 2 + c[:, 4] ** 2 + c[:, 5] ** 2) / 2
    near = seen = 0
    for h, hr in enumerate(holes):
        S = np.array([hr[0], (hr[1] + hr[2]) / 2, hr[3]])
        for j in range(n):
            Lj = np.linalg.norm(Pk[j] - S)
            for k in range(n):
                Lk = np.linalg.norm(Pk[k] - S)
                d = np.linalg.norm((Pk[k] - S) / Lk - (Pk[j] - S) / Lj) - (R[k] / Lk + R[j] / Lj)
                b = h * n * n + j * n + k
                got = bool((int(words[b >> 5]) >> (b & 31)) & 1)
                if abs(d) < 1e-5:
                    near += 1
                    continue
                seen += got
                assert got == (d <= 0), (h, j, k)
    assert near <= 2 and 0 < seen < nh * n * n


def test_voxel_sample_and_inline_criterion_equal_the_full_runs():
    """The oracle's voxel-sample / bitmap-free forms (used at sizes whose bitmaps do not fit)
    give exactly the full runs' numbers: same lists, same per-element arithmetic."""
    g = params.tiny(seed=6, n_det_side=3, n_layers=2, image=(4, 3, 3), holes_side=2)
    N, V = g.n_det, g.n_vox
    pe = oracle.pe(g, sub=(2, 2, 2), precision="f64")
    bc, bh = oracle.bitmap_crystal(g), oracle.bitmap_collimator(g)
    vox = np.array([0, 7, 13, 35, 22], np.int32)
    full_c = oracle.scatter_crystal(g, pe, bc, precision="f64")
    full_h = oracle.scatter_collimator(g, bh, precision="f64")
    assert full_c.max() > 0 and full_h.max() > 0
    assert np.array_equal(oracle.scatter_crystal(g, pe, None, precision="f64"), full_c)
    assert np.array_equal(oracle.scatter_collimator(g, None, precision="f64"), full_h)
    assert np.array_equal(oracle.scatter_crystal(g, pe[:, vox], None, precision="f64", vox=vox), full_c[:, vox])
    assert np.array_equal(oracle.scatter_collimator(g, None, precision="f64", vox=vox, det_range=(3, 9)),
                          full_h[3:9][:, vox])
    pairs = [(0, 5), (2, 17), (4, 9), (16, 17)]
    got = oracle.crystal_bits_for_pairs(g, pairs)
    for p, (i, j) in enumerate(pairs):
        for k in range(N):
            b = (i * N - i * (i + 1) // 2 + j) * N + k
            assert got[p, k] == (int(bc[b >> 5]) >> (b & 31)) & 1
