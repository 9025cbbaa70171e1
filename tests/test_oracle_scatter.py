"""Pins the scatter physics helpers of the oracle (scatter.cu:961-1038) against
closed forms: the Klein-Nishina total cross-section (the 0.01-rad Riemann sum
over [0, pi] must approach it), the Compton energy, the Gaussian window
against math.erf and known fractions, and cone angles of simple geometries."""
import math

import numpy as np
import pytest

import oracle


def kn_total_closed_form(e0):
    """integral_0^pi f(theta) sin(theta) d theta for the unnormalised KN shape f = 2/r_e^2 dsigma/dOmega:
    sigma_KN / (pi r_e^2) = 2 {(1+a)/a^2 [2(1+a)/(1+2a) - ln(1+2a)/a] + ln(1+2a)/(2a) - (1+3a)/(1+2a)^2}."""
    a = e0 / 511.0
    l = math.log(1 + 2 * a)
    return 2 * ((1 + a) / a ** 2 * (2 * (1 + a) / (1 + 2 * a) - l / a) + l / (2 * a) - (1 + 3 * a) / (1 + 2 * a) ** 2)


@pytest.mark.parametrize("e0", [30.0, 140.5, 364.0, 511.0])
def test_kn_riemann_sum_approaches_total_cross_section(e0):
    L = oracle.lib()
    s = L.or_kn_riemann_f64(e0, 0.0, math.pi)
    exact = kn_total_closed_form(e0)
    assert abs(s - exact) / exact < 2e-4, (s, exact)
    # Thomson limit: 8/3 at low energy
    if e0 == 30.0:
        assert abs(exact - 8.0 / 3.0) < 0.35
    s32 = L.or_kn_riemann_f32(e0, 0.0, math.pi)
    assert abs(s32 - s) / s < 1e-4


def test_kn_partial_sums_add_up():
    L = oracle.lib()
    e0 = 140.5
    whole = L.or_kn_riemann_f64(e0, 0.0, 1.0)
    assert whole > 0
    # the sum is inclusive of both ends: a zero-width range is one sample of f(theta) sin(theta) h
    one = L.or_kn_riemann_f64(e0, 0.5, 0.5)
    c = math.cos(0.5)
    a = e0 / 511 * (1 - c)
    f = (1 + c * c) / (1 + a) ** 2 * (1 + a * a / ((1 + c * c) * (1 + a)))
    assert abs(one - f * math.sin(0.5) * float(np.float32(0.01))) < 1e-12


def test_compton_energy_closed_form():
    L = oracle.lib()
    assert abs(L.or_scatter_energy_f32(math.pi, 140.0) - 90.44) < 0.01  # 140 / (1 + 2 * 140 / 511)
    assert abs(L.or_scatter_energy_f32(0.0, 140.0) - 140.0) < 1e-4
    th = np.linspace(0, math.pi, 50)
    e = [L.or_scatter_energy_f32(float(t), 140.5) for t in th]
    assert all(a >= b for a, b in zip(e, e[1:]))
    assert abs(e[-1] - 140.5 / (1 + 2 * 140.5 / 511.0)) < 1e-3  # 90.65 keV back-scatter edge


def test_window_probability_matches_erf():
    L = oracle.lib()
    rng = np.random.default_rng(0)
    for _ in range(50):
        e = float(rng.uniform(80, 160))
        res = float(rng.uniform(0.03, 0.2))
        lo, hi = sorted(rng.uniform(60, 180, 2))
        s = res / 2.35482 * e * math.sqrt(2)
        want = 0.5 * (math.erf((hi - e) / s) - math.erf((lo - e) / s))
        assert abs(L.or_window_probability_f32(e, res, float(lo), float(hi)) - want) < 2e-6
    # wide window -> 1; +-FWHM/2 around the peak -> erf(sqrt(ln 2)) = 0.7610
    assert abs(L.or_window_probability_f32(140.0, 0.1, 0.0, 1000.0) - 1.0) < 1e-6
    fwhm = 0.1 * 140.0
    assert abs(L.or_window_probability_f32(140.0, 0.1, 140.0 - fwhm / 2, 140.0 + fwhm / 2) -
               math.erf(math.sqrt(math.log(2)))) < 1e-4


def test_cone_angle_simple_geometries():
    L = oracle.lib()

    def ang(s, a, d):
        return L.or_cone_angle_f32(np.array(s + a + d, np.float32))
    assert abs(ang([0, 0, 0], [0, 0, 1], [0, 0, 2])) < 1e-3
    assert abs(ang([0, 0, 0], [0, 0, 1], [1, 0, 1]) - math.pi / 2) < 1e-6
    assert abs(ang([0, 0, 0], [0, 0, 1], [0, 0, 0.5]) - math.pi) < 1e-3
    assert abs(ang([0, 0, 0], [0, 0, 1], [1, 0, 2]) - math.pi / 4) < 1e-6
