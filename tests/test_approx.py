"""The scatter site term's FP32 approximations (csrc/approx.hpp) against the FP64 library
functions: erfc_nr (Numerical Recipes erfcc with an FP64 exponent) in relative error over
z in [0, 10], and atan2_upper (Abramowitz & Stegun 4.4.49 after octant reduction) in
absolute error over the upper half-plane.  The header is plain C++ outside HIP, so it is
compiled here with g++; the device build differs only in the reciprocal (v_rcp_f32, 1 ulp)
and exp2 (v_exp_f32), covered by the margins below and by the scatter parity tests."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "gpu-spect-pet-sysmat_amd", "csrc", "approx.hpp")

SRC = r'''
#include "approx.hpp"
extern "C" float erfc_nr_c(float z) { return sysmat::erfc_nr(z); }
extern "C" float atan2_upper_c(float y, float x) { return sysmat::atan2_upper(y, x); }
'''


@pytest.fixture(scope="module")
def ap(tmp_path_factory):
    d = tmp_path_factory.mktemp("approx")
    src = d / "approx.cpp"
    src.write_text(SRC)
    so = d / "approx.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-I", os.path.dirname(HDR), str(src),
                    "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.erfc_nr_c.restype = ctypes.c_float
    L.erfc_nr_c.argtypes = [ctypes.c_float]
    L.atan2_upper_c.restype = ctypes.c_float
    L.atan2_upper_c.argtypes = [ctypes.c_float, ctypes.c_float]
    return L


def test_erfc_nr_relative_error(ap):
    rng = np.random.default_rng(1)
    zs = np.concatenate([np.linspace(0.0, 10.0, 20001), rng.uniform(0.0, 10.0, 20000),
                         [0.0, 1e-30, 1e-7, 0.5, 1.0, 3.0, 9.0, 9.5]]).astype(np.float32)
    worst = 0.0
    for z in zs:
        z = float(z)
        want = math.erfc(z)
        got = ap.erfc_nr_c(z)
        if want > 1e-30:
            worst = max(worst, abs(got - want) / want)
    assert worst < 6e-7, worst


def test_erfc_nr_tail_and_overflow(ap):
    # beyond the FP32 range the result is 0, never NaN; large arguments do not overflow the
    # integer exponent
    for z in (11.0, 20.0, 1e3, 1e19, float("inf")):
        assert ap.erfc_nr_c(z) == 0.0
    assert ap.erfc_nr_c(0.0) == pytest.approx(1.0, rel=2e-7)


def test_erfc_nr_fails_with_a_wrong_coefficient_sign():
    # the check above is sensitive: the same form with one flipped coefficient is far off
    t = 1.0 / (1.0 + 0.5)
    c = [0.17087277, -0.82215223, 1.48851587, -1.13520398, 0.27886807, -0.18628806, 0.09678418,
         0.37409196, 1.00002368, -1.26551223]
    def f(cs):
        p = cs[0]
        for k in cs[1:]:
            p = p * t + k
        return t * math.exp(-1.0 + p)
    assert abs(f(c) - math.erfc(1.0)) / math.erfc(1.0) < 2e-7
    bad = list(c)
    bad[5] = -bad[5]
    assert abs(f(bad) - math.erfc(1.0)) / math.erfc(1.0) > 1e-3


def test_atan2_upper_absolute_error(ap):
    rng = np.random.default_rng(2)
    th = np.concatenate([np.linspace(0.0, math.pi, 40001), rng.uniform(0.0, math.pi, 20000),
                         [0.0, math.pi / 4, math.pi / 2, 3 * math.pi / 4, math.pi]])
    rad = np.concatenate([np.ones(th.size - 5000), rng.uniform(1e-3, 2.0, 5000)])
    worst = 0.0
    for t, r in zip(th, rad):
        y, x = np.float32(r * math.sin(t)), np.float32(r * math.cos(t))
        y = max(float(y), 0.0)
        want = math.atan2(y, float(x))
        worst = max(worst, abs(ap.atan2_upper_c(y, float(x)) - want))
    # the step count of the KN sum decides by FP32 outside a band of 5e-6 rad around each
    # step boundary; the approximation stays well inside it
    assert worst < 5e-7, worst


def test_atan2_upper_special_points(ap):
    assert ap.atan2_upper_c(0.0, 1.0) == 0.0
    assert ap.atan2_upper_c(0.0, 0.0) == 0.0
    assert ap.atan2_upper_c(1.0, 0.0) == pytest.approx(math.pi / 2, abs=2e-7)
    assert ap.atan2_upper_c(0.0, -1.0) == pytest.approx(math.pi, abs=2e-7)
    assert ap.atan2_upper_c(1.0, -1.0) == pytest.approx(3 * math.pi / 4, abs=3e-7)
