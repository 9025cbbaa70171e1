"""The scatter kernel's O(1) Klein-Nishina partial sum (csrc/kn.hpp, kn_sum_cs) against
the oracle's 0.01-rad Riemann sum it replaces (computeComptonIntegral,
scatter.cu:979-993), for start angles and step counts over the whole range.  The
header is plain C++ behind a __HIPCC__ guard, so it is compiled here with g++."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "gpu-spect-pet-sysmat_amd", "csrc", "kn.hpp")

SRC = r'''
#include "kn.hpp"
extern "C" double kn_cs(double e0, double a, int n, double h) {
    const sysmat::KnCoef c = sysmat::kn_coef(e0);
    return sysmat::kn_sum_cs<false>(c, std::cos(a), std::sin(a), std::cos((double)n * h), std::sin((double)n * h), h);
}
extern "C" double kn_cs_series(double e0, double a, int n, double h) {
    const sysmat::KnCoef c = sysmat::kn_coef(e0);
    return sysmat::kn_sum_cs<true>(c, std::cos(a), std::sin(a), std::cos((double)n * h), std::sin((double)n * h), h);
}
extern "C" double kn_log1p(double x) { return sysmat::log1p_series(x); }
extern "C" double kn_series_max_alpha() { return sysmat::kKnSeriesMaxAlpha; }
'''


@pytest.fixture(scope="module")
def kn(tmp_path_factory):
    d = tmp_path_factory.mktemp("kn")
    src = d / "kn_cs.cpp"
    src.write_text(SRC)
    so = d / "kn_cs.so"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-I", os.path.dirname(HDR), str(src), "-o", str(so)],
                   check=True)
    L = ctypes.CDLL(str(so))
    L.kn_cs.restype = ctypes.c_double
    L.kn_cs.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.kn_cs_series.restype = ctypes.c_double
    L.kn_cs_series.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.kn_log1p.restype = ctypes.c_double
    L.kn_log1p.argtypes = [ctypes.c_double]
    L.kn_series_max_alpha.restype = ctypes.c_double
    return L


def test_log1p_series_against_math_log1p(kn):
    """The series form of ln(1 + x) used for the KN antiderivative difference: within 3 eps of
    math.log1p wherever the kernel uses it -- the arguments x = w_b / w_a - 1 with w in
    [1, 1 + 2 alpha] for alpha up to the series bound, the extremes, tiny and zero arguments."""
    a = kn.kn_series_max_alpha()
    lo, hi = 1.0 / (1.0 + 2.0 * a) - 1.0, 2.0 * a
    s_max = a / (1.0 + a)
    assert s_max * s_max <= 0.1 + 1e-15  # the bound the series is accurate for
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(lo, hi, 20000), rng.uniform(-1e-3, 1e-3, 2000),
                         [0.0, 1e-300, -1e-300, 1e-17, lo, hi]])
    for x in xs:
        x = float(x)
        got, want = kn.kn_log1p(x), math.log1p(x)
        assert abs(got - want) <= 3 * 2.220446049250313e-16 * abs(want) + 1e-320, (x, got, want)


@pytest.mark.parametrize("e0,series", [(140.5, False), (364.0, False), (140.5, True), (236.0, True)])
def test_kn_sum_cs_matches_riemann(kn, e0, series):
    L = oracle.lib()
    h = float(np.float32(0.01))
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(300):
        a = float(rng.uniform(0.0, math.pi))
        n = int(rng.integers(0, int((math.pi - a) / h) + 1))
        ref = L.or_kn_riemann_f64(e0, a, a + (n + 0.5) * h)   # i = 0 .. floor(n + 0.5) = n: n + 1 terms
        got = (kn.kn_cs_series if series else kn.kn_cs)(e0, a, n, h)
        if ref > 1e-12:
            worst = max(worst, abs(got - ref) / ref)
    assert worst < 1e-8, worst


def test_kn_sum_cs_single_step_and_whole_range(kn):
    L = oracle.lib()
    h = float(np.float32(0.01))
    e0 = 140.5
    # n = 0: one sample f(a) sin(a) h
    for a in (0.3, 1.7, 3.0):
        assert kn.kn_cs(e0, a, 0, h) == pytest.approx(L.or_kn_riemann_f64(e0, a, a), rel=1e-9)
    # the normalisation range [0, pi]: 314 steps
    n = int(float(np.float32(math.pi)) / h)
    assert kn.kn_cs(e0, 0.0, n, h) == pytest.approx(L.or_kn_riemann_f64(e0, 0.0, (n + 0.5) * h), rel=1e-9)
