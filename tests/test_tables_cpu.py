"""Host-side table building (csrc/tables.hpp, plain C++) compiled with g++ and run on the CPU:
the parameter-array guards reject bad headers before any crystal or hole record is read."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sysmat import params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpu-spect-pet-sysmat_amd", "csrc")

HARNESS = r"""
#include <cstdio>
#include <vector>
#include "tables.hpp"
// argv: col.f32 det.f32 -> prints "ok n_det n_layers n_holes" or "err <message>"
static std::vector<float> load(const char* path) {
    std::vector<float> v;
    FILE* f = std::fopen(path, "rb");
    float x;
    while (f && std::fread(&x, sizeof x, 1, f) == 1) v.push_back(x);
    if (f) std::fclose(f);
    return v;
}
int main(int argc, char** argv) {
    std::vector<float> col = load(argv[1]), det = load(argv[2]);
    sysmat::Tables T;
    const std::string e = sysmat::build_tables(col.data(), det.data(), 25.0f, T);
    if (!e.empty()) std::printf("err %s\n", e.c_str());
    else std::printf("ok %d %d %d\n", T.n_det, T.n_layers, T.n_holes);
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("tables")
    src = d / "harness.cpp"
    src.write_text(HARNESS)
    exe = d / "harness"
    subprocess.run([gxx, "-O1", "-std=c++17", "-fsanitize=address", "-I", CSRC, str(src), "-o", str(exe)],
                   check=True)
    return d, exe


def run(harness, col, det):
    d, exe = harness
    np.asarray(col, np.float32).tofile(d / "col.f32")
    np.asarray(det, np.float32).tofile(d / "det.f32")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    out = subprocess.run([str(exe), str(d / "col.f32"), str(d / "det.f32")], capture_output=True, text=True,
                         env=env, check=True)
    return out.stdout.strip()


def test_valid_geometry_counts(harness):
    g = params.tiny(seed=1, n_det_side=3, n_layers=2, image=(2, 2, 2), holes_side=2, n_plates=2)
    assert run(harness, g.col, g.det) == f"ok {g.n_det} 2 8"


def test_more_than_65535_crystals_rejected_before_reading_them(harness):
    g = params.tiny(seed=1, n_det_side=2)
    det = np.zeros(13, np.float32)      # header claims 65536 crystals, array holds one record
    det[0] = 65536
    assert run(harness, g.col, det) == "err more than 65535 detector crystals"


def test_more_than_65535_holes_rejected_before_reading_them(harness):
    g = params.tiny(seed=1, n_det_side=2, n_plates=2)
    col = np.array(g.col[:100], np.float32)   # no hole records at all
    col[10], col[20] = 40000, 30000
    assert run(harness, col, g.det) == "err more than 65535 collimator holes"
    col[20] = -1
    assert run(harness, col, g.det) == "err negative hole count"
