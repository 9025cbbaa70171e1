"""JSON -> Param_*.dat conversion is byte-exact against the reference's own
converted fixture (PE/Param_*.dat = json_to_dat.py applied to
PE/geometry.annotated.json), and malformed documents raise."""
import copy
import json
import os

import numpy as np
import pytest

from sysmat import paramjson

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "param_json")


def load_doc():
    with open(os.path.join(HERE, "geometry.annotated.json")) as f:
        return json.load(f)


def test_byte_exact_against_reference_fixture(tmp_path):
    paths = paramjson.write_param_files(os.path.join(HERE, "geometry.annotated.json"), str(tmp_path))
    for name, p in paths.items():
        want = open(os.path.join(HERE, f"Param_{name}.dat"), "rb").read()
        got = open(p, "rb").read()
        assert got == want, name
    col, det, img, phy = paramjson.param_arrays(load_doc())
    assert (col.size, det.size, img.size, phy.size) == (136, 13, 12, 10)


def test_geometry_from_json_reads_like_the_drivers():
    g = paramjson.geometry_from_json(load_doc())
    assert g.col.size >= 80000 and g.det.size >= 80000 and g.img.size >= 100
    assert g.n_det == 1 and g.n_layers == int(load_doc()["collimator"]["layers"].__len__())


@pytest.mark.parametrize("mutate, msg", [
    (lambda d: d["image"].pop("num_voxel_x"), "num_voxel_x"),
    (lambda d: d["detector"]["bins"][0].__setitem__("cx", "left"), "cx"),
    (lambda d: d["collimator"]["layers"][0].__setitem__("hole_count", 99), "hole_count"),
    (lambda d: d["detector"].__setitem__("bins", []), "bins"),
])
def test_malformed_documents_raise(mutate, msg):
    d = copy.deepcopy(load_doc())
    mutate(d)
    with pytest.raises(ValueError, match=msg):
        paramjson.param_arrays(d)
