"""Writes tests/golden/flops_per_element.json: algorithmic FLOPs per matrix element.

Counting rule (op-counting build of the FP32 restatement of photodetectorCudaMe,
oracle/pe_oracle.cpp instantiated with Counted<float>): every + - * / sqrt exp
counts 1; comparisons, fabs, min/max, conversions and the sin/cos of fixed
rotation angles count 0; the per-element culling prologue
(PESysMatGen.cu:504-590) is excluded.  Counted exhaustively over seeded
samples of (detector row, voxel plane) pairs.

Run: python tests/golden/make_flops.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "gpu-spect-pet-sysmat_amd")]

import oracle  # noqa: E402
from sysmat import params  # noqa: E402


def sample(geo, nv, sub, n_rows, seed):
    rng = np.random.default_rng(seed)
    rows = rng.choice(geo.n_det, n_rows, replace=False)
    planes = rng.integers(0, geo.shape[2], n_rows)
    X, Y, _ = geo.shape
    flops = cross_det = cross_hole = 0.0
    n = 0
    for r, z in zip(rows, planes):
        _, st = oracle.pe_count(geo, nv=nv, sub=sub, det_range=(int(r), int(r) + 1), z0=int(z), slab_z=1)
        flops += st[0, 0]
        cross_det += st[0, 1]
        cross_hole += st[0, 2]
        n += X * Y
    return {"flops_per_element": flops / n, "culled_crystals_per_element": cross_det / n,
            "culled_holes_per_element": cross_hole / n, "elements_counted": n,
            "sample": f"{n_rows} seeded detector rows x one voxel plane each (seed {seed})"}


def main():
    out = {}
    c1 = params.config1()
    out["c1_pe_ref_sampling"] = sample(c1, 1, (8, 16, 8), 8, 1)
    out["c1_pe_nv4_sub222"] = sample(c1, 4, (2, 2, 2), 8, 2)
    c2 = params.config2()
    out["c2_pe_nv3_sub222_detail"] = sample(c2, 3, (2, 2, 2), 24, 3)
    out["c2_pe_nv3_sub222"] = out["c2_pe_nv3_sub222_detail"]["flops_per_element"]
    with open(os.path.join(HERE, "flops_per_element.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
