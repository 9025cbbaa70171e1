"""Batched crystal + collimator scatter over collimator variants (shared detector
and image), launched eagerly and as one captured hipGraph.

The graph replays must equal the eager launches and the per-variant plans bit
for bit (same kernels, same order), every replay must equal the first (the
collimator term overwrites, the crystal term adds), and one small variant set is
checked against the FP64 oracle (collimator + crystal, 1e-5 relative)."""
import numpy as np
import pytest

import oracle
from sysmat import params
from sysmat import scatter as sc

pytestmark = pytest.mark.gpu


def small_variants(n=3, seed=4):
    return params.config4_variants(seed=seed, n_variants=n, image=(6, 6, 5), voxel=2.0, det_side=8, n_layers=2)


def pe_of(g):
    import torch
    return torch.from_numpy(oracle.pe(g, sub=(2, 2, 2), precision="f64")).cuda()


def test_batch_graph_equals_eager_and_plans(lib):
    import torch
    vs = small_variants()
    assert len({v.det.tobytes() for v in vs}) == 1 and len({v.col.tobytes() for v in vs}) == len(vs)
    N, V = vs[0].n_det, vs[0].n_vox
    pes = [pe_of(g) for g in vs]
    b = sc.ScatterBatch(vs)
    eager = [torch.full((N, V), 7.0, device="cuda") for _ in vs]
    b.run(pes, eager, use_graph=False)
    graph = [torch.full((N, V), -3.0, device="cuda") for _ in vs]
    b.run(pes, graph, use_graph=True)
    first = [t.clone() for t in graph]
    print("graph nodes", b.graph_nodes, flush=True)
    assert b.graph_nodes >= 4 * len(vs)
    for rep in range(3):
        torch.cuda.synchronize()
        _ = [t.sum().item() for t in graph]  # eager work between replays
        b.run(pes, graph, use_graph=True)
        for a, f in zip(graph, first):
            assert torch.equal(a, f), f"replay {rep}"
    for v, g in enumerate(vs):
        assert torch.equal(graph[v], eager[v])
        p = sc.ScatterPlan(g, bitmaps=False)
        col = torch.zeros((N, V), device="cuda")
        cry = torch.zeros((N, V), device="cuda")
        p.collimator(0, N, col)
        p.crystal(pes[v], 0, N, cry)
        torch.cuda.synchronize()
        assert torch.equal(graph[v], col + cry), f"variant {v}"
        assert float(graph[v].sum()) > 0


def test_batch_matches_fp64_oracle(lib):
    import torch
    vs = small_variants(n=2, seed=9)
    N, V = vs[0].n_det, vs[0].n_vox
    pes = [pe_of(g) for g in vs]
    b = sc.ScatterBatch(vs)
    out = [torch.zeros((N, V), device="cuda") for _ in vs]
    b.run(pes, out, use_graph=True)
    for v, g in enumerate(vs):
        pe = pes[v].cpu().numpy()
        ref = (oracle.scatter_collimator(g, oracle.bitmap_collimator(g), precision="f64") +
               oracle.scatter_crystal(g, pe, oracle.bitmap_crystal(g), precision="f64"))
        gpu = out[v].cpu().numpy().astype(np.float64)
        rowmax = ref.max(axis=1, keepdims=True)
        sig = (ref >= 1e-9 * rowmax) & (rowmax > 0)
        err = np.where(sig, np.abs(gpu - ref) / np.maximum(ref, 1e-300), 0).max()
        print(f"variant {v}: max rel err vs FP64 oracle {err:.2e}", flush=True)
        assert err <= 1e-5
