"""Benchmark: system-matrix elements/s on one MI355X.

Headline (``value``): the realistic photoelectric system matrix -- 64^3 voxels
of 0.5 mm, two 5 mm tungsten plates with 8x8 holes each, 4 layers of 64x64
crystals (16384 detectors), 3^3 voxel sub-points x 2^3 crystal sub-cells per
element, output kept as CSR (elements >= 1e-9 of the row maximum).  One step
= one block of detector rows x all 262144 voxels: the PE tracer into an
HBM-resident dense block, then the CSR count/scan/compaction kernels; the
geometry is resident before timing starts.

Extra objects on the same JSON line:
  roofline      the tracer kernel against the FP32 vector peak (algorithmic
                FLOPs per element from the op-counting restatement, counting
                the attenuators that touch the element), HBM traffic per
                launch and executed FLOP/s from the committed rocprofv3 summary;
  cpu_baseline  the FP32 restatement of the reference kernel (oracle/, g++
                -O3 OpenMP) on this host's cores, bounded sample;
  pe_scatter    PE + primary Compton scatter, plain exact kernels vs the
                learned surrogate (exact near field + MLP far field +
                exact fallback), on the collimator-scatter geometry (fully
                measured) and on the surrogate geometry (N = 16384, 128^3 x
                0.4 mm image; PE, collimator scatter and training over every
                row, the exact crystal scatter and the surrogate apply timed on
                the same seeded row sample and projected), with the MLP
                kernel's MFMA roofline; and the batched-variant geometry (6
                collimator variants, crystal + collimator scatter of all of them
                in one captured hipGraph, per-variant rates and FP64 errors).
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gpu-spect-pet-sysmat_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak (dense)
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X published dense BF16 matrix peak
HBM_PEAK_GBS = 8000.0
METRIC = "sysmat elements/s (PE, PE+scatter) plain vs surrogate @1% err, 1 MI355X"


def load_flops(key):
    path = os.path.join(ROOT, "tests", "golden", "flops_per_element.json")
    try:
        with open(path) as f:
            return json.load(f)[key]
    except (OSError, KeyError, ValueError):
        return None


def load_profile_figure(key, name="pe_tile_kernel_c2.md"):
    """A ``key: value`` figure from a committed rocprofv3 summary (None when absent)."""
    try:
        text = open(os.path.join(ROOT, "profiles", name)).read()
    except OSError:
        return None
    m = re.search(re.escape(key) + r":\s*([0-9.eE+-]+)", text)
    return float(m.group(1)) if m else None


def load_traffic(name="pe_tile_kernel_c2.md"):
    """HBM bytes per launch (FETCH + WRITE, calibrated) from the committed rocprofv3 summary."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    r = re.search(r"rows_per_launch:\s*([0-9]+)", text)
    if not m:
        return None
    return float(m.group(1)), int(r.group(1)) if r else None


def stratified_rows(geo, per_layer, seed=0):
    """``per_layer`` seeded detector rows from each detector layer (crystals of equal
    centre y), sorted."""
    import numpy as np
    y = np.asarray(geo.det[1:1 + 12 * geo.n_det], np.float64).reshape(geo.n_det, 12)[:, 1]
    rng = np.random.default_rng(seed)
    rows = []
    for ly in np.unique(y):
        idx = np.nonzero(y == ly)[0]
        rows += sorted(int(r) for r in rng.choice(idx, min(per_layer, idx.size), replace=False))
    return rows


def cpu_baseline(geo, nv, sub, rows, plane, n_planes=10, cores=None):
    """FP32 restatement of the reference kernel (oracle/) timed on a bounded sample: the
    listed detector rows (one per host thread at a time, each single-threaded) x n_planes
    voxel planes."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    cores = min(cores or os.cpu_count() or 1, len(rows))
    X, Y, _ = geo.shape
    oracle.lib()

    def one(r):
        oracle.pe(geo, nv=nv, sub=sub, det_range=(r, r + 1), z0=plane, slab_z=n_planes, precision="f32", threads=1)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(one, rows))
    dt = time.perf_counter() - t0
    n = len(rows) * X * Y * n_planes
    return {"value": round(n / dt, 1), "unit": "elements/s", "cores": cores, "kind": "port",
            "sample": f"detector rows {rows} (the same number from each detector layer, seeded) x voxel planes "
                      f"z={plane}..{plane + n_planes - 1} ({n} elements), FP32 restatement of photodetectorCudaMe "
                      f"(g++ -O3; {cores} host threads, one row each), {dt:.1f} s"}


def sync_t():
    import torch
    torch.cuda.synchronize()
    return time.perf_counter()


def pe_full(geo, nv, sub, dev, block=2048, log=None):
    """Every PE row of ``geo`` into one HBM-resident (N, V) tensor; with ``log`` each row
    block is synchronised and reported (long runs)."""
    import torch
    from sysmat.pe import PePlan
    N, V = geo.n_det, geo.n_vox
    pe = torch.empty((N, V), dtype=torch.float32, device=dev)
    plan = PePlan(geo, voxel_sub=nv, sub=sub)
    t0 = sync_t()
    for d0 in range(0, N, block):
        plan.launch(d0, min(block, N - d0), pe[d0:])
        if log is not None:
            plan.sync()
            log(f"  PE rows {d0 + min(block, N - d0)}/{N} {sync_t() - t0:.1f}s")
    plan.sync()
    t = sync_t() - t0
    plan.close()
    return pe, t


def scatter_c3(dev, log):
    """Collimator-scatter geometry (2048 detectors, 32^3 voxels), everything measured."""
    import torch
    from sysmat import params
    from sysmat.scatter import ScatterPlan
    from sysmat.surrogate import ScatterSurrogate, TrainConfig
    g = params.config3()
    N, V = g.n_det, g.n_vox
    pe, t_pe = pe_full(g, 1, (8, 16, 8), dev)
    log(f"C3 PE {t_pe:.2f}s")
    plan = ScatterPlan(g)
    cry = torch.empty((N, V), dtype=torch.float32, device=dev)
    col = torch.empty_like(cry)
    t0 = sync_t()
    plan.crystal(pe, 0, N, cry)
    t_cry = sync_t() - t0
    t0 = sync_t()
    plan.collimator(0, N, col)
    t_col = sync_t() - t0
    log(f"C3 exact crystal {t_cry:.3f}s collimator {t_col:.3f}s")
    # plate-cell collimator model (0.5 mm cells within 1 mm of a hole wall) on a centre row block
    cell_rows, cell_h, cell_keep = 512, 0.5, 1.0
    c0 = (N - cell_rows) // 2
    cel = torch.empty((cell_rows, V), dtype=torch.float32, device=dev)
    t0 = sync_t()
    plan.collimator_cells(c0, cell_rows, cel, cell_size=cell_h, keep_dist=cell_keep)
    t_cel = sync_t() - t0
    n_cells = len(plan.cell_table(cell_h, cell_keep))
    cells = {"cell_size_mm": cell_h, "keep_dist_mm": cell_keep, "cells": n_cells, "rows": cell_rows,
             "s": round(t_cel, 3), "elements_per_s": round(cell_rows * V / t_cel, 1),
             "site_terms_per_s": float(f"{cell_rows * V * n_cells / t_cel:.4g}"),
             "projected_full_s": round(t_cel * N / cell_rows, 2),
             "sum_ratio_to_hole_sites": float(f"{(cel.double().sum() / col[c0:c0 + cell_rows].double().sum()).item():.4g}")}
    log(f"C3 collimator cells {t_cel:.3f}s for {cell_rows} rows x {n_cells} cells")
    del cel
    sur = ScatterSurrogate(g)
    t0 = sync_t()
    tr = sur.train(pe, TrainConfig(seed=0))
    t_train = sync_t() - t0
    out = torch.empty_like(cry)
    t0 = sync_t()
    n_fb = sur.apply(pe, 0, N, out)
    t_apply = sync_t() - t0
    e, s = cry.double(), out.double()
    full_exact = (pe.double().sum() + col.double().sum() + e.sum()).item()
    tot_err = ((s.sum() - e.sum()) / e.sum()).item()
    full_err = ((s.sum() - e.sum()) / full_exact)
    l1 = ((s - e).abs().sum() / e.sum()).item()
    judge = sur.judge(pe)
    elems = N * V
    plain_t = t_pe + t_cry + t_col
    sur_t = t_pe + t_train + t_apply + t_col
    res = {
        "workload": "collimator-scatter geometry: 32^3 x 1 mm, 1 W plate 6x6 holes, 2 x 32x32 CdTe (2048 det), "
                    "PE at reference sampling (1, 8x16x8)",
        "elements": elems,
        "plain": {"elements_per_s": round(elems / plain_t, 1), "pe_s": round(t_pe, 3),
                  "crystal_scatter_s": round(t_cry, 3), "collimator_scatter_s": round(t_col, 3)},
        "surrogate": {"elements_per_s": round(elems / sur_t, 1), "train_s": round(t_train, 3),
                      "apply_s": round(t_apply, 3), "fallback_tiles": n_fb,
                      "fallback_frac": round(n_fb / (N * sur.n_tiles), 4),
                      "train_pairs": tr["pairs"]},
        "collimator_cells": cells,
        "crystal_scatter_speedup_excl_training": round(t_cry / t_apply, 2),
        "crystal_scatter_speedup_incl_training": round(t_cry / (t_apply + t_train), 2),
        "crystal_total_sum_rel_err": round(tot_err, 6),
        "system_matrix_total_sum_rel_err": round(float(full_err), 7),
        "crystal_l1_rel_err": round(l1, 5),
        "judge_heldout": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in judge.items()},
    }
    plan.close()
    sur.close()
    del pe, cry, col, out
    torch.cuda.empty_cache()
    return res


def scatter_c4(dev, log, sample_rows=256, check_rows=2, check_vox=48):
    """Batched variants: 6 collimator variants (seed 0) sharing a 3 x 48x48 detector and a
    48^3 x 0.75 mm image.  Per variant the PE matrix (1, 2x2x2); then crystal + collimator
    scatter of every variant and every row in ONE captured hipGraph (timed: capture + run);
    per-variant scatter rates from a seeded row sample of each variant's plan (eager,
    projected to all rows); per-variant error against the FP64 oracle on sampled rows x
    voxels of the graph's own output."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import torch
    import oracle
    from sysmat import params
    from sysmat.pe import PePlan
    from sysmat.scatter import ScatterBatch, ScatterPlan
    vs = params.config4_variants(seed=0)
    N, V = vs[0].n_det, vs[0].n_vox
    pes, pe_t = [], []
    for g in vs:
        pe = torch.empty((N, V), dtype=torch.float32, device=dev)
        plan = PePlan(g, voxel_sub=1, sub=(2, 2, 2))
        t0 = sync_t()
        plan.launch(0, N, pe)
        plan.sync()
        pe_t.append(sync_t() - t0)
        plan.close()
        pes.append(pe)
    log(f"C4 PE per variant {[round(t, 2) for t in pe_t]}")
    outs = [torch.empty((N, V), dtype=torch.float32, device=dev) for _ in vs]
    b = ScatterBatch(vs)
    t0 = sync_t()
    b.run(pes, outs, use_graph=True)
    t_all = sync_t() - t0
    nodes = b.graph_nodes
    b.close()
    log(f"C4 all variants in one captured graph ({nodes} nodes): {t_all:.2f}s")
    rng = np.random.default_rng(7)
    rows = np.sort(rng.choice(N, sample_rows, replace=False)).astype(np.int32)
    vox = np.sort(rng.choice(V, check_vox, replace=False)).astype(np.int32)
    vt = torch.from_numpy(vox).long().to(dev)
    per, checks = [], []
    for i, g in enumerate(vs):
        sp = ScatterPlan(g, bitmaps=False)
        cry = torch.empty((sample_rows, V), dtype=torch.float32, device=dev)
        col = torch.empty((sample_rows, V), dtype=torch.float32, device=dev)
        t0 = sync_t()
        sp.crystal_rows(pes[i], rows, cry)
        for j, r in enumerate(rows):
            sp.collimator(int(r), 1, col[j:j + 1])
        t = sync_t() - t0
        sp.close()
        per.append(t * N / sample_rows)
        crow = [int(rows[0]), int(rows[-1])][:check_rows]
        checks.append((g, crow, pes[i].index_select(1, vt).cpu().numpy(),
                       outs[i][torch.as_tensor(crow, device=dev).long()].index_select(1, vt).cpu().numpy()))
        del cry, col
    log(f"C4 per-variant scatter (projected from {sample_rows} rows): {[round(t, 1) for t in per]}")

    def ref(args):
        g, r, pe_cols = args
        c = oracle.scatter_crystal(g, pe_cols, None, det_range=(r, r + 1), precision="f64", vox=vox, threads=1)
        h = oracle.scatter_collimator(g, None, det_range=(r, r + 1), precision="f64", vox=vox, threads=1)
        return (c + h.astype(np.float64))[0]
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4)) as ex:
        refs = list(ex.map(ref, [(g, r, pc) for g, crow, pc, _ in checks for r in crow]))
    errs, k = [], 0
    for g, crow, _, got in checks:
        want = np.stack(refs[k:k + len(crow)]).astype(np.float64)
        k += len(crow)
        rowmax = want.max(axis=1, keepdims=True)
        sig = (want >= 1e-9 * rowmax) & (rowmax > 0)
        errs.append(float(np.where(sig, np.abs(got - want) / np.maximum(want, 1e-300), 0).max()))
    log(f"C4 per-variant max rel err vs FP64 oracle {[f'{e:.1e}' for e in errs]}")
    elems = len(vs) * N * V
    res = {"workload": "6 collimator variants (1-3 plates, 4x4..12x12 holes), shared 3 x 48x48 detector (6912 det), "
                       "48^3 x 0.75 mm image; PE (1, 2x2x2) + crystal + collimator scatter",
           "elements": elems, "pe_s": [round(t, 3) for t in pe_t],
           "scatter_s_all_one_graph": round(t_all, 3), "graph_nodes": nodes,
           "pe_plus_scatter_elements_per_s": round(elems / (sum(pe_t) + t_all), 1),
           "per_variant": [{"plates": g.n_layers, "holes": g.n_holes, "pe_s": round(pe_t[i], 3),
                            "scatter_s_projected": round(per[i], 2),
                            "pe_plus_scatter_elements_per_s": round(N * V / (pe_t[i] + per[i]), 1),
                            "max_rel_err_vs_fp64": float(f"{errs[i]:.3g}")} for i, g in enumerate(vs)],
           "per_variant_basis": f"eager plan on {sample_rows} seeded rows, projected; FP64 error on "
                                f"{check_rows} rows x {check_vox} voxels of the graph output"}
    del pes, outs
    torch.cuda.empty_cache()
    return res


MLP_FLOP_PER_ELEMENT = 2 * (12 * 256 + 3 * 256 * 256 + 256)  # 399,872 MFMA FLOP per element per member


def scatter_c5(dev, log, image=128, voxel=0.4, sample_rows=64, apply_all=False, apply_block=2048, tag="C5"):
    """Surrogate geometry (16384 detectors in 4 layers; image ``image``^3 voxels of ``voxel``
    mm): PE and the collimator scatter of every row, surrogate training on its own exact rows,
    the judge rows; the exact crystal scatter timed on a seeded row sample and projected
    linearly to all rows; the surrogate apply (near field + MLP + exact fallback) on every
    row (``apply_all``) or on the same sampled rows, projected likewise.  The PE matrix stays
    resident (the scatter reads the rows of every scatter site)."""
    import numpy as np
    import torch
    from sysmat import params
    from sysmat.scatter import ScatterPlan
    from sysmat.surrogate import ScatterSurrogate, TrainConfig, usable_rows
    g = params.config5()
    g.img[0:3] = image
    g.img[3:6] = voxel
    N, V = g.n_det, g.n_vox
    long_run = V * N > (1 << 32)
    pe, t_pe = pe_full(g, 1, (2, 2, 2), dev, log=log if long_run else None)
    log(f"{tag} PE {t_pe:.2f}s")
    plan = ScatterPlan(g, bitmaps=False)
    blk = torch.empty((min(apply_block, N), V), dtype=torch.float32, device=dev)
    t0 = sync_t()
    for b in range(0, N, apply_block):
        plan.collimator(b, min(apply_block, N - b), blk)
        if long_run:
            torch.cuda.synchronize()
            log(f"  collimator rows {b + min(apply_block, N - b)}/{N} {sync_t() - t0:.1f}s")
    t_col = sync_t() - t0
    rng = np.random.default_rng(123)
    rows = np.sort(rng.choice(usable_rows(g), sample_rows, replace=False))
    ex = torch.empty((sample_rows, V), dtype=torch.float32, device=dev)
    t0 = sync_t()
    plan.crystal_rows(pe, rows, ex)
    t_ex = sync_t() - t0
    log(f"{tag} exact crystal {sample_rows} rows {t_ex:.2f}s, collimator all rows {t_col:.2f}s")
    sur = ScatterSurrogate(g, use_bitmap=False)
    t0 = sync_t()
    tr = sur.train(pe, TrainConfig(seed=0))
    t_train = sync_t() - t0
    log(f"{tag} surrogate trained {t_train:.2f}s ({tr['exact_rows_s']:.2f}s exact training rows, "
        f"{tr['train_s']:.2f}s optimisation)")
    judge = sur.judge(pe)
    log(f"{tag} judge rows done")
    # surrogate apply: near field, MLP kernel and exact fallback timed separately (HIP events on the
    # launch stream)
    phase = np.zeros(3)
    n_fb = 0
    applied_rows = 0

    def run(b, cnt, out):
        nonlocal n_fb, applied_rows
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        n_fb += sur.apply(pe, b, cnt, out, events=ev)
        torch.cuda.synchronize()
        phase[:] += [ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), ev[2].elapsed_time(ev[3])]
        applied_rows += cnt

    sur_out = torch.empty((sample_rows, V), dtype=torch.float32, device=dev)
    t0 = sync_t()
    if apply_all:
        for b in range(0, N, apply_block):
            run(b, min(apply_block, N - b), blk)
            if long_run:
                log(f"  surrogate rows {b + min(apply_block, N - b)}/{N} {sync_t() - t0:.1f}s")
    else:
        for i, r in enumerate(rows):
            run(int(r), 1, sur_out[i:i + 1])
    t_ap = sync_t() - t0
    scale = N / applied_rows
    log(f"{tag} surrogate apply {applied_rows} rows {t_ap:.2f}s (near {phase[0] / 1e3:.2f}s, MLP "
        f"{phase[1] / 1e3:.2f}s, fallback {phase[2] / 1e3:.2f}s; train {t_train:.2f}s)")
    t_cry_full = t_ex / sample_rows * N
    t_sur_full = t_ap * scale + t_train
    elems = N * V
    plain_t = t_pe + t_col + t_cry_full
    sur_t = t_pe + t_col + t_sur_full
    mlp_s = phase[1] / 1e3
    mfma = applied_rows * V * len(sur.members) * MLP_FLOP_PER_ELEMENT / mlp_s / 1e12
    res = {
        "workload": f"surrogate geometry: 4 x 64x64 crystals (16384 det), 1 W plate 6x6 holes, image "
                    f"{image}^3 x {voxel} mm, PE sampling (1, 2x2x2)",
        "elements": elems,
        "projection": (f"PE, collimator scatter and training over every row; exact crystal scatter timed on "
                       f"{sample_rows} seeded rows and scaled to {N} rows; surrogate apply "
                       + (f"over every row (blocks of {apply_block})" if apply_all else
                          f"timed on the same {sample_rows} rows and scaled to {N} rows")),
        "plain": {"elements_per_s": round(elems / plain_t, 1), "pe_s": round(t_pe, 3),
                  "collimator_scatter_s": round(t_col, 3), "crystal_scatter_s_projected": round(t_cry_full, 2)},
        "surrogate": {"elements_per_s": round(elems / sur_t, 1), "train_s": round(t_train, 3),
                      "train_exact_rows_s": round(tr["exact_rows_s"], 3), "train_optimise_s": round(tr["train_s"], 3),
                      "apply_s": round(t_ap * scale, 3),
                      "apply_phase_s": {"near_field": round(phase[0] / 1e3 * scale, 3),
                                        "mlp_kernel": round(mlp_s * scale, 3),
                                        "exact_fallback": round(phase[2] / 1e3 * scale, 3)},
                      "fallback_frac": round(n_fb / (applied_rows * sur.n_tiles), 4),
                      "train_pairs": tr["pairs"]},
        "crystal_scatter_speedup_incl_training": round(t_cry_full / t_sur_full, 2),
        "pe_plus_scatter_speedup": round(plain_t / sur_t, 2),
        "judge_heldout": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in judge.items()},
        "mlp_roofline": {"kernel": "surrogate_apply_kernel", "bound": "mfma", "achieved": round(mfma, 2),
                         "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(mfma / BF16_MFMA_PEAK_TFLOPS, 4),
                         "flops_per_element": MLP_FLOP_PER_ELEMENT * len(sur.members),
                         "basis": f"{len(sur.members)} members x 2 (12x256 + 3 x 256x256 + 256) MFMA FLOP per "
                                  f"element, {applied_rows} rows x {V} voxels / the kernel's HIP-event time",
                         "kernel_s": round(mlp_s, 4)},
    }
    # total-sum error on the judge rows for the whole system matrix (PE + collimator + crystal)
    jr = sur.judge_rows
    jpe = pe[torch.as_tensor(jr, device=dev).long()].double().sum().item()
    jcol = 0.0
    for r in jr:
        plan.collimator(int(r), 1, blk)
        jcol += blk[0].double().sum().item()
    jex = torch.empty((len(jr), V), dtype=torch.float32, device=dev)
    sur.exact_rows(pe, jr, jex)
    jcry = jex.double().sum().item()
    res["system_matrix_total_sum_rel_err_judge_rows"] = round(judge["total_rel_err"] * jcry / (jpe + jcol + jcry), 7)
    plan.close()
    sur.close()
    del pe, blk, jex, ex, sur_out
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--block", type=int, default=1024, help="detector rows per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-scatter", action="store_true", help="skip the PE+scatter legs")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched collimator-variant leg (~2 minutes)")
    ap.add_argument("--c5-reduced", action="store_true",
                    help="surrogate leg on a 32^3 x 1.6 mm image (every row applied) instead of the 128^3 x 0.4 mm one")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="CPU baseline rows per detector layer (0: host threads / 4, threads <= 16)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (HBM counter calibration under rocprofv3)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="use the default 4x4x4 voxel tile")
    ap.add_argument("--tile", default="", help="voxel tile X,Y,Z for the C2 leg (no autotuning)")
    ap.add_argument("--tune-cache", default=os.environ.get("SYSMAT_TUNE_CACHE",
                                                            os.path.join(ROOT, ".tile_cache.txt")))
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; runs on several GPUs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    def log(msg):
        if not args.quiet:
            print(msg, file=sys.stderr, flush=True)

    import numpy as np
    import torch

    from sysmat import csr as csr_mod
    from sysmat import params
    from sysmat.pe import PePlan, autotune

    geo = params.config2(seed=0)
    nv, sub = 3, (2, 2, 2)
    N, V = geo.n_det, geo.n_vox
    B = min(args.block, N)
    n_blocks = N // B
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    tile, tune_info = (4, 4, 4), None
    if args.tile:
        tile = tuple(int(x) for x in args.tile.split(","))
        if len(tile) != 3 or min(tile) < 1:
            raise SystemExit("--tile takes three positive integers X,Y,Z")
    elif not args.no_autotune:
        os.makedirs(os.path.dirname(args.tune_cache) or ".", exist_ok=True)
        t0 = time.perf_counter()
        tile, ns_el, cached = autotune(geo, sample_rows=64, cache_path=args.tune_cache, voxel_sub=nv, sub=sub)
        tune_info = {"tile": list(tile), "sample_ns_per_element": round(ns_el, 3), "cached": cached,
                     "tuning_s": round(time.perf_counter() - t0, 2)}
        log(f"autotune: tile {tile} ({ns_el:.2f} ns/element on the sample, cached={cached})")
    plan = PePlan(geo, voxel_sub=nv, sub=sub, tile=tile, device=0)
    out = torch.empty((B, V), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def block_of(i):
        # timed step i takes detector block (5 i) mod n_blocks: the default 16 steps cover every
        # detector of the 4-layer C2 detector exactly once, the first 4 steps one block of each layer
        return ((i * 5) % n_blocks) * B

    def step(i, ev=None):
        if ev is not None:
            ev[0].record(stream)
        plan.launch(block_of(i), B, out, stream=stream)
        if ev is not None:
            ev[1].record(stream)
        rp, col, val, rmax = csr_mod.csr_from_dense_device(out, 1e-9, stream=stream)
        if ev is not None:
            ev[2].record(stream)
        return int(col.numel())

    for i in range(args.warmup):
        step(args.steps + i)
    plan.sync()
    torch.cuda.synchronize()

    ev = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(args.steps)]
    nnz = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        nnz.append(step(i, ev[i]))
    plan.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pe_ms = [a.elapsed_time(b) for a, b, _ in ev]
    csr_ms = [b.elapsed_time(c) for _, b, c in ev]
    elems = args.steps * B * V
    value = elems / wall
    ms_per_step = wall * 1e3 / args.steps
    avg_kernel_s = sum(pe_ms) / len(pe_ms) / 1e3
    log(f"C2 PE+CSR: {value:.3e} elements/s, PE {np.mean(pe_ms):.1f} ms, CSR {np.mean(csr_ms):.1f} ms per step")

    fpe = load_flops("c2_pe_nv3_sub222_touched")
    fpe_ref = load_flops("c2_pe_nv3_sub222")
    roof = None
    if fpe:
        achieved = fpe * B * V / avg_kernel_s / 1e12
        traffic = None
        tr = load_traffic()
        if tr and tr[1]:
            traffic = round(tr[0] * B / tr[1])  # per launch of B rows
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "flops_per_element": round(fpe, 1), "kernel": "pe_tile_kernel",
                "avg_launch_ms": round(avg_kernel_s * 1e3, 3), "rows_per_launch": B,
                "algorithmic_output_bytes_per_launch": B * V * 4,
                "flops_basis": "FP32 operations of the reference algorithm (op-counting restatement of "
                               "photodetectorCudaMe) on the culled crystals and holes that touch the element "
                               "(non-zero chord on some sub-ray); the others add exactly zero",
                "flops_per_element_reference_basis": round(fpe_ref, 1) if fpe_ref else None,
                "achieved_reference_basis": round(fpe_ref * B * V / avg_kernel_s / 1e12, 3) if fpe_ref else None,
                "non_packed_peak": FP32_PEAK_TFLOPS / 2,
                "packed_fp32": False,
                "executed_fp32_tflops": load_profile_figure("executed_fp32_tflops"),
                "executed_fp64_tflops": load_profile_figure("executed_fp64_tflops"),
                "valu_issue_frac": load_profile_figure("valu_issue_frac")}

    if args.calib_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    pe_scatter = None
    if not args.no_scatter:
        plan.close()
        del out
        torch.cuda.empty_cache()
        pe_scatter = {}
        try:
            pe_scatter["c3"] = scatter_c3(dev, log)
        except Exception as e:  # reported, not hidden
            pe_scatter["c3"] = {"error": repr(e)}
        try:
            if args.c5_reduced:
                pe_scatter["c5"] = scatter_c5(dev, log, image=32, voxel=1.6, sample_rows=512, apply_all=True,
                                              tag="C5r")
            else:
                pe_scatter["c5"] = scatter_c5(dev, log)
        except Exception as e:
            pe_scatter["c5"] = {"error": repr(e)}


    if not args.no_c4:
        if pe_scatter is None:
            plan.close()
            del out
            torch.cuda.empty_cache()
            pe_scatter = {}
        try:
            pe_scatter["c4"] = scatter_c4(dev, log)
        except Exception as e:
            pe_scatter["c4"] = {"error": repr(e)}

    cpu = None
    if not args.no_cpu:
        rng = np.random.default_rng(0)
        cores = min(16, os.cpu_count() or 8)
        per_layer = args.cpu_rows if args.cpu_rows > 0 else max(1, cores // 4)
        rows = stratified_rows(geo, per_layer, seed=0)
        n_planes = 8  # about 10-25 s of host work at the measured rates
        plane = int(rng.integers(0, geo.shape[2] - n_planes))
        cpu = cpu_baseline(geo, nv, sub, rows, plane, n_planes, cores=cores)

    rec = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulate)",
        "data": "synthetic geometry from seed 0",
        "config": {"workload": "PE realistic: 64^3 x 0.5 mm image, 2 W plates 8x8 holes, 4x64x64 crystals "
                               "(16384 det), 3^3 voxel x 2^3 crystal sub-points, CSR output >= 1e-9 row max",
                   "detectors_per_step": B, "voxels": V, "csr_nnz_per_step": int(np.mean(nnz)),
                   "pe_kernel_ms_per_step": round(float(np.mean(pe_ms)), 3),
                   "voxel_tile": list(tile), "autotune": tune_info,
                   "csr_ms_per_step": round(float(np.mean(csr_ms)), 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "pe_scatter": pe_scatter,
    }
    if cpu:
        rec["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(rec))
    if args.no_scatter and args.no_c4:
        plan.close()


if __name__ == "__main__":
    main()
