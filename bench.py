"""Benchmark: PE system-matrix elements/s on one MI355X.

Workload: the realistic photoelectric system matrix (64^3 voxels of 0.5 mm, two
5 mm tungsten plates with 8x8 holes each, 4 layers of 64x64 crystals =
16384 detectors; 3^3 voxel sub-points x 2^3 crystal sub-cells = 216 sub-rays
per element).  One step = one block of detector rows x all 262144 voxels,
traced into an HBM-resident dense block; consecutive steps walk detector
blocks spread over all four crystal layers.

Prints one JSON line (metric, value = elements/s, roofline of the tracer
kernel against the FP32 vector peak, CPU baseline of the FP32 restatement of
the reference kernel on this host's cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gpu-spect-pet-sysmat_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector/matrix peak (dense)
HBM_PEAK_GBS = 8000.0


def load_flops(key):
    path = os.path.join(ROOT, "tests", "golden", "flops_per_element.json")
    try:
        with open(path) as f:
            return json.load(f)[key]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(geo, nv, sub, row0, n_rows, plane, n_planes=4):
    """FP32 restatement of the reference kernel (oracle/, OpenMP over detector rows)
    timed on a bounded sample: n_rows consecutive detectors x n_planes voxel planes."""
    import oracle
    threads = min(os.cpu_count() or 1, n_rows)
    X, Y, _ = geo.shape
    oracle.lib()
    t0 = time.perf_counter()
    oracle.pe(geo, nv=nv, sub=sub, det_range=(row0, row0 + n_rows), z0=plane, slab_z=n_planes, precision="f32",
              threads=threads)
    dt = time.perf_counter() - t0
    n = n_rows * X * Y * n_planes
    return {"value": round(n / dt, 1), "unit": "elements/s", "cores": threads, "kind": "port",
            "sample": f"detectors {row0}..{row0 + n_rows - 1} x voxel planes z={plane}..{plane + n_planes - 1} "
                      f"({n} elements), FP32 restatement of photodetectorCudaMe (g++ -O3, OpenMP, "
                      f"{threads} threads), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--block", type=int, default=1024, help="detector rows per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=0, help="CPU baseline rows (0: one per host core, <= 64)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from sysmat import params
    from sysmat.pe import PePlan

    geo = params.config2(seed=0)
    nv, sub = 3, (2, 2, 2)
    N, V = geo.n_det, geo.n_vox
    B = min(args.block, N)
    n_blocks = N // B
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    plan = PePlan(geo, voxel_sub=nv, sub=sub, device=0)
    out = torch.empty((B, V), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def block_of(i):
        return ((i * 5) % n_blocks) * B

    for i in range(args.warmup):
        plan.launch(block_of(i), B, out, stream=stream)
    plan.sync()
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        plan.launch(block_of(args.warmup + i), B, out, stream=stream)
        ev[i][1].record(stream)
    plan.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    elems = args.steps * B * V
    value = elems / wall
    ms_per_step = wall * 1e3 / args.steps
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    nnz = int((out > 0).sum().item())

    fpe = load_flops("c2_pe_nv3_sub222")
    roof = None
    if fpe:
        achieved = fpe * B * V / avg_kernel_s / 1e12
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "flops_per_element": fpe, "kernel": "pe_tile_kernel",
                "avg_launch_ms": round(avg_kernel_s * 1e3, 3)}

    if args.calib_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        rng = np.random.default_rng(0)
        n_rows = args.cpu_rows if args.cpu_rows > 0 else min(64, max(8, os.cpu_count() or 8))
        row0 = int(rng.integers(0, N - n_rows))
        plane = int(rng.integers(0, geo.shape[2] - 4))
        cpu = cpu_baseline(geo, nv, sub, row0, n_rows, plane)

    rec = {
        "metric": "sysmat elements/s (PE, PE+scatter) plain vs surrogate @1% err, 1 MI355X",
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
        "config": {"workload": "PE realistic: 64^3 x 0.5 mm image, 2 W plates 8x8 holes, 4x64x64 crystals, "
                               "3^3 voxel x 2^3 crystal sub-points, dense HBM block",
                   "detectors_per_step": B, "voxels": V, "nonzero_in_last_block": nnz},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        rec["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(rec))
    plan.close()


if __name__ == "__main__":
    main()
