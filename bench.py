#!/usr/bin/env python
"""Benchmark: plain MLEM iterations of the XRF self-absorption hot path on one MI355X.

Workload: 256^3 voxels x 360 angles x 4 emission lines (one reference per line),
FP32, attenuation stored in HBM (the form mlem_cuda_batch iterates on,
FL/cuda_lib/atten_cuda.cpp:283-345).  One step = one MLEM iteration of all four
lines: forward projection + ratio, back-projection + update, per line.
Inputs are synthetic: seeded ellipsoid phantoms, attenuation maps, and Poisson
projections at 1e4 peak counts.

Prints one JSON line (rank 0).  --gpus above 1 is refused: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256, help="volume edge N (N^3 voxels)")
    ap.add_argument("--angles", type=int, default=360)
    ap.add_argument("--elements", type=int, default=4)
    ap.add_argument("--cpu-angles", type=int, default=24,
                    help="angles of one line timed on the CPU for the baseline (scaled to the full step)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--synthetic-atten", action="store_true",
                    help="uniform random attenuation in [0.3, 1] instead of the physics pipeline")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the design for sharding slices or angles is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch
    sys.path.insert(0, ROOT)
    from fl_correction_amd import phantom
    from fl_correction_amd.cuda_lib import atten_cuda
    from fl_correction_amd.mlem import MLEMSolver

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N, A, K = args.size, args.angles, args.elements
    B = H = W = N
    R = 1
    pix, rho = 1.0e-4, 1.0

    t0 = time.time()
    C_true = phantom.ellipsoid_phantom_device((B, H, W), K, n_ell=(10, 30), semi_axes=(6.0, 32.0), seed=20260002,
                                              device=dev)
    theta = torch.tensor(np.deg2rad(phantom.angle_list(A)), dtype=torch.float32, device=dev)
    em = torch.full((A, R), rho * pix, dtype=torch.float32, device=dev)
    log(f"phantom {tuple(C_true.shape)} in {time.time() - t0:.1f}s")

    t0 = time.time()
    atten = torch.empty((K, A, B, H, W), dtype=torch.float32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(20260002)
    for e in range(K):
        atten[e].uniform_(0.3, 1.0, generator=gen)
    atten_kind = "uniform random in [0.3, 1]"
    torch.cuda.synchronize()
    log(f"attenuation {tuple(atten.shape)} ({atten_kind}) in {time.time() - t0:.1f}s")

    t0 = time.time()
    I = []
    for e in range(K):
        Pf = atten_cuda.forward_emission_batch(atten[e], C_true[e:e + 1], em, theta)
        s = 1.0e4 / float(Pf.max())
        I.append((torch.poisson(Pf.clamp_min(0) * s, generator=gen) / s).contiguous())
    solvers = [MLEMSolver(atten[e], em, theta, I[e], beta=0.0) for e in range(K)]
    C = [torch.ones((R, B, H, W), dtype=torch.float32, device=dev) for _ in range(K)]
    Cn = [torch.empty_like(c) for c in C]
    torch.cuda.synchronize()
    log(f"projections + sensitivities in {time.time() - t0:.1f}s")

    def step(events=None):
        for e in range(K):
            if events is not None:
                events[e][0].record()
            solvers[e].ratio_step(C[e])
            if events is not None:
                events[e][1].record()
            solvers[e].update_step(C[e], Cn[e])
            if events is not None:
                events[e][2].record()
            C[e], Cn[e] = Cn[e], C[e]

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    evs = [[[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(K)] for _ in range(args.steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record()
    for k in range(args.steps):
        step(evs[k])
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    total_ms = start.elapsed_time(stop)
    ms_per_step = total_ms / max(args.steps, 1)

    fwd_ms = [evs[k][e][0].elapsed_time(evs[k][e][1]) for k in range(args.steps) for e in range(K)]
    upd_ms = [evs[k][e][1].elapsed_time(evs[k][e][2]) for k in range(args.steps) for e in range(K)]
    fwd_avg, upd_avg = float(np.mean(fwd_ms)), float(np.mean(upd_ms))

    n3 = B * H * W
    fwd_bytes = 4.0 * (A * n3 + R * n3 + A * B * W)          # F1 per line
    upd_bytes = 4.0 * (A * n3 + A * B * W + 3 * R * n3)      # B1 + fused update per line
    iter_bytes = K * (fwd_bytes + upd_bytes + 4.0 * (A * B * W))  # + ratio pass
    if upd_avg >= fwd_avg:
        dom, dom_ms, dom_bytes = "bwd_tile_kernel (back-projection + MLEM update)", upd_avg, upd_bytes
    else:
        dom, dom_ms, dom_bytes = "fwd_tile_kernel (forward projection + ratio)", fwd_avg, fwd_bytes
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9

    result = {
        "metric": "MLEM iterations/s at 256^3 voxels x 360 angles x 4 elements (FP32, stored attenuation)",
        "value": 1000.0 / ms_per_step,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: seeded ellipsoid phantom, attenuation {atten_kind}, Poisson 1e4 peak",
        "config": {"workload": "plain MLEM iteration (beta=0), forward + ratio + back-projection + update, "
                               "all emission lines", "volume": [B, H, W], "angles": A, "elements": K,
                   "refs_per_line": R},
        "stage_ms": {"forward_ratio_per_line": fwd_avg, "backproject_update_per_line": upd_avg},
        "hbm_algorithmic_gbs_per_iteration": iter_bytes / (ms_per_step * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "wall_s_timed": wall,
    }

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args, atten, C_true, em, theta, I)

    print(json.dumps(result), flush=True)


def cpu_baseline(args, atten, C_true, em, theta, I):
    """The oracle (the FP32 C restatement of the reference kernels, OpenMP over
    independent outputs) on one line and a subset of the angles at full volume,
    scaled linearly to all angles and lines (angles are independent)."""
    ncores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(ncores))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker's CPU restatement, used here as the timed CPU port

    ka = min(args.cpu_angles, args.angles)
    a_np = atten[0, :ka].cpu().numpy()
    c_np = np.ones((1,) + tuple(C_true.shape[1:]), np.float32)
    em_np = em[:ka].cpu().numpy()
    th_np = theta[:ka].cpu().numpy()
    i_np = I[0][:ka].cpu().numpy()
    sens = oracle.backward_batch_f32(a_np, np.ones_like(i_np), em_np, th_np)  # untimed, as on the GPU
    t0 = time.perf_counter()
    Pf = oracle.forward_batch_f32(a_np, c_np, em_np, th_np)
    ratio = i_np / np.maximum(Pf, np.float32(1e-6))
    back = oracle.backward_batch_f32(a_np, ratio, em_np, th_np)
    c_new = np.maximum(c_np * back / np.maximum(sens, np.float32(1e-6)), 0)
    dt = time.perf_counter() - t0
    del c_new
    scale = (args.angles / ka) * args.elements
    return {"value": 1.0 / (dt * scale), "unit": "iter/s", "cores": ncores, "kind": "port",
            "sample": f"one line, {ka} of {args.angles} angles at full volume: {dt:.2f}s, "
                      f"scaled x{scale:g} (angles and lines are independent)"}


if __name__ == "__main__":
    main()
