"""Time the self-consistent MLEM loop (attenuation recomputed from the current
estimate every few iterations) on a seeded phantom.

    python -m fl_correction_amd.bench_joint --size 128 --elements 3 --angles 180 \
        --cone 5 --half-angle 15 --iters 50 --recompute-every 5

The measured projections are the attenuated forward model of an ellipsoid
phantom (seeded cross-sections, rho = 1, pix = 1 um) with Poisson noise at 1e4
peak counts.  The loop starts from ones and runs joint.self_consistent_mlem,
its first block unattenuated unless --from-ones is given.
Prints one JSON line: total seconds, the seconds spent recomputing attenuation
and in MLEM iterations, and the data residual ||Pf(C) - I|| / ||I|| of the
start and of the result (Pf with the attenuation of the result).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import phantom
from .atten_util import make_param
from .detector_mask import Cone
from .joint import data_residual, self_consistent_mlem
from .model import XRFModel


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--size", type=int, default=128, help="volume edge N (N^3 voxels)")
    ap.add_argument("--elements", type=int, default=3)
    ap.add_argument("--angles", type=int, default=180, help="angles over 0-180 degrees")
    ap.add_argument("--cone", type=int, default=5, help="n x n cone directions")
    ap.add_argument("--half-angle", type=float, default=15.0, help="cone half-angle in degrees")
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--recompute-every", type=int, default=5)
    ap.add_argument("--from-ones", action="store_true",
                    help="recompute the first attenuation from the uniform start too (default: the first block "
                         "runs unattenuated, as the reference's demo does)")
    ap.add_argument("--seed", type=int, default=20260001)
    args = ap.parse_args(argv)

    dev = torch.device("cuda:0")
    K, N, A = args.elements, args.size, args.angles
    rng = np.random.default_rng(args.seed)
    param = make_param([f"E{i}" for i in range(K)], [f"C{i}" for i in range(K)], rng.uniform(20, 400, K),
                       rng.uniform(20, 400, (K, K)), rho=1.0, pix=1e-4)
    angles = np.linspace(0.0, 180.0, A, endpoint=False)
    model = XRFModel(param, angles, Cone(n=args.cone, half_angle_deg=args.half_angle, q=8), store=False)
    C_true = phantom.ellipsoid_phantom_device((N, N, N), K, n_ell=(10, 30), semi_axes=(6.0, 32.0), seed=args.seed,
                                              device=dev)
    Pf = model.forward(C_true)
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)
    s = 1.0e4 / float(Pf.max())
    I = (torch.poisson(Pf.clamp_min(0) * s, generator=gen) / s).contiguous()
    C0 = torch.ones_like(C_true)
    res0 = data_residual(model.forward(C0), I)

    # Split of the time: attenuation recomputes vs MLEM iterations.
    t_att = [0.0]
    attenuation = model.attenuation

    def timed_attenuation(C, out=None):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = attenuation(C, out)
        torch.cuda.synchronize()
        t_att[0] += time.perf_counter() - t
        return r

    model.attenuation = timed_attenuation
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    C = self_consistent_mlem(model, I, C0, n_iter=args.iters, recompute_every=args.recompute_every,
                             start_unattenuated=not args.from_ones)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    model.attenuation = attenuation
    res = data_residual(model.forward(C), I)
    print(json.dumps({
        "volume": [N, N, N], "elements": K, "angles": A,
        "cone": f"{args.cone}x{args.cone} directions, {args.half_angle:g} deg half-angle, slopes in 1/8",
        "iterations": args.iters, "recompute_every": args.recompute_every,
        "first_block": "attenuation of the uniform start" if args.from_ones else "unattenuated",
        "total_s": total, "attenuation_s": t_att[0], "mlem_s": total - t_att[0],
        "data_residual_start": res0, "data_residual_result": res,
        "rel_l2_vs_phantom": float((C - C_true).norm() / C_true.norm()),
    }))


if __name__ == "__main__":
    main()
