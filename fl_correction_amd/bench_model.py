"""Time the differentiable model (XRFModel) at any volume shape on a few angles.

    python -m fl_correction_amd.bench_model --slices 128 --size 512 --elements 5 \
        --angles 720 --timed-angles 4 --cone 7 --half-angle 20

runs forward and vjp on `--timed-angles` of the `--angles` angles (evenly
spread over 0-360 degrees, or 0-180 with --half-turn), with the fluorescence
exponents recomputed per angle (store=False) unless --store is given, and
prints per-angle times, the ratio vjp / forward, the time of one full gradient
evaluation (per-angle time x all angles: angles are independent), the same for
the one-pass Poisson loss and gradient (XRFModel.loss_and_grad) and the peak
device memory.  Inputs: a seeded ellipsoid phantom and seeded cross-sections.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from .atten_util import make_param
from .detector_mask import Cone
from .model import XRFModel
from . import phantom


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--slices", type=int, default=128)
    ap.add_argument("--size", type=int, default=512, help="lateral edge H = W")
    ap.add_argument("--elements", type=int, default=5)
    ap.add_argument("--angles", type=int, default=720)
    ap.add_argument("--half-turn", action="store_true", help="angles over 0-180 degrees instead of 0-360")
    ap.add_argument("--timed-angles", type=int, default=4)
    ap.add_argument("--cone", type=int, default=7, help="n x n cone directions")
    ap.add_argument("--half-angle", type=float, default=20.0, help="cone half-angle in degrees")
    ap.add_argument("--store", action="store_true", help="keep the exponents between forward and vjp")
    ap.add_argument("--seed", type=int, default=20260004)
    args = ap.parse_args(argv)

    dev = torch.device("cuda:0")
    K, B, N, A = args.elements, args.slices, args.size, args.angles
    na = max(1, min(args.timed_angles, A))
    rng = np.random.default_rng(args.seed)
    param = make_param([f"E{i}" for i in range(K)], [f"C{i}" for i in range(K)], rng.uniform(20, 400, K),
                       rng.uniform(20, 400, (K, K)), rho=1.0, pix=1e-4)
    span = 180.0 if args.half_turn else 360.0
    all_angles = np.linspace(0.0, span, A, endpoint=False)
    angles = all_angles[np.linspace(0, A - 1, na).round().astype(int)]
    cone = Cone(n=args.cone, half_angle_deg=args.half_angle, q=8)
    C = phantom.ellipsoid_phantom_device((B, N, N), K, n_ell=(10, 30), semi_axes=(6.0, 32.0), seed=args.seed,
                                         device=dev)
    model = XRFModel(param, angles, cone, store=bool(args.store))

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3, out

    model.forward(C)  # warm-up
    t_f, Pf = timed(lambda: model.forward(C))
    G = torch.randn_like(Pf)
    t_b, g = timed(lambda: model.vjp(C, G))
    I = Pf * (torch.rand_like(Pf) * 0.4 + 0.8)
    model.loss_and_grad(C, I)  # warm-up
    t_lg, (loss, g2, _) = timed(lambda: model.loss_and_grad(C, I))
    ok = all(bool(torch.isfinite(x).all()) for x in (Pf, g, g2, loss))
    model.release()
    exponent_gb = 4.0 * A * K * B * N * N / 1e9
    print(json.dumps({
        "volume": [B, N, N], "elements": K, "angles": A, "angles_timed": na,
        "cone": f"{args.cone}x{args.cone} directions, {args.half_angle:g} deg half-angle, slopes in 1/8",
        "exponents": "kept" if args.store else "recomputed per angle",
        "forward_ms_per_angle": t_f / na, "vjp_ms_per_angle": t_b / na, "vjp_over_forward": t_b / t_f,
        "grad_eval_s": (t_f + t_b) / na * A / 1e3,
        "loss_and_grad_ms_per_angle": t_lg / na,
        "loss_and_grad_eval_s": t_lg / na * A / 1e3,
        "exponents_gb_all_angles": exponent_gb,
        "peak_memory_gb": torch.cuda.max_memory_allocated() / 1e9,
        "finite": ok,
    }))
    if not ok:
        raise SystemExit("non-finite model output")


if __name__ == "__main__":
    main()
