"""Time gradient-based joint reconstruction (an MLEM start, then Adam through the
attenuated model) on a seeded phantom, at any shape.

    python -m fl_correction_amd.bench_adam --slices 128 --size 512 --elements 5 \
        --angles 720 --full-turn --cone 7 --half-angle 20 --mlem-iters 10 --steps 3

The measured projections are the attenuated forward model of an ellipsoid
phantom (seeded cross-sections, rho = 1, pix = 1 um) with Poisson noise at 1e4
peak counts.  joint.mlem_start runs --mlem-iters MLEM iterations per line from
a uniform start on the start's attenuation (one line's maps at a time when all
lines' would not fit), then joint.joint_reconstruct runs --steps Adam steps,
each one XRFModel.loss_and_grad over all angles.  Prints a line per phase and
step, then one JSON line: the seconds of each phase, seconds per Adam step and
the data residual ||Pf - I|| / ||I|| after the start and after each step (and,
with --plain-mlem N, after N plain MLEM iterations from the uniform start).
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
from .joint import data_residual, joint_reconstruct, mlem_start
from .model import XRFModel


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--slices", type=int, default=32)
    ap.add_argument("--size", type=int, default=128, help="lateral edge H = W")
    ap.add_argument("--elements", type=int, default=3)
    ap.add_argument("--angles", type=int, default=180)
    ap.add_argument("--full-turn", action="store_true", help="angles over 0-360 degrees instead of 0-180")
    ap.add_argument("--cone", type=int, default=5, help="n x n cone directions")
    ap.add_argument("--half-angle", type=float, default=15.0, help="cone half-angle in degrees")
    ap.add_argument("--mlem-iters", type=int, default=10)
    ap.add_argument("--steps", type=int, default=5, help="Adam steps timed")
    ap.add_argument("--plain-mlem", type=int, default=0,
                    help="also run this many plain MLEM iterations from the uniform start (on its attenuation) "
                         "and report their data residual beside Adam's")
    ap.add_argument("--seed", type=int, default=20260005)
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[bench_adam] {msg}", flush=True)

    dev = torch.device("cuda:0")
    K, B, N, A = args.elements, args.slices, args.size, args.angles
    rng = np.random.default_rng(args.seed)
    param = make_param([f"E{i}" for i in range(K)], [f"C{i}" for i in range(K)], rng.uniform(20, 400, K),
                       rng.uniform(20, 400, (K, K)), rho=1.0, pix=1e-4)
    angles = np.linspace(0.0, 360.0 if args.full_turn else 180.0, A, endpoint=False)
    model = XRFModel(param, angles, Cone(n=args.cone, half_angle_deg=args.half_angle, q=8), store=False)
    C_true = phantom.ellipsoid_phantom_device((B, N, N), K, n_ell=(10, 30), semi_axes=(6.0, 32.0), seed=args.seed,
                                              device=dev)

    def sync_time():
        torch.cuda.synchronize()
        return time.perf_counter()

    t0 = sync_time()
    Pf = model.forward(C_true)
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)
    s = 1.0e4 / float(Pf.max())
    I = (torch.poisson(Pf.clamp_min(0) * s, generator=gen) / s).contiguous()
    del Pf
    t_data = sync_time() - t0
    log(f"projections of the phantom ({K} lines, {A} angles, {B}x{N}x{N}): {t_data:.1f} s")

    C0 = torch.full_like(C_true, float(C_true.mean()))
    t0 = sync_time()
    C_mlem = mlem_start(model, I, C0, n_iter=args.mlem_iters)
    t_mlem = sync_time() - t0
    log(f"MLEM start, {args.mlem_iters} iterations per line: {t_mlem:.1f} s")

    step_s = []
    last = [sync_time()]

    def on_step(step, C, loss):
        now = sync_time()
        step_s.append(now - last[0])
        last[0] = now
        log(f"Adam step {step + 1}/{args.steps}: {step_s[-1]:.1f} s, loss {loss:.6g}")

    t0 = sync_time()
    last[0] = t0
    C, hist = joint_reconstruct(model, I, C_mlem, n_steps=args.steps, callback=on_step)
    t_adam = sync_time() - t0
    per_step = float(np.mean(step_s)) if step_s else None
    plain = {}
    if args.plain_mlem > 0:
        t0 = sync_time()
        C_plain = mlem_start(model, I, C0, n_iter=args.plain_mlem)
        t_plain = sync_time() - t0
        res_plain = data_residual(model.forward(C_plain), I)
        log(f"{args.plain_mlem} plain MLEM iterations: {t_plain:.1f} s, data residual {res_plain:.4f}")
        plain = {"plain_mlem_iters": args.plain_mlem, "plain_mlem_s": t_plain, "plain_mlem_data_residual": res_plain}
    print(json.dumps({
        "volume": [B, N, N], "elements": K, "angles": A, "angle_span_deg": 360 if args.full_turn else 180,
        "cone": f"{args.cone}x{args.cone} directions, {args.half_angle:g} deg half-angle, slopes in 1/8",
        "projections_s": t_data, "mlem_start_s": t_mlem, "mlem_iters": args.mlem_iters,
        "adam_steps": args.steps, "adam_s": t_adam, "adam_s_per_step": per_step,
        "adam_100_steps_s_extrapolated": None if per_step is None else 100 * per_step,
        "data_residual_after_start": hist[0][2], "data_residual_per_step": [h[2] for h in hist[1:]],
        "loss_per_step": [h[1] for h in hist],
        **plain,
        "peak_memory_gb": torch.cuda.max_memory_allocated() / 1e9,
        "finite": bool(torch.isfinite(C).all()),
    }))


if __name__ == "__main__":
    main()
