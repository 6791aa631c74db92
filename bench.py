#!/usr/bin/env python
"""Benchmark: MLEM iterations of the XRF self-absorption hot path on one MI355X.

Workload: 256^3 voxels x 360 angles x 4 emission lines (one reference per line),
FP32, attenuation resident in HBM (the form mlem_cuda_batch iterates on,
FL/cuda_lib/atten_cuda.cpp:283-345).  One step = one plain MLEM iteration of
all four lines: forward projection + ratio, then back-projection + update, per
line.  The back-projection runs on a slice-major copy of each line's
attenuation (A, H, W, B), made once per attenuation map before the timed steps
(its time is reported as set-up), when the copies fit in HBM.  The lines are independent problems, so each runs on its own stream
(default priority) and their kernels overlap; --sequential-lines runs them one
after another on one stream.  Inputs are synthetic: seeded ellipsoid phantoms,
attenuation maps and Poisson projections at 1e4 peak counts.

The roofline kernel durations come from HIP events around each kernel over K
steps run one line at a time (right after the timed steps, or the timed steps
themselves with --sequential-lines): in the concurrent schedule the kernels of
different lines share the GPU, so their event spans overlap.

Besides the timed steps, the JSON line carries
  * roofline: the dominant kernel's algorithmic HBM bytes over its average
    duration, timed with HIP events on the stream it is launched on;
  * model: attenuation, forward and gradient (vjp) of the differentiable model
    with a 5x5 detector cone over 15 deg, timed over all 360 angles, and the
    cone march's roofline (flop count and its own HBM traffic);
  * cpu_baseline: the oracle's CPU restatement of the same iteration, timed on
    one line on the host cores and scaled to all lines; beside it, at the
    reference's CPU-sized case (64^3, 2 lines, 60 angles), the legacy dense-matrix
    path and the oracle's pure-torch model run by torch on the host cores.

--gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256, help="volume edge N (N^3 voxels)")
    ap.add_argument("--angles", type=int, default=360)
    ap.add_argument("--elements", type=int, default=4)
    ap.add_argument("--cpu-angles", type=int, default=240,
                    help="angles of one line timed on the CPU for the baseline (scaled to the full step)")
    ap.add_argument("--model-angles", type=int, default=-1,
                    help="angles of the differentiable model timed (-1: all of --angles, measured; "
                         "fewer: extrapolated per angle; 0: skip)")
    ap.add_argument("--batch-samples", type=int, default=48,
                    help="samples of the batched uneven-instance workload (0: skip)")
    ap.add_argument("--batch-iters", type=int, default=20)
    ap.add_argument("--batch-shapes", choices=("tuned", "default"), default="tuned",
                    help="tile shapes of the batch: per size bucket from the autotuner, or the library defaults")
    ap.add_argument("--no-batch-compare", action="store_true",
                    help="time only the batch itself (no reference-layout, per-bucket or sequential runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sequential-lines", action="store_true",
                    help="time the lines one after another on one stream (default: one stream per line, "
                         "the lines being independent problems)")
    ap.add_argument("--layout", choices=("auto", "slice_major", "reference"), default="auto",
                    help="back-projection layout of the MLEM solvers (auto: slice-major when its copy of the "
                         "attenuation fits in HBM)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run two elementwise kernels of known traffic (4 GiB read + 4 GiB written, with "
                         "16-byte and with 4-byte accesses), to scale HBM counters")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(slices shard with no communication while the attenuation is fixed; that path is not built).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    sys.path.insert(0, ROOT)
    from fl_correction_amd import phantom
    from fl_correction_amd.cuda_lib import atten_cuda
    from fl_correction_amd.mlem import MLEMSolver

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calibrate_copy:
        src = torch.ones(1 << 30, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        # an elementwise kernel that reads and writes exactly 4 GiB (a plain
        # copy_ of contiguous tensors is a DMA transfer, not a kernel)
        torch.mul(src, 2.0, out=dst)
        # the same with 4-byte accesses: a base one float off 16-byte alignment
        # makes torch fall back to its scalar elementwise kernel
        torch.mul(src[1:], 2.0, out=dst[1:])
        torch.cuda.synchronize()
        del src, dst
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

    # The kernels' cost does not depend on the attenuation values; the physics
    # pipeline for all 360 angles is timed separately below (model timings).
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
    torch.cuda.synchronize()
    log(f"projections in {time.time() - t0:.1f}s")
    # Each solver copies its line's attenuation to the slice-major layout once
    # (when it fits in HBM) and computes the sensitivity; the copy is timed here.
    solvers = []
    copy_ms = []
    for e in range(K):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        solvers.append(MLEMSolver(atten[e], em, theta, I[e], beta=0.0, layout=args.layout))
        e1.record()
        torch.cuda.synchronize()
        copy_ms.append(e0.elapsed_time(e1))
    layouts = sorted({s.layout for s in solvers})
    C = [torch.ones((R, B, H, W), dtype=torch.float32, device=dev) for _ in range(K)]
    Cn = [torch.empty_like(c) for c in C]
    torch.cuda.synchronize()
    log(f"solvers ({', '.join(layouts)} layout): set-up {np.mean(copy_ms):.1f} ms per line")

    def step(events=None):
        # torch.cuda.Event records on torch's current stream, the stream the
        # library launches on: each pair of events brackets exactly one kernel.
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

    # The lines are independent problems: by default each runs on its own stream
    # (default priority), so one line's kernels overlap another's.
    streams = [torch.cuda.Stream(device=dev) for _ in range(K)]

    def step_concurrent():
        for e in range(K):
            with torch.cuda.stream(streams[e]):
                solvers[e].ratio_step(C[e])
                solvers[e].update_step(C[e], Cn[e])
            C[e], Cn[e] = Cn[e], C[e]

    def timed_steps(fn, with_events):
        """K steps between two events on the current stream (every line's stream
        waits for the first and is waited for by the second)."""
        evs = [[[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(K)] for _ in range(args.steps)]
        cur = torch.cuda.current_stream(dev)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t_wall = time.perf_counter()
        start.record(cur)
        for st in streams:
            st.wait_event(start)
        for k in range(args.steps):
            if with_events:
                fn(evs[k])
            else:
                fn()
        for st in streams:
            done = torch.cuda.Event()
            done.record(st)
            cur.wait_event(done)
        stop.record(cur)
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / max(args.steps, 1), time.perf_counter() - t_wall, evs

    concurrent = not args.sequential_lines
    for _ in range(args.warmup):
        step_concurrent() if concurrent else step()
    torch.cuda.synchronize()

    if concurrent:
        ms_per_step, wall, _ = timed_steps(step_concurrent, False)
        # Kernel durations for the roofline: the same K steps one line at a time
        # right after the timed region (in the timed steps the lines' kernels
        # share the GPU, so their event spans overlap).
        seq_ms_per_step, _, evs = timed_steps(step, True)
    else:
        ms_per_step, wall, evs = timed_steps(step, True)
        seq_ms_per_step = ms_per_step
    log(f"{args.steps} steps: {ms_per_step:.2f} ms/step ({'one stream per line' if concurrent else 'one stream'}); "
        f"one line at a time {seq_ms_per_step:.2f} ms/step")

    fwd_ms = [evs[k][e][0].elapsed_time(evs[k][e][1]) for k in range(args.steps) for e in range(K)]
    upd_ms = [evs[k][e][1].elapsed_time(evs[k][e][2]) for k in range(args.steps) for e in range(K)]
    fwd_avg, upd_avg = float(np.mean(fwd_ms)), float(np.mean(upd_ms))

    # Algorithmic HBM bytes per launch (each array read or written once).
    n3 = B * H * W
    fwd_bytes = 4.0 * (A * n3 + R * n3 + 2 * A * B * W)      # atten, C, I in, ratio out
    upd_bytes = 4.0 * (A * n3 + A * B * W + 3 * R * n3)      # atten, ratio, C, sens in, C_next out
    iter_bytes = K * (fwd_bytes + upd_bytes)
    slice_major = layouts == ["slice_major"]
    if slice_major:
        upd_name = ("slice-major back-projection + MLEM update: sm_fused_kernel<4,true> (interior blocks and "
                    "edge strips), sm_edge_reduce_kernel<4>, sm_finalize_kernel (one library call)")
        upd_key = "slice-major back-projection + update"
        fwd_name = "fwd_strip_kernel<1,8> (forward projection + ratio, written slice-major)"
    else:
        upd_name, upd_key = "bwd_tile_kernel<1,31> (back-projection + MLEM update)", "bwd_tile_kernel<1,31>"
        fwd_name = "fwd_strip_kernel<1,8> (forward projection + ratio)"
    fwd_key = "fwd_strip_kernel<1,8>"
    if upd_avg >= fwd_avg:
        dom, dom_ms, dom_bytes, dom_key = upd_name, upd_avg, upd_bytes, upd_key
    else:
        dom, dom_ms, dom_bytes, dom_key = fwd_name, fwd_avg, fwd_bytes, fwd_key
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    stage_roofline = {
        name: {"kernel": kname, "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": measured_traffic(key, N, A),
               "bytes_per_launch": b, "avg_launch_ms": ms}
        for name, kname, key, b, ms in (("forward_ratio", fwd_name, fwd_key, fwd_bytes, fwd_avg),
                                        ("backproject_update", upd_name, upd_key, upd_bytes, upd_avg))}

    result = {
        "metric": "MLEM iter time & grad-evals/s at 256^3 voxels x 360 angles x 4 elements; HBM GB/s vs peak",
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
        "config": {"workload": "plain MLEM iteration (beta=0) with attenuation resident in HBM: forward + ratio, "
                               "back-projection + update, all emission lines",
                   "volume": [B, H, W], "angles": A, "elements": K, "refs_per_line": R,
                   "backprojection_layout": layouts[0] if len(layouts) == 1 else layouts},
        "schedule": "one stream per line (the lines are independent)" if concurrent else "one stream",
        "sequential_ms_per_step": seq_ms_per_step,
        "setup_ms_per_line": {"value": float(np.mean(copy_ms)),
                              "what": "solver set-up before the timed steps: slice-major copy of the line's "
                                      "attenuation (slice-major layout only) and the sensitivity back-projection; "
                                      "done once per attenuation map"},
        "stage_ms": {"forward_ratio_per_line": fwd_avg, "backproject_update_per_line": upd_avg,
                     "timing": "kernel events over K steps run one line at a time" +
                               (", right after the timed steps" if concurrent else " (the timed steps)")},
        "hbm_algorithmic_gbs_per_iteration": iter_bytes / (ms_per_step * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom_key, N, A),
                     "bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms},
        "stage_roofline": stage_roofline,
        "wall_s_timed": wall,
    }

    # free the MLEM state before the model timings
    del solvers, C, Cn, I, atten
    torch.cuda.empty_cache()

    if args.model_angles != 0:
        result["model"] = model_timings(args, C_true, pix)

    if args.batch_samples > 0:
        result["batch"] = batch_timings(args)

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args)
        result["cpu_baseline"]["legacy_dense"] = legacy_dense_small(args)
        result["cpu_baseline"]["torch_cpu"] = torch_cpu_small(
            result["cpu_baseline"]["legacy_dense"]["gpu_ms_per_iteration"])
        if "model" in result:
            # the iteration that recomputes the attenuation for every angle first
            m = result["model"]
            cpu_att = cpu_attenuation_sample(args)
            gpu_ms = ms_per_step + (m["attenuation_ms_all_angles"] if m.get("attenuation_ms_all_angles")
                                    else A * m["attenuation_ms_per_angle"])
            cpu_s = 1.0 / result["cpu_baseline"]["value"] + A * cpu_att["s_per_angle"]
            result["iteration_with_attenuation"] = {
                "gpu_ms": gpu_ms, "cpu_s": cpu_s, "speedup": cpu_s * 1e3 / gpu_ms,
                "cpu_attenuation": cpu_att,
                "note": "MLEM step plus attenuation maps of all lines at every angle (5x5 cone over 15 deg); "
                        "GPU attenuation timed over all angles in the model timings (per angle scaled when "
                        "--model-angles is smaller), CPU from the oracle's direct FP64 march on a slab, scaled "
                        "per angle"}

    print(json.dumps(result), flush=True)


def measured_traffic(kernel, N, A):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (read and write requests by size class, see the file), for the
    256^3 x 360-angle workload it was measured on; None for other sizes."""
    import re
    if (N, A) != (256, 360):
        return None
    path = os.path.join(ROOT, "profiles", "mlem_iteration_256cubed_360x4_hbm_counters.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic " + re.escape(kernel) + r": ([0-9.e+]+) bytes per launch", text, re.M)
    return float(m.group(1)) if m else None


def model_timings(args, C_true, pix):
    """Attenuation, forward and vjp of the differentiable model (XRFModel) over all
    the angles (or, with --model-angles N, over N of them, the per-angle times then
    scaled to all angles and marked extrapolated), and the cone march's roofline."""
    import torch
    from fl_correction_amd import _lib
    from fl_correction_amd.atten_util import make_param
    from fl_correction_amd.detector_mask import Cone
    from fl_correction_amd.model import XRFModel

    K, A = args.elements, args.angles
    rng = np.random.default_rng(20260002)
    cs_x = rng.uniform(20, 400, K)
    cs_line = rng.uniform(20, 400, (K, K))
    param = make_param([f"E{i}" for i in range(K)], [f"C{i}" for i in range(K)], cs_x, cs_line, rho=1.0, pix=pix)
    na = A if args.model_angles < 0 else min(args.model_angles, A)
    measured = na == A
    angles = np.linspace(0.0, 180.0, na, endpoint=False)
    cone = Cone(n=5, half_angle_deg=15.0, q=8)
    Cm = (C_true * 0.5).contiguous()

    def timed(fn, warm=None):
        # warm-up on a few angles' model (kernels, workspaces), then the timed call
        (warm or fn)()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        out = fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), out

    # the warm-up model's angle chunks (16 angles) have the timed model's sizes, so
    # that the timed calls find their scratch sizes in torch's cache
    wmodel = XRFModel(param, angles[:min(16, na)], cone, store=False)
    Gw = torch.randn_like(wmodel.forward(Cm))
    model = XRFModel(param, angles, cone, store=False)
    t_att, att = timed(lambda: model.attenuation(Cm), lambda: wmodel.attenuation(Cm))
    del att
    t_fwd, Pf = timed(lambda: model.forward(Cm), lambda: wmodel.forward(Cm))
    G = torch.randn_like(Pf)
    t_vjp, _ = timed(lambda: model.vjp(Cm, G), lambda: wmodel.vjp(Cm, Gw))
    # Poisson loss, gradient and projections in one pass (what joint reconstruction runs)
    Iw = torch.rand_like(Gw) + 0.5
    Im = Pf * (torch.rand_like(Pf) * 0.4 + 0.8)
    t_lg, _ = timed(lambda: model.loss_and_grad(Cm, Im), lambda: wmodel.loss_and_grad(Cm, Iw))
    model.release()
    del model, Im
    # The same with the fluorescence exponents kept from forward to vjp.
    smodel = XRFModel(param, angles, cone, store=True)
    t_sfwd, _ = timed(lambda: smodel.forward(Cm), lambda: wmodel.forward(Cm))
    t_svjp, _ = timed(lambda: smodel.vjp(Cm, G), lambda: wmodel.vjp(Cm, Gw))
    smodel.release()
    del smodel, wmodel
    torch.cuda.empty_cache()
    per = lambda t: t / na
    log(f"model: {na} angles: attenuation {t_att:.1f} ms, forward {t_fwd:.1f} ms, vjp {t_vjp:.1f} ms, "
        f"loss and gradient in one pass {t_lg:.1f} ms; exponents kept: forward {t_sfwd:.1f} ms, "
        f"vjp {t_svjp:.1f} ms")
    B, H, W = Cm.shape[1:]
    store_bytes = 4.0 * A * K * B * H * W  # all angles' exponents
    total = torch.cuda.get_device_properties(0).total_memory
    grad = {"grad_eval_ms": t_fwd + t_vjp, "grad_evals_per_s": 1000.0 / (t_fwd + t_vjp)} if measured else \
        {"grad_eval_ms_extrapolated": per(t_fwd + t_vjp) * A,
         "grad_evals_per_s_extrapolated": 1000.0 / (per(t_fwd + t_vjp) * A)}
    fused = {"grad_eval_ms": t_lg, "grad_evals_per_s": 1000.0 / t_lg} if measured else \
        {"grad_eval_ms_extrapolated": per(t_lg) * A, "grad_evals_per_s_extrapolated": 1000.0 / (per(t_lg) * A)}
    kept = {"grad_eval_ms": t_sfwd + t_svjp, "grad_evals_per_s": 1000.0 / (t_sfwd + t_svjp)} if measured else \
        {"grad_evals_per_s_extrapolated": 1000.0 / (per(t_sfwd + t_svjp) * A)}
    return {
        "angles_timed": na, "measured_over_all_angles": measured,
        "cone": "5x5 directions, 15 deg half-angle, slopes in 1/8",
        "attenuation_ms_per_angle": per(t_att),
        "attenuation_ms_all_angles": t_att if measured else None,
        "forward_ms_per_angle": per(t_fwd),
        "vjp_ms_per_angle": per(t_vjp),
        "adjoint_over_forward": t_vjp / t_fwd,
        **grad,
        "note": "forward recomputes the attenuation for every angle (nothing stored); vjp recomputes it too "
                "and differentiates through it",
        "loss_and_grad": {
            "ms_per_angle": per(t_lg),
            **fused,
            "note": "Poisson loss, its gradient and the projections in one pass over the angles "
                    "(XRFModel.loss_and_grad, what joint reconstruction runs): each chunk of 16 angles is "
                    "projected, its cotangent formed and its adjoint run on the chunk's exponents, so the "
                    "cone runs twice per angle",
        },
        "exponents_kept": {
            "forward_ms_per_angle": per(t_sfwd),
            "vjp_ms_per_angle": per(t_svjp),
            "adjoint_over_forward": t_svjp / t_sfwd,
            **kept,
            "stored_gb_at_all_angles": store_bytes / 1e9,
            "fits_default_budget": bool(store_bytes <= 0.4 * total),
            "note": "forward keeps every angle's fluorescence exponents and vjp reuses them (the model's "
                    "store mode; its default 'auto' keeps them when they fit 40% of free HBM)",
        },
        "cone_roofline": cone_roofline(K, B, H, W, cone),
    }


def cone_roofline(K, B, H, W, cone, reps=5):
    """The cone march of one angle for all lines (fl_cone_sweep on the mask frame
    (K, B, W, H)), timed with HIP events on its stream, against both roofs: the
    ray-form flop count 2*K*N^3*sum_d taps_d (taps_d = the bilinear taps of
    direction d's Q-step stencil), and the sweep's own HBM traffic (mu read once,
    T written once, each direction's running ray sum written and read once)."""
    import torch
    from fl_correction_amd import _lib
    dev = torch.device("cuda:0")
    arrs = [np.ascontiguousarray(a) for a in (cone.pz, cone.py, cone.qd, cone.weight)]
    ptrs = [a.ctypes.data_as(_lib.ctypes.c_void_p) for a in arrs]
    Z, X, Y = B, W, H
    nb = int(_lib.load().fl_cone_sweep_workspace_bytes(K, Z, X, Y, *ptrs[:3], cone.ndir))
    mu = torch.rand((K, Z, X, Y), device=dev) * 0.01
    T = torch.empty_like(mu)
    ws = torch.empty((nb + 3) // 4, device=dev)
    st = _lib.stream_of(dev)
    run = lambda: _lib.call("fl_cone_sweep", _lib.ptr(mu), _lib.ptr(T), K, Z, X, Y, *ptrs, cone.ndir, 0,
                            _lib.ptr(ws), ws.numel() * 4, st)
    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    Q = int(max(cone.qd))
    taps = 0
    for pz, py, q in zip(cone.pz, cone.py, cone.qd):
        m = Q // int(q)
        for t in range(1, Q + 1):
            fz = (t * int(pz) * m) % Q != 0
            fy = (t * int(py) * m) % Q != 0
            taps += (2 if fz else 1) * (2 if fy else 1)
    vox = K * Z * X * Y
    flops = 2.0 * vox * taps
    traffic = 4.0 * vox * (2 + 2 * cone.ndir)
    del mu, T, ws
    return {
        "kernel": "cone_sweep_kernel<8,8,2> (the 5x5 grid's compiled-in stencil; X/Q launches: one angle, all lines)",
        "ms_per_angle": ms,
        "taps_per_voxel": taps,
        "flops": flops, "achieved_tflops": flops / ms / 1e9, "fp32_peak_tflops": FP32_PEAK_TFLOPS,
        "flop_frac": flops / ms / 1e9 / FP32_PEAK_TFLOPS,
        "traffic_bytes": traffic, "achieved_gbs": traffic / ms / 1e6, "hbm_peak_gbs": HBM_PEAK_GBS,
        "hbm_frac": traffic / ms / 1e6 / HBM_PEAK_GBS,
        "bound": "hbm (the running ray sums: 8 bytes per voxel and direction)",
    }


def batch_timings(args):
    """Batched uneven instances: seeded samples with lateral size in {64, 96, 128,
    160, 192}, 16-64 slices, 2-4 emission lines and 90-360 angles (multiples of
    30); one problem per line, the lines of a sample sharing its theta tensor.
    Tile shapes per size bucket from the autotuner; all problems iterate
    together, two iterations per replay of the captured hipGraph, in the
    layout BatchMLEM picks (slice-major: each sample's lines packed along the
    slice axis) and again in the reference layout; each size bucket alone
    (its GB/s); and the same problems one after another through the
    per-problem path (fl_mlem_ratio / fl_mlem_update launches)."""
    import tempfile
    import torch
    from fl_correction_amd.batch import Autotuner, BatchMLEM, Problem
    from fl_correction_amd.cuda_lib import atten_cuda
    from fl_correction_amd.mlem import MLEMSolver

    dev = torch.device("cuda:0")
    torch.cuda.empty_cache()  # the earlier legs' cached blocks: the batch workspaces are large
    rng = np.random.default_rng(20260003)
    gen = torch.Generator(device=dev)
    gen.manual_seed(20260003)
    probs, shapes_of = [], []
    t0 = time.time()
    for _ in range(args.batch_samples):
        L = int(rng.choice([64, 96, 128, 160, 192]))
        Bs = int(rng.integers(16, 65))
        K = int(rng.integers(2, 5))
        A = int(rng.choice(np.arange(90, 361, 30)))
        theta = torch.tensor(np.deg2rad(np.linspace(0, 180, A, endpoint=False)), dtype=torch.float32, device=dev)
        for _ in range(K):
            atten = torch.empty((A, Bs, L, L), device=dev).uniform_(0.3, 1.0, generator=gen)
            em = torch.full((A,), 1.0, device=dev)
            C_true = torch.rand((1, Bs, L, L), device=dev, generator=gen)
            I = atten_cuda.forward_emission_batch(atten, C_true, em[:, None].contiguous(), theta)
            probs.append(Problem(atten, em, theta, I.contiguous(), torch.ones((Bs, L, L), device=dev)))
            shapes_of.append((A, Bs, L, L))
    torch.cuda.synchronize()
    n_vox = sum(B * H * W for _, B, H, W in shapes_of)
    atten_bytes = sum(4.0 * A * B * H * W for A, B, H, W in shapes_of)
    log(f"batch: {args.batch_samples} samples, {len(probs)} problems, {atten_bytes / 1e9:.1f} GB attenuation, "
        f"set up in {time.time() - t0:.1f}s")
    n_it = args.batch_iters
    tuner = Autotuner(path=os.path.join(tempfile.mkdtemp(), "autotune.json")) if args.batch_shapes == "tuned" \
        else None

    def timed(sel, layout, iters):
        """ms per iteration of BatchMLEM over probs[sel] (graph replays), its
        layout, and the set-up time in ms (slice-major copies, sensitivity)."""
        sub = [probs[i] for i in sel]
        for p in sub:
            p.C.fill_(1.0)
        torch.cuda.synchronize()
        t_s = time.time()
        run = BatchMLEM(sub, beta=0.0, shapes=tuner, layout=layout)
        torch.cuda.synchronize()
        setup_ms = (time.time() - t_s) * 1e3
        run.run(2, graph=True)  # capture + first replay
        for p in sub:
            p.C.fill_(1.0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(run.stream)
        run.run(iters, graph=True)
        e1.record(run.stream)
        torch.cuda.synchronize()
        lay = run.layout
        run.close()
        del run
        torch.cuda.empty_cache()
        return e0.elapsed_time(e1) / iters, lay, setup_ms

    if tuner is not None:
        t0 = time.time()
        for s_ in shapes_of:  # tune every bucket in both layouts before timing
            tuner.shapes(*s_, device=dev, layout="slice_major")
            tuner.shapes(*s_, device=dev, layout="reference")
        log(f"batch: autotuned {len(tuner.table)} (bucket, layout) pairs in {time.time() - t0:.1f}s: {tuner.table}")

    every = list(range(len(probs)))
    batch_ms, layout, setup_ms = timed(every, "auto", n_it)
    C_batch = [p.C.clone() for p in probs[:4]]
    it_bytes = 2 * atten_bytes  # atten read by the forward and by the back-projector
    out = {
        "workload": "batched uneven MLEM: samples with lateral size in {64..192}, 16-64 slices, 2-4 lines, "
                    "90-360 angles; one problem per line; beta=0",
        "samples": args.batch_samples, "problems": len(shapes_of), "voxels": n_vox,
        "attenuation_gb": atten_bytes / 1e9, "iterations": n_it,
        "layout": layout,
        "ms_per_iteration": batch_ms,
        "atten_stream_gbs": it_bytes / (batch_ms * 1e-3) / 1e9,
        "setup_ms": setup_ms,
        "graph": "two iterations per replay",
        "tile_shapes": {k.split("|", 1)[1]: v for k, v in tuner.table.items()} if tuner else "library defaults",
    }
    log(f"batch: {n_it} iterations {batch_ms:.1f} ms per iteration ({layout})")
    if args.no_batch_compare:
        return out
    ref_ms, _, _ = timed(every, "reference", n_it)
    buckets = {}
    for L in sorted({s_[2] for s_ in shapes_of}):
        sel = [i for i, s_ in enumerate(shapes_of) if s_[2] == L]
        gb = sum(4.0 * A * B * H * W for A, B, H, W in (shapes_of[i] for i in sel)) / 1e9
        ms, _, _ = timed(sel, layout, max(4, n_it // 4))
        buckets[f"{L}x{L}"] = {"problems": len(sel), "attenuation_gb": gb, "ms_per_iteration": ms,
                               "atten_stream_gbs": 2 * gb / (ms * 1e-3)}
    log(f"batch: reference layout {ref_ms:.1f} ms; per bucket "
        + ", ".join(f"{k} {v['atten_stream_gbs']:.0f} GB/s" for k, v in buckets.items()))

    # the same problems one after another (default tile shapes, one launch per pass per problem)
    solvers = [MLEMSolver(p.atten, p.em_cs[:, None].contiguous(), p.theta, p.I, layout="reference") for p in probs]
    Cs = [torch.ones_like(p.C)[None] for p in probs]
    Cn = [torch.empty_like(c) for c in Cs]
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(n_it):
        for i, sv in enumerate(solvers):
            sv.ratio_step(Cs[i])
            sv.update_step(Cs[i], Cn[i])
            Cs[i], Cn[i] = Cn[i], Cs[i]
    e1.record()
    torch.cuda.synchronize()
    seq_ms = e0.elapsed_time(e1)
    agree = max(float((a - b[0]).norm() / b[0].norm()) for a, b in zip(C_batch, Cs[:4]))
    del solvers, Cs, Cn, probs
    torch.cuda.empty_cache()
    log(f"batch: sequential {seq_ms / n_it:.1f} ms per iteration")
    out.update({
        "reference_layout_ms_per_iteration": ref_ms,
        "sequential_ms_per_iteration": seq_ms / n_it,
        "speedup_vs_sequential": seq_ms / n_it / batch_ms,
        "per_bucket": buckets,
        "rel_l2_batch_vs_sequential_first4": agree,
    })
    return out


def cpu_attenuation_sample(args):
    """The oracle's direct FP64 cone march (every voxel sums its rays' bilinear
    samples over the full depth, the per-voxel form of the reference's CPU
    attenuation, FL_correction_core.py:1498-1518) on a slab of 4 slices of one
    line, all host cores; scaled to all slices and lines of one angle."""
    ncores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(ncores))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the oracle's CPU restatement, used here as the timed CPU port
    from fl_correction_amd.detector_mask import Cone

    N, K, Zs = args.size, args.elements, 4
    cone = Cone(n=5, half_angle_deg=15.0, q=8)
    mu = np.random.default_rng(5).uniform(0, 0.05, (1, Zs, N, N)).astype(np.float32)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.cone_march_f64(mu, cone.pz, cone.py, cone.qd, cone.weight.astype(np.float64))
        runs.append(time.perf_counter() - t0)
    dt = float(np.median(runs))
    scale = (N / Zs) * K
    log(f"cpu attenuation sample: median {dt:.2f}s of {len(runs)} runs for {Zs} slices of one line")
    return {"s_per_angle": dt * scale, "cores": ncores,
            "sample": f"{Zs} of {N} slices, 1 of {K} lines, one angle at {N}^3: median of 3 runs {dt:.2f}s, "
                      f"scaled x{scale:g}"}


def cpu_baseline(args):
    """The oracle (the FP32 C restatement of the reference kernels, OpenMP over
    independent outputs) on one line at full volume, scaled to all lines."""
    import torch
    ncores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(ncores))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the oracle's CPU restatement, used here as the timed CPU port

    N = args.size
    ka = min(args.cpu_angles, args.angles)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(7)
    a_np = (torch.rand((ka, N, N, N), device="cuda:0", generator=g) * 0.7 + 0.3).cpu().numpy()
    i_np = ((torch.rand((ka, N, N), device="cuda:0", generator=g) * 1.5 + 0.5) * 1e-2).cpu().numpy()
    c_np = np.ones((1, N, N, N), np.float32)
    em_np = np.full((ka, 1), 1e-4, np.float32)
    th_np = np.deg2rad(np.linspace(0, 180, ka, endpoint=False)).astype(np.float32)
    sens = oracle.backward_batch_f32(a_np, np.ones_like(i_np), em_np, th_np)  # untimed, as on the GPU
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        Pf = oracle.forward_batch_f32(a_np, c_np, em_np, th_np)
        ratio = i_np / np.maximum(Pf, np.float32(1e-6))
        back = oracle.backward_batch_f32(a_np, ratio, em_np, th_np)
        c_new = np.maximum(c_np * back / np.maximum(sens, np.float32(1e-6)), 0)
        runs.append(time.perf_counter() - t0)
        del c_new
    dt = float(np.median(runs))
    scale = (args.angles / ka) * args.elements
    log(f"cpu baseline: median {dt:.2f}s of 3 runs for one line, {ka} angles, {ncores} threads")
    return {"value": 1.0 / (dt * scale), "unit": "iter/s", "cores": ncores, "kind": "port",
            "sample": f"one line, {ka} of {args.angles} angles at {N}^3: median of 3 runs {dt:.2f}s on {ncores} threads, "
                      f"scaled x{scale:g} (lines are independent)"}


def legacy_dense_small(args, N=64, A=60, K=2, n_slices=8):
    """The reference's legacy CPU path (dense system matrix per slice, FP64
    generate_H_jit + mlem_matrix, restated in the oracle) on its own CPU-sized
    case, 64^3 x 60 angles x 2 lines, beside this library's MLEM iteration on
    the GPU at the same size.  CPU: n_slices slices of one line, median of 3
    runs, scaled to all slices and lines."""
    import torch
    from fl_correction_amd.mlem import MLEMSolver
    ncores = len(os.sched_getaffinity(0))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the oracle's legacy restatement, timed as the legacy-faithful CPU path

    rng = np.random.default_rng(20260000)
    theta = np.deg2rad(np.linspace(0, 180, A, endpoint=False))
    slices = [rng.uniform(0.3, 1.0, (A, N, N)) for _ in range(n_slices)]
    y = [rng.uniform(0.5, 2.0, A * N) for _ in range(n_slices)]
    c0 = np.ones((N, N), np.float32)
    t_h, t_it = [], []
    for _ in range(3):
        th = ti = 0.0
        for k in range(n_slices):
            t0 = time.perf_counter()
            Hm = oracle.legacy_H_f64(slices[k], theta)
            t1 = time.perf_counter()
            oracle.legacy_mlem_f64(Hm, y[k], c0, 1)
            t2 = time.perf_counter()
            oracle.legacy_mlem_f64(Hm, y[k], c0, 3)
            t3 = time.perf_counter()
            # per iteration: two more iterations; the column sums b_sum cancel
            th, ti = th + (t1 - t0), ti + ((t3 - t2) - (t2 - t1)) / 2
            del Hm
        t_h.append(th)
        t_it.append(ti)
    scale = (N / n_slices) * K
    cpu_h, cpu_it = float(np.median(t_h)) * scale, float(np.median(t_it)) * scale

    dev = torch.device("cuda:0")
    th_t = torch.tensor(theta, dtype=torch.float32, device=dev)
    em = torch.ones((A, 1), dtype=torch.float32, device=dev)
    sol = [MLEMSolver(torch.rand((A, N, N, N), device=dev) * 0.7 + 0.3, em, th_t,
                      torch.rand((A, N, N), device=dev) + 0.5) for _ in range(K)]
    C = [torch.ones((1, N, N, N), device=dev) for _ in range(K)]
    Cn = [torch.empty_like(c) for c in C]

    def it():
        for e in range(K):
            sol[e].ratio_step(C[e])
            sol[e].update_step(C[e], Cn[e])
            C[e], Cn[e] = Cn[e], C[e]

    it()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        it()
    e1.record()
    torch.cuda.synchronize()
    gpu_ms = e0.elapsed_time(e1) / 20
    log(f"legacy dense CPU path at {N}^3 x {A} x {K}: H {cpu_h:.2f}s, {cpu_it:.2f}s per iteration; "
        f"GPU {gpu_ms:.3f} ms per iteration")
    return {"workload": f"{N}^3 voxels x {A} angles x {K} lines (the reference's CPU-sized case)",
            "cpu_h_build_s": cpu_h, "cpu_s_per_iteration": cpu_it, "gpu_ms_per_iteration": gpu_ms,
            "speedup_per_iteration": cpu_it / (gpu_ms * 1e-3), "cores": ncores,
            "sample": f"{n_slices} of {N} slices of one line, median of 3 runs, scaled x{scale:g}; FP64 dense "
                      f"matrix of {A * N} x {N * N} per slice; per iteration = (3 iterations - 1 iteration) / 2"}


def torch_cpu_small(gpu_mlem_ms, n_att=10):
    """The pure-torch restatement of the model in the oracle (oracle/torch_model.py:
    the reference's own torch ops for the rotation, the incident product and the
    mask correlation; the projector as an explicit gather), run by torch on all
    host cores in FP32 at the reference's CPU-sized case: 64^3, 2 lines, 60 angles,
    3x3 cone over 10 deg.  Timed: the attenuation maps (n_att angles, scaled to
    all), and one MLEM iteration of both lines (projection, its transpose through
    autograd, the update; the sensitivity untimed, as on the GPU), each the median
    of 3 runs.  Beside them on the GPU: the model's attenuation maps of every angle
    at the same size, and the MLEM iteration of the legacy leg (same sizes)."""
    import torch
    from fl_correction_amd.workloads import PLAIN_64 as wl
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import torch_model as tm  # the oracle's pure-torch restatement, timed as the torch-CPU path

    ncores = len(os.sched_getaffinity(0))
    threads0 = torch.get_num_threads()
    torch.set_num_threads(ncores)
    K, (B, H, W) = wl.K, wl.shape
    angles = wl.angles()
    A = len(angles)
    C = torch.from_numpy(wl.phantom())
    cx, cxrf = (torch.from_numpy(c) for c in wl.coefs())
    mask = torch.from_numpy(np.asarray(wl.cone().splat_mask(W), np.float32))

    def atten_angle(ang):
        Fr = tm.rotate(C, ang)
        Ax = tm.incident(Fr * cx[:, None, None, None])
        mu = torch.einsum("jbhw,je->ebhw", Fr, cxrf)
        return tm.xrf(mu, mask) * Ax

    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        for ang in angles[:n_att]:
            atten_angle(ang)
        runs.append(time.perf_counter() - t0)
    cpu_att = float(np.median(runs)) * A / n_att

    # MLEM on attenuation of the same shape (the cost does not depend on its values)
    gen = torch.Generator().manual_seed(20260000)
    att = torch.rand((K, A, B, H, W), generator=gen) * 0.7 + 0.3

    def project_all(k, Cv):
        return torch.stack([tm.project(att[k, a], Cv, 1.0, ang) for a, ang in enumerate(angles)])

    I = [project_all(k, C[k]) for k in range(K)]
    sens = []
    for k in range(K):
        Cv = torch.zeros((B, H, W), requires_grad=True)
        P = project_all(k, Cv)
        sens.append(torch.autograd.grad(P, Cv, torch.ones_like(P))[0].clamp_min(1e-6))

    def mlem_iteration(Cs):
        out = []
        for k in range(K):
            Cv = Cs[k].clone().requires_grad_(True)
            P = project_all(k, Cv)
            back = torch.autograd.grad(P, Cv, I[k] / P.detach().clamp_min(1e-6))[0]
            out.append(Cs[k] * back / sens[k])
        return torch.stack(out)

    Cs = torch.ones((K, B, H, W))
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        Cs = mlem_iteration(Cs)
        runs.append(time.perf_counter() - t0)
    cpu_it = float(np.median(runs))
    torch.set_num_threads(threads0)

    dev = torch.device("cuda:0")
    model = wl.model()
    Cd = C.to(dev)
    model.attenuation(Cd)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    model.attenuation(Cd)
    e1.record()
    torch.cuda.synchronize()
    gpu_att_ms = e0.elapsed_time(e1)
    log(f"torch CPU at {B}^3 x {A} x {K}: attenuation {cpu_att:.2f}s (GPU {gpu_att_ms:.2f} ms), "
        f"MLEM iteration {cpu_it:.3f}s (GPU {gpu_mlem_ms:.3f} ms), {ncores} threads")
    return {"workload": f"{B}^3 voxels x {A} angles x {K} lines, {wl.cone_n}x{wl.cone_n} cone over "
                        f"{wl.cone_half:g} deg (the reference's CPU-sized case)",
            "cpu_attenuation_s": cpu_att, "gpu_attenuation_ms": gpu_att_ms,
            "speedup_attenuation": cpu_att / (gpu_att_ms * 1e-3),
            "cpu_s_per_iteration": cpu_it, "gpu_ms_per_iteration": gpu_mlem_ms,
            "speedup_per_iteration": cpu_it / (gpu_mlem_ms * 1e-3), "cores": ncores, "dtype": "f32",
            "sample": f"torch {torch.__version__} on the host, FP32, {ncores} threads, median of 3 runs; attenuation "
                      f"{n_att} of {A} angles scaled x{A / n_att:g} (mask correlation with the cone's splat mask); "
                      f"MLEM one iteration of both lines, the transpose through autograd"}


if __name__ == "__main__":
    main()
