"""Shared helpers for the test suite (inputs, error metrics, oracle import)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402  (oracle/oracle.py: the CPU oracle)

RTOL = 1e-5  # relative L2 bound of every FP32 result against the same-order FP32 oracle


def rel_l2(x, y):
    return oracle.rel_l2(x, y)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def projector_inputs(A, B, R, H, W, seed, thetas=None):
    rng = np.random.default_rng(seed)
    atten = rng.uniform(0.05, 1.0, size=(A, B, H, W)).astype(np.float32)
    C = rng.uniform(0.0, 1.0, size=(R, B, H, W)).astype(np.float32)
    em = rng.uniform(0.5, 2.0, size=(A, R)).astype(np.float32)
    if thetas is None:
        theta = rng.uniform(0, np.pi, size=A).astype(np.float32)
    else:
        theta = np.asarray(thetas, np.float32)
    ratio = rng.uniform(0.2, 3.0, size=(A, B, W)).astype(np.float32)
    return atten, C, em, theta, ratio


# ---------------------------------------------------------------------------
# The differentiable model's 32^3 crop case: 4 lines, 5x5 cone over 15 deg, the
# 256^3 workload's phantom cut at MODEL_CROP_BOX, 12 angles over 180 deg
# ---------------------------------------------------------------------------
MODEL_CROP_BOX = ((80, 112), (128, 160), (32, 64))
MODEL_CROP_ANGLES = 12
MODEL_CROP_FD_STEP = 1e-4


def model_crop_case():
    """Inputs of the crop case: dict with C_true, C_eval (the point where the gradient is
    taken: 0.5*C_true + 0.05, positive everywhere), angles, coef_x (A,K), coef_xrf (A,K,K),
    em (A,K), cone, param, I (noisy projections of C_true from the FP64 model)."""
    sys.path.insert(0, ROOT)
    from fl_correction_amd import workloads
    W = workloads.MODEL_256
    C_true = W.phantom(box=MODEL_CROP_BOX)
    A = MODEL_CROP_ANGLES
    angles = np.linspace(0.0, 180.0, A, endpoint=False)
    cx, cf = W.coefs()
    coef_x, coef_xrf = np.tile(cx, (A, 1)), np.tile(cf, (A, 1, 1))
    em = np.full((A, W.K), workloads.PIX, np.float32)
    cone = W.cone()
    Pf = oracle.model_forward_f64(C_true, angles, coef_x, coef_xrf, em, cone)
    I = workloads.noisy(Pf, seed=W.seed + 7)
    C_eval = (0.5 * C_true + 0.05).astype(np.float32)
    return dict(C_true=C_true, C_eval=C_eval, angles=angles, coef_x=coef_x, coef_xrf=coef_xrf, em=em, cone=cone,
                param=W.param(), I=I, K=W.K)


def poisson_loss_f64(Pf, I):
    """mean(Pf - I*log Pf) (FL/ml_lib/ml_loss.py:6-8; Pf > 0 here, so the clamp is idle)."""
    return float(np.mean(Pf - I * np.log(Pf)))
