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
