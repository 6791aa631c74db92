"""Pins the pure-torch reference model of the attenuated forward model (the
reference for the hand-written adjoint): FP64 central finite differences
(step 1e-4) against its autograd gradient, and its forward against the oracle's
FP32 same-order kernels."""
import numpy as np
import pytest
import torch

from helpers import oracle, rel_l2
import torch_model as tm  # oracle/torch_model.py


def setup(seed=0, K=2, B=4, N=12, A=3):
    rng = np.random.default_rng(seed)
    C = rng.uniform(0, 1, (K, B, N, N))
    angles = np.array([0.0, 40.0, 115.0])[:A]
    coef_x = rng.uniform(0.01, 0.04, (A, K)).astype(np.float32)
    coef_xrf = rng.uniform(0.005, 0.03, (A, K, K)).astype(np.float32)
    em = rng.uniform(0.5, 1.5, (A, K)).astype(np.float32)
    mask = oracle.generate_detector_mask(20, 20, N)
    return C, angles, coef_x, coef_xrf, em, mask


def test_finite_differences():
    C, angles, cx, cf, em, mask = setup()
    Rw = np.random.default_rng(1).uniform(0.5, 1.5, (C.shape[0], len(angles), C.shape[1], C.shape[3]))
    Ct = torch.tensor(C, dtype=torch.float64, requires_grad=True)
    loss = (tm.forward(Ct, angles, cx, cf, em, mask) * torch.tensor(Rw)).sum()
    loss.backward()
    g = Ct.grad.numpy()
    rng = np.random.default_rng(2)
    idx = [tuple(int(rng.integers(0, n)) for n in C.shape) for _ in range(16)]
    step = 1e-4
    for ix in idx:
        Cp, Cm = C.copy(), C.copy()
        Cp[ix] += step
        Cm[ix] -= step
        with torch.no_grad():
            lp = float((tm.forward(torch.tensor(Cp), angles, cx, cf, em, mask) * torch.tensor(Rw)).sum())
            lm = float((tm.forward(torch.tensor(Cm), angles, cx, cf, em, mask) * torch.tensor(Rw)).sum())
        fd = (lp - lm) / (2 * step)
        assert abs(fd - g[ix]) <= 1e-6 * max(1.0, abs(fd)), (ix, fd, g[ix])


def test_forward_matches_oracle_composition():
    C, angles, cx, cf, em, mask = setup(seed=3)
    Pf = tm.forward(torch.tensor(C), angles, cx, cf, em, mask).numpy()
    K = C.shape[0]
    for a, ang in enumerate(angles):
        rot = oracle.rotate_torch(C.astype(np.float32), np.float32(ang / 180 * np.pi))
        Ax = oracle.incident_f32(oracle.mu_xray(rot, cx[a]))
        Af = oracle.xrf_atten_from_mask(oracle.mu_xrf(rot, cf[a]), mask)
        At = Af * Ax
        th = np.array([ang / 180 * np.pi], np.float32)
        for e in range(K):
            ref = oracle.forward_batch_f32(At[e][None], C[e][None].astype(np.float32), em[a:a + 1, e:e + 1], th)[0]
            assert rel_l2(Pf[e, a], ref) < 1e-5
