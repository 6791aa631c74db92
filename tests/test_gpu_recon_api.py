"""The numpy-level drivers against the oracle: torch_recon.torch_mlem_recon[_batch]
(FL/torch_recon.py:5-98) and core.re_projection_cuda (FL/core.py:104-156).

They only convert their arguments (degrees -> radians in float64 then float32,
em_cs scaled by rho*pix, 1-D em_cs -> a column, 3-D inputs promoted to the
slice-batch form, defaults of ones) and call the library, so each test builds
the converted arrays by hand, runs the oracle on them and asks for the bar of
the kernels themselves: relative L2 <= 1e-5 against the FP32 oracle.
"""
import numpy as np
import pytest

from helpers import RTOL, oracle, rel_l2

pytestmark = pytest.mark.gpu


def _theta(angles_deg):
    return (np.asarray(angles_deg) / 180.0 * np.pi).astype(np.float32)


def _case(A, R, B, H, W, seed):
    rng = np.random.default_rng(seed)
    angles = np.sort(rng.uniform(0.0, 180.0, A))
    atten = rng.uniform(0.3, 1.0, (A, B, H, W)).astype(np.float32)
    C_true = rng.uniform(0.0, 1.0, (R, B, H, W)).astype(np.float32)
    return angles, atten, C_true


def test_torch_mlem_recon_single_slice():
    from fl_correction_amd.torch_recon import torch_mlem_recon
    A, R, H, W = 18, 1, 36, 44
    angles, atten, C_true = _case(A, R, 1, H, W, seed=51)
    em_1d = np.random.default_rng(52).uniform(0.5, 2.0, A)  # 1-D: one reference
    rho, pix = 2.0, 0.25
    em = (em_1d[:, None] * rho * pix).astype(np.float32)
    I = oracle.forward_batch_f32(atten, C_true, em, _theta(angles))[:, 0]  # (A, W)
    C0 = np.ones((R, H, W), np.float32)
    print(f"torch_mlem_recon A={A} H={H} W={W}", flush=True)
    got = torch_mlem_recon(C0, I, angles, atten=atten[:, 0], em_cs=em_1d, rho=rho, pix=pix, n_iter=6, beta=0.0)
    ref = oracle.mlem_batch_f32(C0[:, None], atten, em, _theta(angles), I[:, None], 6, 0.0, 0.01)[:, 0]
    assert got.shape == (R, H, W)
    assert rel_l2(got, ref) <= RTOL


def test_torch_mlem_recon_defaults_two_references():
    """atten None -> ones, em_cs None -> ones (n_angle, n_ref), the reference's
    default OSL prior beta = 1e-3, delta = 0.01."""
    from fl_correction_amd.torch_recon import torch_mlem_recon
    A, R, H, W = 12, 2, 30, 30
    angles, _, C_true = _case(A, R, 1, H, W, seed=53)
    ones = np.ones((A, 1, H, W), np.float32)
    em = np.ones((A, R), np.float32)
    I = oracle.forward_batch_f32(ones, C_true, em, _theta(angles))[:, 0]
    C0 = np.full((R, H, W), 0.5, np.float32)
    got = torch_mlem_recon(C0, I, angles, n_iter=4)
    ref = oracle.mlem_batch_f32(C0[:, None], ones, em, _theta(angles), I[:, None], 4, 1e-3, 0.01)[:, 0]
    assert rel_l2(got, ref) <= RTOL


@pytest.mark.parametrize("three_d", [False, True])
def test_torch_mlem_recon_batch(three_d):
    from fl_correction_amd.torch_recon import torch_mlem_recon_batch
    A, R = 15, 1
    B = 1 if three_d else 3
    H, W = 40, 52
    angles, atten, C_true = _case(A, R, B, H, W, seed=54 + int(three_d))
    em_2d = np.random.default_rng(56).uniform(0.5, 2.0, (A, R))
    rho, pix = 1.5, 0.5
    em = (em_2d * rho * pix).astype(np.float32)
    I = oracle.forward_batch_f32(atten, C_true, em, _theta(angles))  # (A, B, W)
    C0 = np.ones((R, B, H, W), np.float32)
    print(f"torch_mlem_recon_batch three_d={three_d}", flush=True)
    if three_d:  # (n_ref, H, W), (n_angle, W), (n_angle, H, W): promoted to one slice
        got = torch_mlem_recon_batch(C0[:, 0], I[:, 0], angles, atten=atten[:, 0], em_cs=em_2d, rho=rho, pix=pix,
                                     n_iter=5, beta=1e-3)
    else:
        got = torch_mlem_recon_batch(C0, I, angles, atten=atten, em_cs=em_2d, rho=rho, pix=pix, n_iter=5, beta=1e-3)
    ref = oracle.mlem_batch_f32(C0, atten, em, _theta(angles), I, 5, 1e-3, 0.01)
    assert got.shape == (R, B, H, W)
    assert rel_l2(got, ref) <= RTOL


@pytest.mark.parametrize("use_ref", [False, True])
def test_re_projection_cuda(use_ref):
    """em = em_cs[elem] * pix (or em_cs[elem_ref{i}] per reference), C = img * rho_compound."""
    from fl_correction_amd.core import re_projection_cuda
    A = 10
    R = 2 if use_ref else 1
    B, H, W = 3, 34, 46
    angles, atten, C_true = _case(A, R, B, H, W, seed=57 + int(use_ref))
    rng = np.random.default_rng(59)
    pix, rho_c = 0.3, 1.7
    if use_ref:
        cs = {f"Ni_ref{i + 1}": rng.uniform(0.5, 2.0, A) for i in range(R)}
        em = np.stack([cs[f"Ni_ref{i + 1}"] * pix for i in range(R)], axis=1)
        img = C_true
    else:
        cs = {"Ni": rng.uniform(0.5, 2.0, A)}
        em = (cs["Ni"] * pix)[:, None]
        img = C_true[0]  # (sli, H, W)
    param = {"pix": pix, "em_cs": cs}
    got = re_projection_cuda(img, angles, param=param, rho_compound=rho_c, atten_coef=atten, elem="Ni",
                             use_ref=use_ref)
    ref = oracle.forward_batch_f32(atten, (C_true * rho_c).astype(np.float32), em.astype(np.float32),
                                   _theta(angles))
    assert got.shape == (A, B, W)
    assert rel_l2(got, ref) <= RTOL


def test_re_projection_cuda_no_param_no_atten():
    """param None -> em = 1, atten None -> ones."""
    from fl_correction_amd.core import re_projection_cuda
    A, B, H, W = 8, 2, 24, 24
    angles, _, C_true = _case(A, 1, B, H, W, seed=60)
    got = re_projection_cuda(C_true[0], angles)
    ones = np.ones((A, B, H, W), np.float32)
    ref = oracle.forward_batch_f32(ones, C_true, np.ones((A, 1), np.float32), _theta(angles))
    assert rel_l2(got, ref) <= RTOL
