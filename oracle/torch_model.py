"""Pure-torch restatement of the attenuated XRF forward model -- TEST INFRASTRUCTURE ONLY.

Differentiable end to end with torch autograd (run in float64 on the CPU), so
its gradient is the reference for the hand-written adjoint.  Pinned by FP64
central differences and by the oracle's kernels (tests/test_torch_model.py).  Each step is the
reference's own tensor code path:
  rotation     F.affine_grid + F.grid_sample            FL/torch_image_util.py:77-104
  incident     1 - mu_x, flip, cumprod, flip           FL/torch_functions.py:92-102
  fluorescence rot90 'clock', mask correlation, rot90 back, exp(-.)
                                                       FL/torch_functions.py:140-174 (atten_kernel as conv3d)
  product      A_xrf * A_x                              FL/core.py:99
  projector    clamp-bilinear sample of C at the rotated detector grid, times
               atten, summed over rows                 FL/cuda_lib/atten_cuda_kernel.cu:298-337
               (written as an explicit gather: grid_sample's border padding is
               not the kernel's clamp to H-1.001)
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F


def rotate(x, angle_deg):
    """torch_rotate_image_nd defaults on (K, B, H, W)."""
    K, B, H, W = x.shape
    th = float(np.float32(angle_deg / 180.0 * np.pi))
    c, s = float(np.float32(math.cos(th))), float(np.float32(math.sin(th)))
    mat = torch.tensor([[c, -s, 0.0], [s, c, 0.0]], dtype=x.dtype).unsqueeze(0).repeat(K * B, 1, 1)
    grid = F.affine_grid(mat, size=(K * B, 1, H, W), align_corners=False)
    y = F.grid_sample(x.reshape(K * B, 1, H, W), grid, mode="bilinear", padding_mode="zeros", align_corners=False)
    return y.reshape(K, B, H, W)


def incident(mux):
    f = 1 - mux
    cp = torch.flip(torch.cumprod(torch.flip(f, dims=(2,)), dim=2), dims=(2,))
    out = torch.ones_like(mux)
    return torch.cat([cp[:, :, 1:], out[:, :, -1:]], dim=2)


def mask_correlation(mu, mask):
    """atten_kernel on mu (K, Z, X, Y) as a conv3d (cross-correlation) with zero padding.
    A mask with few non-zero taps (a cone's splat mask) is applied tap by tap instead:
    the same sum, as shifted multiply-adds of the padded volume."""
    Mz, Mx, My = mask.shape
    hz, hy = Mz // 2, My // 2
    t = F.pad(mu[:, None], (hy, My - 1 - hy, 0, Mx - 1, hz, Mz - 1 - hz))
    nz = torch.nonzero(mask)
    if len(nz) > 0.25 * mask.numel():
        return F.conv3d(t, mask[None, None])[:, 0]
    K, Z, X, Y = mu.shape
    t = t[:, 0]
    out = torch.zeros_like(mu)
    for dz, dx, dy in nz.tolist():
        out = out + mask[dz, dx, dy] * t[:, dz:dz + Z, dx:dx + X, dy:dy + Y]
    return out


def xrf(mu, mask, side="r"):
    """exp(-(mask (*) mu)) in the original frame; mu (K, B, H, W)."""
    if side == "r":
        mr = torch.flip(torch.transpose(mu, 2, 3), dims=(3,))     # rot90 'clock'
        T = mask_correlation(mr, mask)
        T = torch.flip(torch.transpose(T, 2, 3), dims=(2,))       # rot90 'c-clock'
    else:
        mr = torch.flip(torch.transpose(mu, 2, 3), dims=(2,))     # rot90 'c-clock'
        T = mask_correlation(mr, mask)
        T = torch.flip(torch.transpose(T, 2, 3), dims=(3,))       # rot90 'clock'
    return torch.exp(-T)


def project(atten, C, em, angle_deg):
    """Pf (B, W) = em * sum_h atten[b,h,w] * S(C[b])(h,w) for one angle and one line."""
    B, H, W = C.shape
    th = float(np.float32(angle_deg / 180.0 * np.pi))
    c, s = math.cos(-th), math.sin(-th)
    cx, cy = (W - 1) * 0.5, (H - 1) * 0.5
    h = torch.arange(H, dtype=torch.float64)[:, None]
    w = torch.arange(W, dtype=torch.float64)[None, :]
    x = (c * (h - cy) - s * (w - cx) + cy).clamp(0, H - 1.001)
    y = (s * (h - cy) + c * (w - cx) + cx).clamp(0, W - 1.001)
    x0, y0 = torch.floor(x).long(), torch.floor(y).long()
    fx, fy = (x - x0).to(C.dtype), (y - y0).to(C.dtype)
    Cf = C.reshape(B, H * W)
    g = lambda r, q: Cf[:, (r * W + q).reshape(-1)].reshape(B, H, W)
    val = ((1 - fx) * (1 - fy) * g(x0, y0) + (1 - fx) * fy * g(x0, y0 + 1) + fx * (1 - fy) * g(x0 + 1, y0)
           + fx * fy * g(x0 + 1, y0 + 1))
    return em * (atten * val).sum(dim=1)


def forward(C, angles_deg, coef_x, coef_xrf, em, mask, side="r"):
    """Pf (K, A, B, W) for C (K, B, H, W).  coef_x (A, K), coef_xrf (A, K, K), em (A, K)."""
    K = C.shape[0]
    out = []
    for a, ang in enumerate(angles_deg):
        Fr = rotate(C, ang)
        Ax = incident(Fr * torch.as_tensor(coef_x[a], dtype=C.dtype)[:, None, None, None])
        mu = torch.einsum("jbhw,je->ebhw", Fr, torch.as_tensor(coef_xrf[a], dtype=C.dtype))
        At = xrf(mu, torch.as_tensor(mask, dtype=C.dtype), side) * Ax
        out.append(torch.stack([project(At[e], C[e], float(em[a, e]), ang) for e in range(K)]))
    return torch.stack(out, dim=1)
