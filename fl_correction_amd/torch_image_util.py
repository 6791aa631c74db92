"""Drop-in for FL/torch_image_util.py: in-plane rotation and rot90 frame changes.

``torch_rotate_image_nd`` runs the HIP rotation kernel (fl_rotate) for the
options the hot path uses (the reference's defaults: crop to the input size,
align_corners=False, bilinear, zero padding).  Other options belong to the
detector-offset branch of torch_xrf_atten_3D, which the reference itself cannot
run (its crop arithmetic is broken, FL/torch_functions.py:161), and raise.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib


def _angle_value(theta):
    if isinstance(theta, torch.Tensor):
        return float(theta.detach().cpu().to(torch.float32))
    return float(theta)


def rotation_entries(theta):
    """float32 (cos, sin) entries of the affine matrix, computed as the reference does:
    math.cos of the (float32) angle, rounded to float32 (torch_image_util.py:401-409)."""
    th = _angle_value(theta)
    return float(np.float32(math.cos(th))), float(np.float32(math.sin(th)))


def torch_rotate_image_nd(x, theta, enable_crop=True, align_corners=False, mode="bilinear", padding_mode="zeros"):
    """Rotate (..., H, W) by theta radians (positive = counter-clockwise) about the centre
    (FL/torch_image_util.py:337-438)."""
    if not (enable_crop and not align_corners and mode == "bilinear" and padding_mode == "zeros"):
        raise NotImplementedError("only enable_crop=True, align_corners=False, bilinear, zeros are implemented")
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise RuntimeError("x must be a CUDA tensor")
    if x.dim() < 2:
        raise ValueError("Input must have at least 2 dimensions")
    if x.dtype != torch.float32:
        raise RuntimeError("x must be float32")
    xc = x.contiguous()
    H, W = xc.shape[-2:]
    N = xc.numel() // max(H * W, 1)
    out = torch.empty_like(xc)
    c, s = rotation_entries(theta)
    with torch.cuda.device(xc.device):
        _lib.call("fl_rotate", _lib.ptr(xc), _lib.ptr(out), N, H, W, c, s, _lib.stream_of(xc.device))
    return out


def torch_rot90_3D(img_raw, ax=0, mode="c-clock"):
    """90-degree rotation of a 3-D tensor about axis `ax` (FL/torch_image_util.py:441-471):
    a transpose and a flip, as in the reference."""
    img_r = img_raw.clone()
    if ax == 0:
        img_r = torch.transpose(img_r, 1, 2)
        if mode == "clock":
            img_r = img_r.flip(dims=(2,))
        elif mode == "c-clock":
            img_r = img_r.flip(dims=(1,))
    if ax == 1:
        img_r = torch.transpose(img_r, 0, 2)
        if mode == "clock":
            img_r = img_r.flip(dims=(2,))
        elif mode == "c-clock":
            img_r = img_r.flip(dims=(0,))
    if ax == 2:
        img_r = torch.transpose(img_r, 0, 1)
        if mode == "clock":
            img_r = img_r.flip(dims=(0,))
        elif mode == "c-clock":
            img_r = img_r.flip(dims=(1,))
    return img_r


def torch_rot90_4D(img4D, ax=0, mode="c-clock"):
    """torch_rot90_3D applied to each element of a 4-D stack (FL/torch_image_util.py:474-479)."""
    return torch.stack([torch_rot90_3D(img4D[i], ax, mode) for i in range(img4D.shape[0])])
