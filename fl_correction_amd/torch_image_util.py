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
    math.cos of the (float32) angle, rounded to float32 (FL/torch_image_util.py:77-85)."""
    th = _angle_value(theta)
    return float(np.float32(math.cos(th))), float(np.float32(math.sin(th)))


def torch_rotate_image_nd(x, theta, enable_crop=True, align_corners=False, mode="bilinear", padding_mode="zeros"):
    """Rotate (..., H, W) by theta radians (positive = counter-clockwise) about the centre
    (FL/torch_image_util.py:13-114)."""
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
    """Quarter turn of a 3-D tensor about axis `ax` (FL/torch_image_util.py:117-147).

    With the other two axes (p, q) in increasing order, 'c-clock' and 'clock' are
    torch.rot90 by +1 and -1 quarter turns in the plane (p, q), except about axis 2,
    where the reference turns the other way (its positive direction runs left to
    right).  Any other mode only swaps p and q.  The index identities, for
    img (B, H, W) about axis 0:
        clock    out[b, i, j] = img[b, H-1-j, i]
        c-clock  out[b, i, j] = img[b, j, W-1-i]
    """
    if ax not in _ROT90_PLANE:
        return img_raw.clone()
    plane = _ROT90_PLANE[ax]
    turns = _ROT90_TURNS.get(mode)
    if turns is None:
        return img_raw.transpose(*plane).clone()
    return torch.rot90(img_raw, turns if ax != 2 else -turns, plane)


# the plane each axis turns, and the quarter turns of each mode
_ROT90_PLANE = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
_ROT90_TURNS = {"c-clock": 1, "clock": -1}


def torch_rot90_4D(img4D, ax=0, mode="c-clock"):
    """torch_rot90_3D of every element of a 4-D stack at once (FL/torch_image_util.py:150-155)."""
    if ax not in _ROT90_PLANE:
        return img4D.clone()
    plane = tuple(p + 1 for p in _ROT90_PLANE[ax])
    turns = _ROT90_TURNS.get(mode)
    if turns is None:
        return img4D.transpose(*plane).clone()
    return torch.rot90(img4D, turns if ax != 2 else -turns, plane)
