"""Element fractions for the self-consistent correction loop, on the GPU.

``cal_frac`` follows FL/core.py:159-211 and ``adaptive_threshold`` follows
FL/image_util.py:337-375, with the reference's argument names and defaults.
The sum over elements, the multi-Otsu threshold, hole filling, dilation and
erosion, and the per-element fraction arithmetic run as HIP kernels
(csrc/frac.hip); the quantile picks sort on the device.

Types follow the reference's numpy arithmetic: the element sum and the
threshold run in the input's float type, and everything after the float64 mask
multiplies in (img_sum, ratio, the fractions) is float64.  numpy inputs give
numpy outputs (the drop-in form); CUDA tensors give CUDA tensors.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

def _as_device(x):
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            raise RuntimeError("cal_frac: tensors must be on a CUDA device")
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    if t.dtype not in (torch.float32, torch.float64):
        t = t.double()
    return t.contiguous()


def _f64_flag(t):
    return 1 if t.dtype == torch.float64 else 0


def adaptive_threshold(img, fill_hole=False, dilation=0, erosion=0):
    """Mask (float64 0/1) of a 2-D image or a stack of 2-D images: img >= the lowest
    three-class multi-Otsu threshold, holes filled (5x5 structure), then dilated and
    eroded with the 4-neighbour cross (FL/image_util.py:337-375).  An image whose
    threshold fails (fewer than three distinct histogram bins, or a non-finite
    value) gets an all-zero mask, as the reference's except branch sets it.
    Returns a CUDA tensor (numpy for numpy input)."""
    numpy_in = not isinstance(img, torch.Tensor)
    t = _as_device(img)
    if t.dim() not in (2, 3):
        raise ValueError("image shape not recognized")
    stack = t if t.dim() == 3 else t[None]
    n, H, W = stack.shape
    mask, _, _ = _adaptive_mask(stack, fill_hole, dilation, erosion)
    mask = mask if t.dim() == 3 else mask[0]
    return mask.cpu().numpy() if numpy_in else mask


def _adaptive_mask(stack, fill_hole, dilation, erosion):
    n, H, W = stack.shape
    dev = stack.device
    mask = torch.empty((n, H, W), dtype=torch.float64, device=dev)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    thr = torch.empty(n, dtype=stack.dtype, device=dev)
    with torch.cuda.device(dev):
        _lib.call("fl_adaptive_mask", _lib.ptr(stack), _lib.ptr(mask), n, H, W, int(bool(fill_hole)), int(dilation),
                  int(erosion), _f64_flag(stack), _lib.ptr(status), _lib.ptr(thr), _lib.stream_of(dev))
    return mask, status, thr


def _median_sorted(v):
    """np.median of an already sorted 1-D tensor (the mean of the two middle values
    for an even length)."""
    n = v.numel()
    if n % 2 == 1:
        return v[n // 2]
    return (v[n // 2 - 1] + v[n // 2]) / 2


def _pre_treat(args):
    """pre_treat (FL/util.py:75-118) for several 3-D arrays or one 4-D array:
    stacked to float64 with NaN and inf set to 0."""
    ts = [_as_device(a) for a in args]
    if len(ts) == 1:
        data = ts[0].double()
        if data.dim() == 3:
            data = data[None]
    else:
        data = torch.stack([t.double() for t in ts])
    return torch.nan_to_num(data, nan=0.0, posinf=0.0, neginf=0.0).contiguous()


def cal_frac(*args, scale_range=[0.02, 0.98], enable_scale=True, scale_limit=0.95):  # noqa: B006 (reference default)
    """Fraction of each element in the element sum, masked and rescaled
    (FL/core.py:159-211).  One argument: a stack (n_elem, H, W), a single (H, W)
    image, or (n_elem, B, H, W); several arguments are stacked as pre_treat does.
    Returns {'img_sum', 'frac', 'pix_median', 'pix_max', 'img_mask'}."""
    numpy_in = not any(isinstance(a, torch.Tensor) for a in args)
    if len(args) == 1:
        img = _as_device(args[0])
        if img.dim() == 2:
            img = img[None]
    else:
        img = _pre_treat(args)
    img = img.contiguous()
    n = img.shape[0]
    plane = tuple(img.shape[1:])
    npix = int(np.prod(plane))
    dev = img.device
    st = _lib.stream_of(dev)
    with torch.cuda.device(dev):
        # img_sum = np.sum(img, axis=0) + 1e-12, in the input's type
        s0 = torch.empty(plane, dtype=img.dtype, device=dev)
        _lib.call("fl_frac_sum", _lib.ptr(img), _lib.ptr(s0), n, npix, _f64_flag(img), st)
        # adaptive_threshold(img_sum, fill_hole=True, dilation=2): per slice for a 3-D sum
        stack = s0 if s0.dim() == 3 else s0[None]
        if stack.dim() != 3:
            raise ValueError("image shape not recognized")
        mask, _, _ = _adaptive_mask(stack, True, 2, 0)
        mask = mask.reshape(plane)
        mask_is_one = bool(mask.max().item() == 0)  # fails in adaptive_threshold: img_mask = 1
        # img_sum * img_mask: float64 with the float64 mask; with the integer 1 the
        # sum keeps the input's type, and so does everything computed from it
        img_sum = s0 if mask_is_one else s0.double() * mask
        single = 1 if (mask_is_one and img.dtype == torch.float32) else 0
        pos = img_sum[img_sum > 0]
        sum_sort, _ = torch.sort(pos)
        l = sum_sort.numel()
        if l == 0:
            raise IndexError("cal_frac: the masked element sum has no positive value")
        pix_max = sum_sort[int(l * scale_limit)]
        ratio = img_sum / pix_max if enable_scale else None
        pix_median = _median_sorted(sum_sort)
        img_sum_d = img_sum.double().contiguous()

        tmp = torch.empty((n,) + plane, dtype=torch.float64, device=dev)
        _lib.call("fl_frac_ratio", _lib.ptr(img), _lib.ptr(img_sum_d), _lib.ptr(tmp), n, npix, _f64_flag(img),
                  single, st)
        # clip bounds per element: the 2% and 98% picks of its positive values
        lohi = torch.tensor([[1.0, 0.0]] * n, dtype=torch.float64, device=dev)
        if len(scale_range):
            for i in range(n):
                ts, _ = torch.sort(tmp[i][tmp[i] > 0])
                m = ts.numel()
                if m > 100:
                    sel = ts[int(m * scale_range[0]):int(m * scale_range[1])]
                    lohi[i, 0] = sel[0]
                    lohi[i, 1] = sel[-1]
        frac = torch.empty_like(tmp)
        ratio_d = None if ratio is None else ratio.double().contiguous()
        _lib.call("fl_frac_final", _lib.ptr(tmp), None if mask_is_one else _lib.ptr(mask),
                  None if ratio_d is None else _lib.ptr(ratio_d), _lib.ptr(lohi), _lib.ptr(frac), n, npix, single, st)
    res = {"img_sum": img_sum, "frac": frac, "pix_median": pix_median, "pix_max": pix_max,
           "img_mask": 1 if mask_is_one else mask}
    if numpy_in:
        out = {}
        for k, v in res.items():
            if isinstance(v, torch.Tensor):
                out[k] = v.cpu().numpy()[()]  # 0-d values as numpy scalars
            else:
                out[k] = v
        return out
    return res
