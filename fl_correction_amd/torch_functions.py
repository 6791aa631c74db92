"""Drop-in for the attenuation and projection helpers of FL/torch_functions.py.

torch_xray_atten_3D  -> fl_incident_atten (HIP)
torch_xrf_atten_3D   -> fl_mu_xrf_rot + (fl_cal_atten | fl_cone_sweep | fl_cone_atten) + fl_xrf_combine (HIP)
torch_3D_forward_projection -> the projector op set

``detector_mask`` may be a dense mask array (the reference's form: correlated
exactly as atten_kernel does) or a :class:`~fl_correction_amd.detector_mask.Cone`
(direction-sampled cone, marched by the exact rational-step recursion).
``torch_extract_3D_mu_*`` return the linear attenuation coefficients with the
reference's own elementwise tensor arithmetic; the HIP pipeline computes them
inside its kernels and does not call these.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .atten_util import coef_xray, coef_xrf
from .cuda_lib.atten_cuda import forward_emission, forward_emission_batch
from .detector_mask import Cone


def _as_device_f32(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float32, device=device)


def torch_extract_3D_mu_xray(frac_4D, param_angle_energy, dict_use_ref_rotate={}):
    """mu_x[j] = frac[j] * cs[comp_j-x] * rho * pix (FL/torch_functions.py:8-52)."""
    if len(dict_use_ref_rotate) > 0 and dict_use_ref_rotate.get("use_ref", False):
        raise NotImplementedError("reference-spectrum (use_ref) attenuation is not supported")
    coef = torch.as_tensor(coef_xray(param_angle_energy), device=frac_4D.device, dtype=frac_4D.dtype)
    return frac_4D * coef.view(-1, *([1] * (frac_4D.dim() - 1)))


def torch_extract_3D_mu_xrf(frac_4D, param_angle_energy):
    """mu[i] = sum_j frac[j] * cs[comp_j-elem_i] * rho * pix (FL/torch_functions.py:55-81)."""
    coef = coef_xrf(param_angle_energy)
    mu = torch.zeros((coef.shape[1],) + tuple(frac_4D.shape[1:]), dtype=frac_4D.dtype, device=frac_4D.device)
    for i in range(coef.shape[1]):
        for j in range(coef.shape[0]):
            mu[i] += frac_4D[j] * float(coef[j, i])
    return mu


def incident_atten(frac_4D, coef_x, out=None):
    """Device incident-beam attenuation (K, B, H, W) from frac (K, B, H, W) and coef_x (K,)."""
    frac = _as_device_f32(frac_4D, frac_4D.device if isinstance(frac_4D, torch.Tensor) else "cuda")
    K, B, H, W = frac.shape
    coef = _as_device_f32(coef_x, frac.device)
    if out is None:
        out = torch.empty_like(frac)
    with torch.cuda.device(frac.device):
        _lib.call("fl_incident_atten", _lib.ptr(frac), _lib.ptr(coef), _lib.ptr(out), K, B, H, W, B * H * W,
                  out.stride(0), _lib.stream_of(frac.device))
    return out


def torch_xray_atten_3D(frac_4D, param_angle_energy, dict_use_ref_rotate={}, flag_detach=True):
    """Incident attenuation prod_{m>r} (1 - mu_x) along rows, last row 1
    (FL/torch_functions.py:83-106).  Returns numpy when flag_detach, as the reference."""
    if len(dict_use_ref_rotate) > 0 and dict_use_ref_rotate.get("use_ref", False):
        raise NotImplementedError("reference-spectrum (use_ref) attenuation is not supported")
    if not (isinstance(frac_4D, torch.Tensor) and frac_4D.is_cuda):
        raise RuntimeError("frac_4D must be a CUDA tensor")
    out = incident_atten(frac_4D, coef_xray(param_angle_energy))
    return out.detach().cpu().numpy() if flag_detach else out


def _side(position_det):
    if position_det == "r":
        return 0
    if position_det == "l":
        return 1
    raise ValueError("position_det must be 'r' or 'l'")


def xrf_exponent_rot(frac, coef, detector_mask, side=0, work=None):
    """Line integral T in the mask frame, (K, B, W, H), for frac (Kc, B, H, W) on the device."""
    Kc, B, H, W = frac.shape
    K = coef.shape[1]
    dev = frac.device
    coef_d = _as_device_f32(coef, dev)
    mu_rot = work if work is not None else torch.empty((K, B, W, H), dtype=torch.float32, device=dev)
    T = torch.empty((K, B, W, H), dtype=torch.float32, device=dev)
    st = _lib.stream_of(dev)
    with torch.cuda.device(dev):
        _lib.call("fl_mu_xrf_rot", _lib.ptr(frac), _lib.ptr(coef_d), _lib.ptr(mu_rot), Kc, K, B, H, W, side, st)
        if isinstance(detector_mask, Cone):
            cone = detector_mask
            arrs = [np.ascontiguousarray(a) for a in (cone.pz, cone.py, cone.qd, cone.weight)]
            ptrs = [a.ctypes.data_as(_lib.ctypes.c_void_p) for a in arrs]
            # all directions in one sweep where it applies, else direction by direction
            nb = int(_lib.load().fl_cone_sweep_workspace_bytes(K, B, W, H, *ptrs[:3], cone.ndir))
            if nb:
                ws = torch.empty((nb + 3) // 4, dtype=torch.float32, device=dev)
                _lib.call("fl_cone_sweep", _lib.ptr(mu_rot), _lib.ptr(T), K, B, W, H, *ptrs, cone.ndir, 0,
                          _lib.ptr(ws), ws.numel() * 4, st)
            else:
                _lib.call("fl_cone_atten", _lib.ptr(mu_rot), _lib.ptr(T), K, B, W, H, *ptrs, cone.ndir, st)
        else:
            mask = _as_device_f32(detector_mask, dev)
            if mask.dim() != 3:
                raise RuntimeError("detector_mask must be 3-D (Mz, Mx, My) or a Cone")
            Mz, Mx, My = mask.shape
            _lib.call("fl_cal_atten", _lib.ptr(mu_rot), _lib.ptr(mask), _lib.ptr(T), K, B, W, H, Mz, Mx, My, st)
    return T


def torch_xrf_atten_3D(frac_4D, param_angle_energy, detector_mask, detector_offset_angle=0, position_det="r",
                       return_numpy=True):
    """Fluorescence-path attenuation exp(-(M (*) mu)) for every element line
    (FL/torch_functions.py:110-176).  frac_4D (K, B, H, W) is the already-rotated
    concentration.  Returns numpy (the reference's return type) unless return_numpy=False."""
    if abs(float(detector_offset_angle)) / 180.0 * np.pi > 1e-2:
        raise NotImplementedError("a detector offset angle is not supported (the reference's crop at "
                                  "torch_functions.py:161 is broken)")
    if not (isinstance(frac_4D, torch.Tensor) and frac_4D.is_cuda):
        raise RuntimeError("frac_4D must be a CUDA tensor")
    frac = frac_4D.to(torch.float32).contiguous()
    side = _side(position_det)
    K, B, H, W = frac.shape
    T = xrf_exponent_rot(frac, coef_xrf(param_angle_energy), detector_mask, side)
    out = torch.empty((T.shape[0], B, H, W), dtype=torch.float32, device=frac.device)
    with torch.cuda.device(frac.device):
        _lib.call("fl_xrf_combine", _lib.ptr(T), None, _lib.ptr(out), T.shape[0], B, H, W, 0, B * H * W, side,
                  _lib.stream_of(frac.device))
    return out.cpu().numpy() if return_numpy else out


def torch_3D_forward_projection(C, angle_list, em_cs, atten, mode="batch"):
    """Projections (n_angle, n_sli, W) of C (n_ref, n_sli, H, W) (FL/torch_functions.py:220-244)."""
    dev = C.device
    theta = torch.tensor(np.asarray(angle_list) / 180.0 * np.pi, dtype=torch.float32, device=dev)
    em = _as_device_f32(em_cs, dev)
    att = _as_device_f32(atten, dev)
    C = C.to(torch.float32).contiguous()
    if mode == "batch":
        return forward_emission_batch(att, C, em, theta).cpu().numpy()
    n_angle = len(angle_list)
    n_ref, n_sli, H, W = C.shape
    prj = np.zeros((n_angle, n_sli, W))
    for i in range(n_sli):
        prj[:, i] = forward_emission(att[:, i].contiguous(), C[:, i].contiguous(), em, theta).cpu().numpy()
    return prj
