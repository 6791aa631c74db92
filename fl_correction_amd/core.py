"""Drop-in for the per-angle attenuation drivers and the projection simulator of FL/core.py.

The reference computes each angle's maps, copies them to the host and writes
them as TIFF files (FL/core.py:19-101, FL/file_io.py:69-126), then reads them
back for MLEM.  The drivers here do the same when given `fsave` (the reference's
default './Angle_prj' included), and always also return the maps as one device
tensor of shape (K, A, B, H, W) (element, angle, slice, row, col), i.e. one
(A, B, H, W) attenuation stack per emission line, the layout ``mlem_cuda_batch``
and the projector consume.  ``cal_atten_at_angles`` fuses the three reference
drivers (fluorescence path, incident beam, product) into one pass per angle
and writes no files.
"""
from __future__ import annotations

import glob
import math
import os

import numpy as np
import torch

from . import _lib, file_io
from .atten_util import coef_xray, coef_xrf, extract_single_value_param, update_dict_use_ref_at_angle
from .cuda_lib.atten_cuda import forward_emission_batch
from .torch_functions import _side, xrf_exponent_rot


def angle_entries(angle_deg):
    """float32 rotation-matrix entries for an angle in degrees, as the reference builds them:
    theta = torch.tensor(angle/180*pi) (float32, FL/core.py:35), then math.cos/sin of it
    rounded to float32 (FL/torch_image_util.py:77-85)."""
    th32 = float(np.float32(angle_deg / 180.0 * np.pi))
    return float(np.float32(math.cos(th32))), float(np.float32(math.sin(th32)))


def _prep(frac_4D, angle_list, dict_use_ref, detector_offset_angle=0):
    if len(dict_use_ref) > 0 and dict_use_ref.get("use_ref", False):
        raise NotImplementedError("reference-spectrum (use_ref) attenuation is not supported")
    if abs(float(detector_offset_angle)) / 180.0 * np.pi > 1e-2:
        raise NotImplementedError("a detector offset angle is not supported")
    if not (isinstance(frac_4D, torch.Tensor) and frac_4D.is_cuda):
        raise RuntimeError("frac_4D must be a CUDA tensor")
    frac = frac_4D.to(torch.float32).contiguous()
    return frac, np.asarray(angle_list, np.float64).reshape(-1)


def _alloc(out, K, A, B, H, W, dev):
    if out is None:
        return torch.empty((K, A, B, H, W), dtype=torch.float32, device=dev)
    if tuple(out.shape) != (K, A, B, H, W) or out.dtype != torch.float32 or not out.is_contiguous():
        raise RuntimeError(f"out must be a contiguous float32 tensor of shape {(K, A, B, H, W)}")
    return out


def _rotate(frac, angle_deg, buf):
    K, B, H, W = frac.shape
    c, s = angle_entries(angle_deg)
    _lib.call("fl_rotate", _lib.ptr(frac), _lib.ptr(buf), K * B, H, W, c, s, _lib.stream_of(frac.device))
    return buf


def _write_maps(out, elem_type, fsave, mode):
    """Every angle's (B, H, W) map of each element as the reference names it
    (write_attenuation_fl / _xray / write_attenuation, FL/file_io.py:69-126)."""
    if fsave is None:
        return
    file_io.write_attenuation_stack(out, list(elem_type), fsave, mode)


def cal_xrf_atten_at_angles(frac_4D, param, angle_list, detector_mask, detector_offset_angle=0, position_det="r",
                            dict_use_ref={}, fsave="./Angle_prj", out=None):
    """Fluorescence-path attenuation at every angle (FL/core.py:19-52): rotate frac by the
    angle, exp(-(mask (*) mu_line)) per element.  As the reference, writes
    atten_fl_{elem}_prj_{id:04d}.tiff into `fsave` (default './Angle_prj'); fsave=None
    keeps the maps on the device only.  Returns the device stack (K, A, B, H, W)."""
    frac, angles = _prep(frac_4D, angle_list, dict_use_ref, detector_offset_angle)
    K, B, H, W = frac.shape
    A = len(angles)
    side = _side(position_det)
    dev = frac.device
    out = _alloc(out, K, A, B, H, W, dev)
    rot = torch.empty_like(frac)
    work = torch.empty((K, B, W, H), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        for a, ang in enumerate(angles):
            update_dict_use_ref_at_angle(dict_use_ref, ang)
            p = extract_single_value_param(param, a, dict_use_ref)
            _rotate(frac, ang, rot)
            T = xrf_exponent_rot(rot, coef_xrf(p), detector_mask, side, work)
            _lib.call("fl_xrf_combine", _lib.ptr(T), None, _lib.ptr(out[0, a]), K, B, H, W, 0, out.stride(0), side,
                      _lib.stream_of(dev))
    _write_maps(out, param["elem_type"], fsave, "fl")
    return out


def cal_xray_atten_at_angles(frac_4D, param, angle_list, dict_use_ref={}, fsave="./Angle_prj", out=None):
    """Incident-beam attenuation at every angle (FL/core.py:55-81).  Writes
    atten_incident_xray_{elem}_prj_{id:04d}.tiff into `fsave` unless it is None.
    Returns the device stack (K, A, B, H, W)."""
    frac, angles = _prep(frac_4D, angle_list, dict_use_ref)
    K, B, H, W = frac.shape
    A = len(angles)
    dev = frac.device
    out = _alloc(out, K, A, B, H, W, dev)
    rot = torch.empty_like(frac)
    with torch.cuda.device(dev):
        for a, ang in enumerate(angles):
            p = extract_single_value_param(param, a, dict_use_ref)
            _rotate(frac, ang, rot)
            coef = torch.as_tensor(coef_xray(p), device=dev)
            _lib.call("fl_incident_atten", _lib.ptr(rot), _lib.ptr(coef), _lib.ptr(out[0, a]), K, B, H, W, B * H * W,
                      out.stride(0), _lib.stream_of(dev))
    _write_maps(out, param["elem_type"], fsave, "xray")
    return out


def cal_total_atten_at_angles(fsave, param):
    """Total attenuation files from the two per-path file sets (FL/core.py:84-101): the
    angle count is the number of atten_fl_* files over the element count; for every
    element and angle, atten_{elem}_prj_{id:04d}.tiff = fl map * incident map (a float32
    product: both maps are float32, so the reference's float64 product rounds to it)."""
    elem_type = list(param["elem_type"])
    n_angle = len(glob.glob(os.path.join(fsave, "atten_fl*"))) // len(elem_type)
    for elem in elem_type:
        for a in range(n_angle):
            fl = file_io.read_attenuation_fl(a, fsave, elem)
            xr = file_io.read_attenuation_xray(a, fsave, elem)
            if fl.shape != xr.shape:
                raise RuntimeError(f"{elem} angle {a}: fluorescence map {fl.shape} and incident map {xr.shape} differ")
            file_io.write_attenuation(elem, fl * xr, a, fsave)


def cal_total_atten_stack(atten_xrf, atten_xray):
    """Product of the two device stacks (K, A, B, H, W): cal_total_atten_at_angles
    without the file round trip."""
    if atten_xrf.shape != atten_xray.shape:
        raise RuntimeError("attenuation stacks differ in shape")
    return atten_xrf * atten_xray


def cal_atten_at_angles(frac_4D, param, angle_list, detector_mask, position_det="r", dict_use_ref={}, out=None,
                        angle_ids=None):
    """Total attenuation A[e] = A_xrf[e] * A_x[e] at every angle in one pass per angle: the
    three reference drivers (FL/core.py:19-101) fused, maps kept on the device.
    Returns (K, A, B, H, W); with angle_ids, only those angle indices are computed."""
    frac, angles = _prep(frac_4D, angle_list, dict_use_ref)
    K, B, H, W = frac.shape
    A = len(angles)
    side = _side(position_det)
    dev = frac.device
    out = _alloc(out, K, A, B, H, W, dev)
    rot = torch.empty_like(frac)
    inc = torch.empty_like(frac)
    work = torch.empty((K, B, W, H), dtype=torch.float32, device=dev)
    ids = range(A) if angle_ids is None else angle_ids
    with torch.cuda.device(dev):
        st = _lib.stream_of(dev)
        for a in ids:
            p = extract_single_value_param(param, a, dict_use_ref)
            _rotate(frac, angles[a], rot)
            cx = torch.as_tensor(coef_xray(p), device=dev)
            _lib.call("fl_incident_atten", _lib.ptr(rot), _lib.ptr(cx), _lib.ptr(inc), K, B, H, W, B * H * W,
                      B * H * W, st)
            T = xrf_exponent_rot(rot, coef_xrf(p), detector_mask, side, work)
            _lib.call("fl_xrf_combine", _lib.ptr(T), _lib.ptr(inc), _lib.ptr(out[0, a]), K, B, H, W, B * H * W,
                      out.stride(0), side, st)
    return out


def re_projection_cuda(img3D, angle_list, param=None, rho_compound=1, atten_coef=None, elem="Ni", use_ref=False,
                       device="cuda"):
    """XRF projections of a volume at all angles (FL/core.py:104-156): em = em_cs * pix,
    C = img * rho, atten = 1 when not given.  Returns numpy (n_angle, n_sli, W)."""
    n_angle = len(angle_list)
    if param is None:
        em = np.ones((n_angle, 1))
    else:
        pix = param["pix"]
        em_cs_dict = param["em_cs"]
        if use_ref:
            n_ref = img3D.shape[0]
            em = np.zeros([n_angle, n_ref])
            for i in range(n_ref):
                em[:, i] = em_cs_dict[f"{elem}_ref{i + 1}"] * pix
        else:
            em = np.asarray(em_cs_dict[elem]) * pix
            em = np.broadcast_to(em.reshape(-1, 1), (n_angle, 1)) if em.size == 1 else em.reshape((len(em), 1))
    img3D = img3D * rho_compound
    if len(img3D.shape) == 3:
        img3D = img3D[np.newaxis]
    C = torch.as_tensor(np.asarray(img3D), dtype=torch.float32, device=device) if not isinstance(img3D, torch.Tensor) \
        else img3D.to(device=device, dtype=torch.float32).contiguous()
    n_ref, n_sli, H, W = C.shape
    if atten_coef is None:
        atten = torch.ones(n_angle, n_sli, H, W, dtype=torch.float32, device=device)
    else:
        if len(atten_coef.shape) == 3:
            atten_coef = atten_coef[:, np.newaxis]
        atten = torch.as_tensor(atten_coef, dtype=torch.float32, device=device).contiguous()
    theta = torch.tensor(np.asarray(angle_list) / 180.0 * np.pi, dtype=torch.float32, device=device)
    em_t = torch.as_tensor(np.ascontiguousarray(em), dtype=torch.float32, device=device)
    return forward_emission_batch(atten, C, em_t, theta).cpu().numpy()
