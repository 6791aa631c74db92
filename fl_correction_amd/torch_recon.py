"""Drop-in for FL/torch_recon.py: MLEM reconstruction from numpy inputs.

Same signatures, defaults and conversions as the reference: angles in degrees
(theta = angle/180*pi, cast to float32), atten None -> ones, em_cs None -> ones,
1-D em_cs -> (n_angle, 1), em_cs scaled by rho*pix, 3-D inputs promoted to the
batch form, result returned as numpy.  Note the reference's defaults run
One-Step-Late MAP-EM with a Huber prior (beta=1e-3); pass beta=0 for plain MLEM.
"""
from __future__ import annotations

import numpy as np
import torch

from .cuda_lib.atten_cuda import mlem_cuda, mlem_cuda_batch


def _t(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float32).contiguous()
    return torch.tensor(np.asarray(x), dtype=torch.float, device=device)


def torch_mlem_recon(C_init, prj_sli, angle_list, atten=None, em_cs=None, rho=1, pix=1, n_iter=50, beta=1e-3,
                     delta=0.01, device="cuda"):
    """One slice (FL/torch_recon.py:5-47): C_init (n_ref, H, W), prj_sli (n_angle, W)."""
    n_ref, H, W = C_init.shape
    n_angle = len(angle_list)
    theta = np.asarray(angle_list) / 180.0 * np.pi
    theta_cuda = torch.tensor(theta, dtype=torch.float, device=device)
    atten_cuda = torch.ones(n_angle, H, W, dtype=torch.float, device=device) if atten is None else _t(atten, device)
    if em_cs is None:
        em_cs = np.ones((n_angle, n_ref))
    elif len(np.shape(em_cs)) == 1:
        em_cs = np.asarray(em_cs)[:, np.newaxis]
    em_cs = np.asarray(em_cs) * rho * pix
    rec = mlem_cuda(_t(C_init, device), atten_cuda, _t(em_cs, device), theta_cuda, _t(prj_sli, device), n_iter,
                    beta, delta)
    return rec.cpu().numpy()


def torch_mlem_recon_batch(C_init, prjs, angle_list, atten=None, em_cs=None, rho=1, pix=1, n_iter=50, beta=1e-3,
                           delta=0.01, device="cuda"):
    """Slice batch (FL/torch_recon.py:50-98): C_init (n_ref, n_sli, H, W) or (n_ref, H, W),
    prjs (n_angle, n_sli, W) or (n_angle, W), atten (n_angle, n_sli, H, W) or (n_angle, H, W)."""
    if len(C_init.shape) == 3:
        C_init = C_init[:, np.newaxis]
    if len(prjs.shape) == 2:
        prjs = prjs[:, np.newaxis]
    n_ref, n_sli, H, W = C_init.shape
    n_angle = len(angle_list)
    theta = np.asarray(angle_list) / 180.0 * np.pi
    theta_cuda = torch.tensor(theta, dtype=torch.float, device=device)
    if atten is None:
        atten_cuda = torch.ones(n_angle, n_sli, H, W, dtype=torch.float, device=device)
    else:
        if len(atten.shape) == 3:
            atten = atten[:, np.newaxis]
        atten_cuda = _t(atten, device)
    if em_cs is None:
        em_cs = np.ones((n_angle, n_ref))
    elif len(np.shape(em_cs)) == 1:
        em_cs = np.asarray(em_cs)[:, np.newaxis]
    em_cs = np.asarray(em_cs) * rho * pix
    rec = mlem_cuda_batch(_t(C_init, device), atten_cuda, _t(em_cs, device), theta_cuda, _t(prjs, device), n_iter,
                          beta, delta)
    return rec.cpu().numpy()
