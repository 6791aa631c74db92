"""Reconstruction through the attenuated model: the self-consistent MLEM loop
and gradient-based joint reconstruction.

Self-consistent MLEM (``self_consistent_mlem``): the attenuation maps are
recomputed from the current estimate every k iterations and each emission line
runs k MLEM iterations on them (the outer loop of the reference's demo,
demo.py:244-336, with the estimate itself as the concentrations).

Gradient-based joint reconstruction (``joint_reconstruct``):

All element maps C (K, B, H, W) are fitted to the measured projections
I (K, A, B, W) through ``XRFModel``: the attenuation is recomputed from C at
every evaluation, so the gradient includes its dependence on C (the term the
reference's autograd wrapper, FL/cuda_lib/autograd.py:40-74, leaves out).
Loss: the reference's Poisson negative log-likelihood ``poisson_loss_image``
(FL/ml_lib/ml_loss.py:6-8), mean(P - I*log P) with P = max(Pf, eps).
Optimiser: Adam with a projection onto C >= 0 after every step.  The start is
an MLEM estimate (``mlem_start``): per emission line, ``MLEMSolver`` on the
attenuation of the initial estimate.

Every model evaluation runs the library's kernels (``XRFModel.loss_and_grad``:
forward, the loss's cotangent and the adjoint per chunk of angles); the loss and
the Adam update are elementwise torch operations on the device.
"""
from __future__ import annotations

import torch

from .mlem import MLEMSolver
from .model import XRFModel, XRFModelFunction


def poisson_loss(Pf, I, eps=1e-6):
    """mean(P - I*log P), P = max(Pf, eps)   (FL/ml_lib/ml_loss.py:6-8)."""
    P = Pf.clamp_min(eps)
    return (P - I * torch.log(P)).mean()


def data_residual(Pf, I):
    """Relative L2 misfit ||Pf - I|| / ||I|| of the projections."""
    return float(torch.linalg.vector_norm(Pf - I) / torch.linalg.vector_norm(I))


def mlem_start(model: XRFModel, I, C_init, n_iter=10, beta=0.0, delta=0.01, per_line=None):
    """n_iter MLEM iterations per emission line, on the attenuation of C_init.

    I (K, A, B, W), C_init (K, B, H, W); returns a new (K, B, H, W) tensor.
    per_line: compute and keep one line's attenuation (A, B, H, W) at a time
    (each recompute runs the cone sweep for all lines) instead of all K lines'
    at once; None chooses it when the K lines' maps would take more than a third
    of the free device memory (5 lines at 720 x 128 x 512^2 take 483 GB)."""
    C_init = C_init.to(model.device, torch.float32).contiguous()
    I = I.to(model.device, torch.float32).contiguous()
    K, B, H, W = C_init.shape
    if per_line is None:
        free, _ = torch.cuda.mem_get_info(model.device)
        per_line = 4.0 * K * model.A * B * H * W > free / 3
    out = torch.empty_like(C_init)
    atten = None if per_line else model.attenuation(C_init)  # (K, A, B, H, W)
    for e in range(model.K):
        a_e = model.attenuation(C_init, lines=[e])[0] if per_line else atten[e]
        em = model.em_T[e][:, None].contiguous()  # (A, 1): one reference per line
        solver = MLEMSolver(a_e, em, model.theta, I[e].contiguous(), beta=beta, delta=delta)
        out[e] = solver.run(C_init[e:e + 1], n_iter)[0]
        del solver, a_e
    del atten
    return out


def element_fractions(C, scale_range=(0.02, 0.98), enable_scale=True, scale_limit=0.95):
    """cal_frac's 'frac' of the element maps C (K, B, H, W) as float32 (K, B, H, W):
    the maps the reference demo recomputes the attenuation from
    (FL/example/demo.py:255-257, FL/core.py:159-211)."""
    from .frac import cal_frac
    res = cal_frac(C.contiguous(), scale_range=list(scale_range), enable_scale=enable_scale, scale_limit=scale_limit)
    return res["frac"].to(torch.float32).contiguous()


def self_consistent_mlem(model: XRFModel, I, C_init, n_iter=50, recompute_every=5, beta=0.0, delta=0.01,
                         callback=None, start_unattenuated=False, atten_from=None):
    """n_iter MLEM iterations per emission line, the attenuation recomputed from
    the current estimate before every block of `recompute_every` iterations.

    I (K, A, B, W), C_init (K, B, H, W); returns a new (K, B, H, W) tensor.
    callback(iteration, C) is called after every block if given.

    atten_from: a function of the estimate C (K, B, H, W) giving the maps the
    attenuation is recomputed from (None: C itself).  ``element_fractions``
    recomputes it from cal_frac's masked, rescaled element fractions, as the
    reference demo's outer loop does (FL/example/demo.py:255-257).

    start_unattenuated: the first block runs without attenuation (atten = 1),
    as the reference's demo makes its first reconstruction (FL/example/demo.py:
    186-228, atten=None) before it recomputes attenuation from an estimate.  A
    uniform start over the whole volume attenuates far more than a sparse
    sample does, and MLEM then raises the shadowed voxels until the next
    recompute overflows; the unattenuated first block avoids that."""
    if recompute_every < 1:
        raise ValueError("recompute_every must be >= 1")
    C = C_init.to(model.device, torch.float32).clone().contiguous()
    I = I.to(model.device, torch.float32).contiguous()
    done = 0
    while done < n_iter:
        k = min(int(recompute_every), int(n_iter) - done)
        if start_unattenuated and done == 0:
            # atten=None: ones, one (A, B, H, W) array shared by every line
            ones = torch.ones((model.A, *C.shape[1:]), dtype=torch.float32, device=model.device)
            atten = [ones] * model.K
        else:
            src = C if atten_from is None else atten_from(C)
            atten = model.attenuation(src)  # (K, A, B, H, W) for the current estimate
        for e in range(model.K):
            em = model.em_T[e][:, None].contiguous()
            solver = MLEMSolver(atten[e], em, model.theta, I[e].contiguous(), beta=beta, delta=delta)
            C[e] = solver.run(C[e:e + 1], k)[0]
            del solver
        del atten
        done += k
        if callback is not None:
            callback(done, C)
    return C


def joint_reconstruct(model: XRFModel, I, C0, n_steps=100, lr=None, eps=1e-6, callback=None, streams=2):
    """Adam on all element maps with a non-negativity projection.

    I (K, A, B, W) measured projections, C0 (K, B, H, W) start.  lr defaults to
    2% of the start's mean value.  Returns (C, history) with history a list of
    (step, loss, data residual) before each update and after the last one.
    callback(step, C, loss) is called after every step if given.  `streams` is
    the model's stream count for every forward and vjp.

    The loss reported is the reference's Poisson loss, a mean over the
    measurements (FL/ml_lib/ml_loss.py:6-8).  Adam is given the gradient of
    their sum (the mean's times I.numel()): Adam does not depend on the
    gradient's scale except through its eps (1e-8), and the mean's gradient
    shrinks with the number of measurements -- about 7e-11 per voxel at
    128 x 512^2 x 720 angles x 5 lines, where eps would all but stop the
    steps."""
    I = I.to(model.device, torch.float32).contiguous()
    C = C0.to(model.device, torch.float32).clone().contiguous().requires_grad_(True)
    if lr is None:
        lr = 0.02 * float(C0.mean().clamp_min(1e-6))
    opt = torch.optim.Adam([C], lr=lr)
    n_meas = float(I.numel())
    history = []
    for step in range(int(n_steps) + 1):
        # loss, gradient and projections in one pass over the angles (the cone
        # twice per angle instead of three times for forward then vjp)
        loss, grad, Pf = model.loss_and_grad(C, I, eps, streams=streams)
        history.append((step, float(loss), data_residual(Pf, I)))
        if step == n_steps:
            break
        C.grad = grad * n_meas
        opt.step()
        with torch.no_grad():
            C.clamp_(min=0.0)
        if callback is not None:
            callback(step, C, float(loss.detach()))
    model.release()
    return C.detach(), history
