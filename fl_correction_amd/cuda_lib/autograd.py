"""Drop-in for FL/cuda_lib/autograd.py: autograd wrappers of the projector.

Same classes and entry points as the reference (autograd.py:5-77).  The
reference returns a gradient for C only; these also return d/d(atten) when
atten requires grad (the projector is linear in atten: G[a,b,w] times the
sampled, emission-weighted C).  For the gradient through the attenuation's
dependence on the concentrations use fl_correction_amd.model.XRFModelFunction.
"""
import torch

from .. import _lib
from . import atten_cuda


def _grad_atten(atten, C, em_cs, theta, G):
    A, B, H, W = atten.shape
    R = C.shape[0]
    out = torch.empty((A, B, H, W), dtype=torch.float32, device=atten.device)
    Cc, em, th, Gc = (t.contiguous().to(torch.float32) for t in (C, em_cs, theta, G))
    with torch.cuda.device(atten.device):
        _lib.call("fl_sample_grad", _lib.ptr(Cc), _lib.ptr(em), _lib.ptr(th), _lib.ptr(Gc), _lib.ptr(out), A, B, R, H,
                  W, B * H * W, _lib.stream_of(atten.device))
    return out


class ForwardEmissionFunction(torch.autograd.Function):
    """Pf (n_angle, W) = forward_emission(atten (n_angle, H, W), C (n_ref, H, W), em_cs, theta)."""

    @staticmethod
    def forward(ctx, atten, C, em_cs, theta):
        ctx.save_for_backward(atten, C, em_cs, theta)
        return atten_cuda.forward_emission(atten, C, em_cs, theta)

    @staticmethod
    def backward(ctx, grad_output):
        atten, C, em_cs, theta = ctx.saved_tensors
        g = grad_output.contiguous()
        grad_C = atten_cuda.backward_emission(atten, g, em_cs, theta) if ctx.needs_input_grad[1] else None
        grad_atten = None
        if ctx.needs_input_grad[0]:
            grad_atten = _grad_atten(atten.unsqueeze(1), C.unsqueeze(1), em_cs, theta, g.unsqueeze(1)).squeeze(1)
        return grad_atten, grad_C, None, None


def forward_emission_autograd(atten, C, em_cs, theta):
    return ForwardEmissionFunction.apply(atten, C, em_cs, theta)


class ForwardEmissionBatchFunction(torch.autograd.Function):
    """Pf (n_angle, B, W) = forward_emission_batch(atten (n_angle, B, H, W), C (n_ref, B, H, W), em_cs, theta)."""

    @staticmethod
    def forward(ctx, atten, C, em_cs, theta):
        ctx.save_for_backward(atten, C, em_cs, theta)
        return atten_cuda.forward_emission_batch(atten, C, em_cs, theta)

    @staticmethod
    def backward(ctx, grad_output):
        atten, C, em_cs, theta = ctx.saved_tensors
        g = grad_output.contiguous()
        grad_C = atten_cuda.backward_emission_batch(atten, g, em_cs, theta) if ctx.needs_input_grad[1] else None
        grad_atten = _grad_atten(atten, C, em_cs, theta, g) if ctx.needs_input_grad[0] else None
        return grad_atten, grad_C, None, None


def forward_emission_batch_autograd(atten, C, em_cs, theta):
    return ForwardEmissionBatchFunction.apply(atten, C, em_cs, theta)
