"""Drop-in for the reference op set ``FL.cuda_lib.atten_cuda``.

Same function names, positional order, keyword names and defaults as the
pybind11 module of FL/cuda_lib/atten_cuda.cpp:351-392; same error behaviour
(``RuntimeError`` for non-device tensors, wrong ranks, mismatched shapes).
Every op allocates its outputs (the callee allocates, as at atten_cuda.cpp:22,
104, 128, 172, 204, 254, 303), runs on the current stream of the inputs'
device, and never modifies its inputs.

Differences from the reference, all deliberate:
  * float32 is required of every tensor argument and checked up front (the
    reference fails later inside ``data_ptr<float>()``);
  * every input is made contiguous (the reference's backward ops read
    non-contiguous inputs as if they were contiguous);
  * launches go to torch's current stream under a device guard, not to the
    legacy default stream;
  * results are deterministic (no float atomics) and indices are 64-bit.
"""
from __future__ import annotations

import torch

from .. import _lib


def _check(cond, msg):
    if not cond:
        raise RuntimeError(msg)


def _dev(name, t):
    _check(isinstance(t, torch.Tensor), f"{name} must be a tensor")
    _check(t.is_cuda, f"{name} must be CUDA")
    _check(t.dtype == torch.float32, f"{name} must be float32")
    return t.contiguous()


def _same_device(*ts):
    d = ts[0].device
    for t in ts[1:]:
        _check(t.device == d, "all tensors must be on the same device")
    return d


def _workspace(nbytes, device):  # noqa: D401  (scratch for the MLEM entry point)
    return torch.empty((max(int(nbytes), 1) + 3) // 4, dtype=torch.float32, device=device)


# ---------------------------------------------------------------------------
# Detector-mask correlation: atten_cuda.cpp:12-28
# ---------------------------------------------------------------------------


def cal_atten_cuda(mu, mask):
    """out[k,z,x,y] = sum_{dz,dx,dy} mask[dz,dx,dy] * mu[k, z+dz-Mz/2, x+dx, y+dy-My/2] with
    out-of-range taps skipped (atten_cuda_kernel.cu:10-57).  mu (K,Z,X,Y), mask (Mz,Mx,My)."""
    _check(isinstance(mu, torch.Tensor) and mu.is_cuda, "mu must be a CUDA tensor")
    _check(isinstance(mask, torch.Tensor) and mask.is_cuda, "mask must be a CUDA tensor")
    _check(mu.dtype == torch.float32, "mu must be float32")
    _check(mask.dtype == torch.float32, "mask must be float32")
    _check(mu.dim() == 4, "mu must be (K, Z, X, Y)")
    _check(mask.dim() == 3, "mask must be (Mz, Mx, My)")
    mu, mask = mu.contiguous(), mask.contiguous()
    dev = _same_device(mu, mask)
    out = torch.zeros_like(mu)
    K, Z, X, Y = mu.shape
    Mz, Mx, My = mask.shape
    if mu.numel() == 0:
        return out
    with torch.cuda.device(dev):
        _lib.call("fl_cal_atten", _lib.ptr(mu), _lib.ptr(mask), _lib.ptr(out), K, Z, X, Y, Mz, Mx, My,
                  _lib.stream_of(dev))
    return out


# ---------------------------------------------------------------------------
# Projector: atten_cuda.cpp:82-214
# ---------------------------------------------------------------------------


def forward_emission(atten, C, em_cs, theta):
    """Pf (n_angle, W) from atten (n_angle, H, W), C (n_ref, H, W), em_cs (n_angle, n_ref),
    theta (n_angle) radians.  atten_cuda.cpp:82-108."""
    atten, C, em_cs, theta = (_dev(n, t) for n, t in
                              (("atten", atten), ("C", C), ("em_cs", em_cs), ("theta", theta)))
    _check(atten.dim() == 3, "atten must be (n_angle, H, W)")
    _check(C.dim() == 3, "C must be (n_ref, H, W)")
    _check(em_cs.dim() == 2, "em_cs must be (n_angle, n_ref)")
    _check(theta.dim() == 1, "theta must be (n_angle)")
    Pf = forward_emission_batch(atten.unsqueeze(1), C.unsqueeze(1), em_cs, theta)
    return Pf.squeeze(1)


def backward_emission(atten, ratio, em_cs, theta):
    """out (n_ref, H, W) = transpose of forward_emission applied to ratio (n_angle, W).
    n_ref = em_cs.size(1).  atten_cuda.cpp:113-135."""
    for n, t in (("atten", atten), ("ratio", ratio), ("em_cs", em_cs), ("theta", theta)):
        _check(isinstance(t, torch.Tensor) and t.is_cuda, f"{n} must be CUDA")
    _check(atten.dim() == 3, "atten must be (n_angle, H, W)")
    _check(ratio.dim() == 2, "ratio must be (n_angle, W)")
    out = backward_emission_batch(atten.unsqueeze(1), ratio.unsqueeze(1), em_cs, theta)
    return out.squeeze(1)


def forward_emission_batch(atten, C, em_cs, theta):
    """Pf (n_angle, B, W) from atten (n_angle, B, H, W), C (n_ref, B, H, W), em_cs (n_angle, n_ref),
    theta (n_angle).  atten_cuda.cpp:142-182."""
    atten, C, em_cs, theta = (_dev(n, t) for n, t in
                              (("atten", atten), ("C", C), ("em_cs", em_cs), ("theta", theta)))
    _check(atten.dim() == 4, "atten must be (n_angle, B, H, W)")
    _check(C.dim() == 4, "C must be (n_ref, B, H, W)")
    _check(em_cs.dim() == 2, "em_cs must be (n_angle, n_ref)")
    _check(theta.dim() == 1, "theta must be (n_angle)")
    A, B, H, W = atten.shape
    R = C.size(0)
    _check(C.size(1) == B, "batch size mismatch")
    _check(C.size(2) == H, "H mismatch")
    _check(C.size(3) == W, "W mismatch")
    _check(em_cs.size(0) == A, "n_angle mismatch")
    _check(em_cs.size(1) == R, "n_ref mismatch")
    _check(theta.size(0) == A, "theta length mismatch")
    dev = _same_device(atten, C, em_cs, theta)
    Pf = torch.empty((A, B, W), dtype=torch.float32, device=dev)
    if A == 0 or B == 0 or W == 0:
        return Pf.zero_()
    with torch.cuda.device(dev):
        _lib.call("fl_forward_emission_batch", _lib.ptr(atten), _lib.ptr(C), _lib.ptr(em_cs), _lib.ptr(theta),
                  _lib.ptr(Pf), A, B, R, H, W, _lib.stream_of(dev))
    return Pf


def backward_emission_batch(atten, ratio, em_cs, theta):
    """out (n_ref, B, H, W), n_ref = em_cs.size(1), from atten (n_angle, B, H, W) and
    ratio (n_angle, B, W).  atten_cuda.cpp:188-214."""
    for n, t in (("atten", atten), ("ratio", ratio), ("em_cs", em_cs), ("theta", theta)):
        _check(isinstance(t, torch.Tensor) and t.is_cuda, f"{n} must be CUDA")
    atten, ratio, em_cs, theta = (_dev(n, t) for n, t in
                                  (("atten", atten), ("ratio", ratio), ("em_cs", em_cs), ("theta", theta)))
    _check(atten.dim() == 4, "atten must be (n_angle, B, H, W)")
    _check(em_cs.dim() == 2, "em_cs must be (n_angle, n_ref)")
    A, B, H, W = atten.shape
    R = em_cs.size(1)
    _check(ratio.shape == (A, B, W), "ratio must be (n_angle, B, W)")
    _check(em_cs.size(0) == A and theta.numel() == A, "n_angle mismatch")
    dev = _same_device(atten, ratio, em_cs, theta)
    out = torch.empty((R, B, H, W), dtype=torch.float32, device=dev)
    if B == 0 or R == 0:
        return out
    with torch.cuda.device(dev):
        _lib.call("fl_backward_emission_batch", _lib.ptr(atten), _lib.ptr(ratio), _lib.ptr(em_cs), _lib.ptr(theta),
                  _lib.ptr(out), A, B, R, H, W, _lib.stream_of(dev))
    return out


def huber_grad_batch(C, delta=0.01):
    """4-neighbour Huber gradient (atten_cuda_kernel.cu:564-616).  Not exported by the
    reference's module (FL/cuda_lib/__init__.py:10-11); provided for tests and callers."""
    C = _dev("C", C)
    squeeze = C.dim() == 3
    if squeeze:
        C = C.unsqueeze(1)
    _check(C.dim() == 4, "C must be (n_ref, B, H, W)")
    R, B, H, W = C.shape
    g = torch.empty_like(C)
    with torch.cuda.device(C.device):
        _lib.call("fl_huber_grad_batch", _lib.ptr(C), _lib.ptr(g), R, B, H, W, float(delta),
                  _lib.stream_of(C.device))
    return g.squeeze(1) if squeeze else g


# ---------------------------------------------------------------------------
# MLEM: atten_cuda.cpp:244-345
# ---------------------------------------------------------------------------


def mlem_cuda_batch(C_init, atten, em_cs, theta, I, n_iter, beta=1e-3, delta=0.01):
    """Batch MLEM reconstruction with optional Huber regularization (atten_cuda.cpp:283-345).
    C_init (n_ref, B, H, W), atten (n_angle, B, H, W), em_cs (n_angle, n_ref), theta (n_angle),
    I (n_angle, B, W).  Returns a new tensor; C_init is not modified."""
    for n, t in (("C_init", C_init), ("atten", atten), ("em_cs", em_cs), ("theta", theta), ("I", I)):
        _check(isinstance(t, torch.Tensor) and t.is_cuda, f"{n} must be CUDA")
    _check(C_init.dim() == 4, "C_init must be (n_ref, B, H, W)")
    _check(atten.dim() == 4, "atten must be (n_angle, B, H, W)")
    _check(I.dim() == 3, "I must be (n_angle, B, W)")
    C_init, atten, em_cs, theta, I = (_dev(n, t) for n, t in (("C_init", C_init), ("atten", atten),
                                                               ("em_cs", em_cs), ("theta", theta), ("I", I)))
    R, B, H, W = C_init.shape
    A = atten.size(0)
    _check(atten.shape == (A, B, H, W), "atten shape does not match C_init")
    _check(em_cs.shape == (A, R), "em_cs must be (n_angle, n_ref)")
    _check(theta.numel() == A, "theta length mismatch")
    _check(I.shape == (A, B, W), "I must be (n_angle, B, W)")
    n_iter = int(n_iter)
    _check(n_iter >= 0, "n_iter must be >= 0")
    dev = _same_device(C_init, atten, em_cs, theta, I)
    out = torch.empty_like(C_init)
    with torch.cuda.device(dev):
        # the slice-major back-projection when its copy of atten fits in free
        # memory (fl_mlem_batch takes it when handed a workspace that large)
        from ..mlem import slice_major_fits
        nbytes = _lib.load().fl_mlem_workspace_bytes(A, B, R, H, W)
        if slice_major_fits(A, B, R, H, W, dev, extra_bytes=nbytes):
            nbytes = _lib.load().fl_mlem_workspace_bytes_sm(A, B, R, H, W)
        ws = _workspace(nbytes, dev)
        _lib.call("fl_mlem_batch", _lib.ptr(C_init), _lib.ptr(atten), _lib.ptr(em_cs), _lib.ptr(theta), _lib.ptr(I),
                  _lib.ptr(out), A, B, R, H, W, n_iter, float(beta), float(delta), _lib.ptr(ws), ws.numel() * 4,
                  _lib.stream_of(dev))
    return out


def mlem_cuda(C_init, atten, em_cs, theta, I, n_iter, beta=1e-3, delta=0.01):
    """MLEM reconstruction with optional Huber regularization (atten_cuda.cpp:244-281).
    C_init (n_ref, H, W), atten (n_angle, H, W), I (n_angle, W)."""
    for n, t in (("C_init", C_init), ("atten", atten), ("em_cs", em_cs), ("theta", theta), ("I", I)):
        _check(isinstance(t, torch.Tensor) and t.is_cuda, f"{n} must be CUDA")
    _check(C_init.dim() == 3, "C_init must be (n_ref, H, W)")
    _check(atten.dim() == 3, "atten must be (n_angle, H, W)")
    _check(I.dim() == 2, "I must be (n_angle, W)")
    out = mlem_cuda_batch(C_init.unsqueeze(1), atten.unsqueeze(1), em_cs, theta, I.unsqueeze(1), n_iter, beta, delta)
    return out.squeeze(1)
