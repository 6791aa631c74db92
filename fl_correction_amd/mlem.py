"""Persistent MLEM driver over the C ABI.

``mlem_cuda_batch`` (the drop-in op) allocates its workspace per call, as the
reference allocates its temporaries per call (atten_cuda.cpp:303-316).  For
callers that iterate many times, step by step, or interleave several emission
lines, ``MLEMSolver`` keeps the sensitivity, the ratio buffer and the ping-pong
estimate resident, and exposes the two stages of
an iteration separately so that their kernels can be timed.
"""
from __future__ import annotations

import torch

from . import _lib


class MLEMSolver:
    """One emission line: atten (A,B,H,W), em_cs (A,R), theta (A), I (A,B,W)."""

    def __init__(self, atten, em_cs, theta, I, R=1, beta=0.0, delta=0.01):
        for name, t in (("atten", atten), ("em_cs", em_cs), ("theta", theta), ("I", I)):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                raise RuntimeError(f"{name} must be a contiguous float32 device tensor")
        self.atten, self.em_cs, self.theta, self.I = atten, em_cs, theta, I
        self.A, self.B, self.H, self.W = (int(x) for x in atten.shape)
        self.R = int(em_cs.shape[1])
        self.beta, self.delta = float(beta), float(delta)
        dev = atten.device
        self.device = dev
        self.sens = torch.empty((self.R, self.B, self.H, self.W), dtype=torch.float32, device=dev)
        self.ratio = torch.empty((self.A, self.B, self.W), dtype=torch.float32, device=dev)
        self._dims = (self.A, self.B, self.R, self.H, self.W)
        with torch.cuda.device(dev):
            _lib.call("fl_mlem_sensitivity", _lib.ptr(atten), _lib.ptr(em_cs), _lib.ptr(theta), _lib.ptr(self.sens),
                      *self._dims, _lib.stream_of(dev))

    def ratio_step(self, C):
        """ratio = I / max(F C, 1e-6)."""
        _lib.call("fl_mlem_ratio", _lib.ptr(self.atten), _lib.ptr(C), _lib.ptr(self.em_cs), _lib.ptr(self.theta),
                  _lib.ptr(self.I), _lib.ptr(self.ratio), *self._dims, _lib.stream_of(self.device))

    def update_step(self, C, C_next):
        """C_next = OSL/MLEM update of C from the current ratio."""
        _lib.call("fl_mlem_update", _lib.ptr(self.atten), _lib.ptr(self.ratio), _lib.ptr(self.em_cs),
                  _lib.ptr(self.theta), _lib.ptr(C), _lib.ptr(self.sens), _lib.ptr(C_next), *self._dims,
                  self.beta, self.delta, _lib.stream_of(self.device))

    def run(self, C_init, n_iter):
        """n_iter iterations from C_init; returns a new tensor."""
        cur = C_init.clone()
        nxt = torch.empty_like(cur)
        with torch.cuda.device(self.device):
            for _ in range(int(n_iter)):
                self.ratio_step(cur)
                self.update_step(cur, nxt)
                cur, nxt = nxt, cur
        return cur
