"""Persistent MLEM driver over the C ABI.

``mlem_cuda_batch`` (the drop-in op) allocates its workspace per call, as the
reference allocates its temporaries per call (atten_cuda.cpp:303-316).  For
callers that iterate many times, step by step, or interleave several emission
lines, ``MLEMSolver`` keeps the sensitivity, the ratio buffer and the ping-pong
estimate resident, and exposes the two stages of
an iteration separately so that their kernels can be timed.

Layouts of the back-projection: "reference" runs it on atten (A, B, H, W) as
given; "slice_major" first copies atten to (A, H, W, B) once and back-projects
from the copy (one sample geometry per detector point shared by all slices,
contiguous loads over slices), at the cost of a second attenuation array in
HBM.  "auto" takes the slice-major layout for one reference per line, at least
``SLICE_MAJOR_MIN_SLICES`` slices and when the copy fits in free device memory
with ``SLICE_MAJOR_HEADROOM`` bytes to spare.
"""
from __future__ import annotations

import torch

from . import _lib

SLICE_MAJOR_MIN_SLICES = 32
SLICE_MAJOR_HEADROOM = 2 << 30


def slice_major_fits(A, B, R, H, W, device, extra_bytes=0):
    """Whether the slice-major back-projection applies to (A, B, R, H, W) and its
    copy of the attenuation plus scratch (and `extra_bytes`) fit in free memory."""
    if R != 1 or B < SLICE_MAJOR_MIN_SLICES:
        return False
    if int(_lib.load().fl_mlem_workspace_bytes_sm(A, B, R, H, W)) == 0:
        return False
    need = 4 * A * B * H * W + int(_lib.load().fl_backproject_sm_scratch_bytes(A, B, H, W)) + int(extra_bytes)
    free, _ = torch.cuda.mem_get_info(device)
    return need + SLICE_MAJOR_HEADROOM <= free


class MLEMSolver:
    """One emission line: atten (A,B,H,W), em_cs (A,R), theta (A), I (A,B,W)."""

    def __init__(self, atten, em_cs, theta, I, R=1, beta=0.0, delta=0.01, layout="auto"):
        for name, t in (("atten", atten), ("em_cs", em_cs), ("theta", theta), ("I", I)):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                raise RuntimeError(f"{name} must be a contiguous float32 device tensor")
        self.atten, self.em_cs, self.theta, self.I = atten, em_cs, theta, I
        self.A, self.B, self.H, self.W = (int(x) for x in atten.shape)
        self.R = int(em_cs.shape[1])
        self.beta, self.delta = float(beta), float(delta)
        dev = atten.device
        self.device = dev
        A, B, R, H, W = self.A, self.B, self.R, self.H, self.W
        if layout not in ("auto", "slice_major", "reference"):
            raise ValueError(f"layout must be 'auto', 'slice_major' or 'reference', not {layout!r}")
        with torch.cuda.device(dev):
            if layout == "auto":
                layout = "slice_major" if slice_major_fits(A, B, R, H, W, dev) else "reference"
            if layout == "slice_major" and (R != 1 or int(_lib.load().fl_mlem_workspace_bytes_sm(A, B, R, H, W)) == 0):
                raise RuntimeError("the slice-major layout needs one reference per line and H*W*B below 2^29")
            self.layout = layout
            self.sens = torch.empty((R, B, H, W), dtype=torch.float32, device=dev)
            self._dims = (A, B, R, H, W)
            st = _lib.stream_of(dev)
            if layout == "slice_major":
                self.atten_sm = torch.empty((A, H, W, B), dtype=torch.float32, device=dev)
                self.ratio = torch.empty((A, W, B), dtype=torch.float32, device=dev)
                nbytes = int(_lib.load().fl_backproject_sm_scratch_bytes(A, B, H, W))
                self.scratch = torch.empty((nbytes + 3) // 4, dtype=torch.float32, device=dev)
                self._sm = (A, B, H, W)
                _lib.call("fl_slice_major", _lib.ptr(atten), _lib.ptr(self.atten_sm), A, B, H * W, st)
                _lib.call("fl_mlem_sensitivity_sm", _lib.ptr(self.atten_sm), _lib.ptr(em_cs), _lib.ptr(theta),
                          _lib.ptr(self.sens), *self._sm, _lib.ptr(self.scratch), self.scratch.numel() * 4, st)
            else:
                self.ratio = torch.empty((A, B, W), dtype=torch.float32, device=dev)
                _lib.call("fl_mlem_sensitivity", _lib.ptr(atten), _lib.ptr(em_cs), _lib.ptr(theta),
                          _lib.ptr(self.sens), *self._dims, st)

    def ratio_step(self, C):
        """ratio = I / max(F C, 1e-6); stored (A, W, B) in the slice-major layout."""
        if self.layout == "slice_major":
            _lib.call("fl_mlem_ratio_sm", _lib.ptr(self.atten), _lib.ptr(C), _lib.ptr(self.em_cs),
                      _lib.ptr(self.theta), _lib.ptr(self.I), _lib.ptr(self.ratio), *self._sm,
                      _lib.stream_of(self.device))
            return
        _lib.call("fl_mlem_ratio", _lib.ptr(self.atten), _lib.ptr(C), _lib.ptr(self.em_cs), _lib.ptr(self.theta),
                  _lib.ptr(self.I), _lib.ptr(self.ratio), *self._dims, _lib.stream_of(self.device))

    def update_step(self, C, C_next):
        """C_next = OSL/MLEM update of C from the current ratio."""
        if self.layout == "slice_major":
            _lib.call("fl_mlem_update_sm", _lib.ptr(self.atten_sm), _lib.ptr(self.ratio), _lib.ptr(self.em_cs),
                      _lib.ptr(self.theta), _lib.ptr(C), _lib.ptr(self.sens), _lib.ptr(C_next), *self._sm,
                      self.beta, self.delta, _lib.ptr(self.scratch), self.scratch.numel() * 4,
                      _lib.stream_of(self.device))
            return
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
