"""Batched MLEM over many independent problems of uneven size.

A problem is one emission line of one sample (the reference's
``torch_mlem_recon_batch`` with one reference, FL/torch_recon.py:50-98):
atten (A,B,H,W), em_cs (A,), theta (A,) in radians, measured I (A,B,W) and
the estimate C (B,H,W), which is updated in place.  The reference
reconstructs such problems one call at a time; ``BatchMLEM`` hands all of
them to the library's ``fl_mlem_multi_*`` entry points, which run every pass
of an iteration as one persistent launch over all problems and replay a
captured hipGraph of two iterations.

Tile shapes are chosen per problem by ``Autotuner``: for each size bucket
(H and W rounded up to multiples of 32) and layout it times full iterations
of a synthetic batch of that bucket -- with the reference layout under every
combination of forward angles per workgroup (8, 4) and back-projector tile
height (31, 15); with the slice-major layout, whose back-projector shape
follows from each pack's slice count, under each forward angle count on
three lines of one sample -- keeps the fastest, and caches the choice in a
JSON file keyed by device name, layout and bucket.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass

import torch

from . import _lib

FWD_ANGLES = (8, 4)
BWD_ROWS = (31, 15)
SLICE_MAJOR_HEADROOM = 2 << 30  # bytes of device memory left free by the 'auto' layout


@dataclass
class Problem:
    atten: torch.Tensor  # (A,B,H,W)
    em_cs: torch.Tensor  # (A,)
    theta: torch.Tensor  # (A,) radians
    I: torch.Tensor      # (A,B,W)
    C: torch.Tensor      # (B,H,W), updated in place

    def dims(self):
        A, B, H, W = (int(x) for x in self.atten.shape)
        return A, B, H, W


def _check(p: Problem, i: int):
    A, B, H, W = p.dims()
    want = {"atten": (A, B, H, W), "em_cs": (A,), "theta": (A,), "I": (A, B, W), "C": (B, H, W)}
    dev = p.atten.device
    for name, shape in want.items():
        t = getattr(p, name)
        if not isinstance(t, torch.Tensor) or tuple(t.shape) != shape:
            raise ValueError(f"problem {i}: {name} must have shape {shape}")
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.device == dev):
            raise ValueError(f"problem {i}: {name} must be a contiguous float32 tensor on {dev}")


def _structs(problems, shapes, thetas=None):
    arr = (_lib.MLEMProblem * len(problems))()
    for i, (p, (fa, br)) in enumerate(zip(problems, shapes)):
        A, B, H, W = p.dims()
        th = p.theta if thetas is None else thetas[i]
        arr[i] = _lib.MLEMProblem(p.atten.data_ptr(), p.em_cs.data_ptr(), th.data_ptr(), p.I.data_ptr(),
                                  p.C.data_ptr(), A, B, H, W, int(fa), int(br))
    return arr


PACK_MERGE_SLICES = 256  # slices up to which samples with equal angles share a pack


def merged_thetas(problems, cap=PACK_MERGE_SLICES):
    """The theta tensor each problem is planned with.  The plan packs problems
    that share a theta tensor and (A, H, W) along the slice axis.  Here the
    samples (the problems sharing one theta tensor) whose angles are equal value
    for value, with equal (A, H, W), are given one tensor, in order, as long as
    their slices total at most `cap`: more lanes of each wavefront then hold
    slices.  Reads every distinct theta tensor once (host copies)."""
    units = {}  # (theta pointer, H, W) -> [slices, (A, H, W), problem indices, theta]
    for i, p in enumerate(problems):
        A, B, H, W = p.dims()
        u = units.setdefault((p.theta.data_ptr(), H, W), [0, (A, H, W), [], p.theta])
        u[0] += B
        u[2].append(i)
    out = [p.theta for p in problems]
    groups = {}  # (A, H, W, angle bytes) -> [canonical tensor, slices so far]
    for slices, dims, members, theta in units.values():
        key = dims + (theta.detach().cpu().numpy().tobytes(),)
        g = groups.get(key)
        if g is None or g[1] + slices > cap:
            groups[key] = g = [theta, 0]
        g[1] += slices
        for i in members:
            out[i] = g[0]
    return out


def workspace_bytes(problems, shapes=None):
    """Bytes of device workspace fl_mlem_multi_create needs (host arithmetic only)."""
    shapes = shapes or [(0, 0)] * len(problems)
    arr = _structs(problems, shapes) if problems and isinstance(problems[0], Problem) else problems
    return int(_lib.load().fl_mlem_multi_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), len(arr)))


class BatchMLEM:
    """All problems of a batch, iterated together.

    shapes: None (library defaults), a list of (fwd_angles, bwd_rows) per
    problem, or an ``Autotuner``.  beta > 0 selects the One-Step-Late Huber
    update (atten_cuda.cpp:330-341), as in ``mlem_cuda_batch``.

    layout: 'slice_major' packs the problems that share a theta tensor and
    (A, H, W) -- the emission lines of one sample, and samples whose angles are
    equal, up to PACK_MERGE_SLICES slices (merged_thetas) -- along the slice
    axis and back-projects each pack from a slice-major copy of its attenuation (the
    workspace then holds that copy: about the attenuation's size again);
    'reference' back-projects every problem from its own attenuation; 'auto'
    takes 'slice_major' when its workspace fits the free device memory with
    2 GiB to spare.
    """

    def __init__(self, problems, beta=0.0, delta=0.01, shapes=None, layout="auto"):
        if not problems:
            raise ValueError("empty batch")
        for i, p in enumerate(problems):
            _check(p, i)
        self.device = problems[0].atten.device
        if any(p.atten.device != self.device for p in problems):
            raise ValueError("all problems must be on one device")
        if layout not in ("auto", "slice_major", "reference"):
            raise ValueError(f"layout must be 'auto', 'slice_major' or 'reference', not {layout!r}")
        self.problems = list(problems)  # keeps every array alive while the plan uses it
        self._thetas = merged_thetas(self.problems)
        # the workspace sizes do not depend on the tile shapes
        lib = _lib.load()
        arr = ctypes.cast(_structs(self.problems, [(0, 0)] * len(problems), self._thetas), ctypes.c_void_p)
        nbytes = int(lib.fl_mlem_multi_workspace_bytes(arr, len(problems)))
        if nbytes == 0:
            raise ValueError("invalid batch")
        sm_bytes = int(lib.fl_mlem_multi_workspace_bytes_sm(arr, len(problems)))
        if layout == "auto":
            free, _ = torch.cuda.mem_get_info(self.device)
            layout = "slice_major" if 0 < sm_bytes and sm_bytes + SLICE_MAJOR_HEADROOM <= free else "reference"
        if layout == "slice_major":
            if sm_bytes == 0:
                raise RuntimeError("the slice-major layout does not apply: a pack has H*W*sum(B) >= 2^29")
            nbytes = sm_bytes
        self.layout = layout
        if isinstance(shapes, Autotuner):
            shapes = [shapes.shapes(*p.dims(), device=self.device, layout=layout) for p in problems]
        self.shapes = list(shapes) if shapes is not None else [(0, 0)] * len(problems)
        if len(self.shapes) != len(problems):
            raise ValueError("one tile shape per problem")
        self._arr = _structs(self.problems, self.shapes, self._thetas)
        # 256-byte aligned workspace inside a slightly larger buffer
        self._ws_buf = torch.empty(nbytes + 256, dtype=torch.uint8, device=self.device)
        off = (-self._ws_buf.data_ptr()) % 256
        self._ws = self._ws_buf.data_ptr() + off
        self._nbytes = nbytes
        # graph capture needs a stream other than the NULL stream
        self.stream = torch.cuda.Stream(self.device)
        self._plan = ctypes.c_void_p()
        cur = torch.cuda.current_stream(self.device)
        self.stream.wait_stream(cur)
        with torch.cuda.device(self.device):
            _lib.call("fl_mlem_multi_create", ctypes.cast(self._arr, ctypes.c_void_p), len(problems),
                      float(beta), float(delta), ctypes.c_void_p(self._ws), nbytes,
                      ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(self._plan))
        cur.wait_stream(self.stream)

    def run(self, n_iter, graph=True):
        """n_iter iterations of every problem; C of each problem holds the result."""
        if not self._plan:
            raise RuntimeError("plan released")
        cur = torch.cuda.current_stream(self.device)
        self.stream.wait_stream(cur)
        with torch.cuda.device(self.device):
            _lib.call("fl_mlem_multi_run", self._plan, int(n_iter), 1 if graph else 0,
                      ctypes.c_void_p(self.stream.cuda_stream))
        cur.wait_stream(self.stream)

    def close(self):
        if self._plan:
            self.stream.synchronize()
            _lib.call("fl_mlem_multi_destroy", self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _default_cache():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "fl_correction_amd", "autotune.json")


class Autotuner:
    """Per size bucket tile shapes for the batched MLEM kernels.

    The bucket of a problem is (H, W) rounded up to multiples of 32.  A bucket
    is tuned once per device and layout: a synthetic batch of that bucket
    (enough slices to fill the GPU several times over) runs ``reps`` timed
    iterations under each candidate; the fastest is cached in ``path``.
    Reference layout: every (fwd_angles, bwd_rows) pair.  Slice-major layout:
    each fwd_angles (bwd_rows is unused there and recorded as 31), on three
    emission lines sharing one theta, which the plan packs.
    """

    def __init__(self, path=None, reps=3, tune_angles=120):
        self.path = path or os.environ.get("FL_AUTOTUNE_CACHE") or _default_cache()
        self.reps = int(reps)
        self.tune_angles = int(tune_angles)
        self.table = {}
        try:
            with open(self.path) as f:
                self.table = json.load(f)
        except (OSError, ValueError):
            self.table = {}
        self.timings = {}

    @staticmethod
    def bucket(H, W):
        return (-(-int(H) // 32) * 32, -(-int(W) // 32) * 32)

    @staticmethod
    def key(device_name, H, W, layout="reference"):
        bh, bw = Autotuner.bucket(H, W)
        return f"{device_name}|{layout}|{bh}x{bw}"

    def shapes(self, A, B, H, W, device=None, layout="reference"):
        if layout not in ("slice_major", "reference"):
            raise ValueError(f"layout must be 'slice_major' or 'reference', not {layout!r}")
        device = torch.device(device or "cuda")
        k = self.key(torch.cuda.get_device_name(device), H, W, layout)
        if not self.valid(self.table.get(k)):  # missing, or from another library version
            self.table[k] = list(self._tune(*self.bucket(H, W), device, layout))
            self._save()
        fa, br = self.table[k]
        return int(fa), int(br)

    @staticmethod
    def valid(entry):
        """A cached choice this library accepts: [angles per forward workgroup,
        rows per back-projector tile]."""
        return (isinstance(entry, (list, tuple)) and len(entry) == 2 and entry[0] in FWD_ANGLES
                and entry[1] in BWD_ROWS)

    def _save(self):
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # the choice is still used for this process

    def _tune(self, H, W, device, layout="reference"):
        A = self.tune_angles
        # slices so that the back-projector has >= 4 tiles per resident workgroup
        props = torch.cuda.get_device_properties(device)
        tiles_per_slice = -(-H // 31) * -(-W // 63)
        B = max(4, min(64, -(-4 * 4 * props.multi_processor_count // tiles_per_slice)))
        g = torch.Generator(device=device).manual_seed(1234)
        atten = torch.rand((A, B, H, W), generator=g, device=device) * 0.9 + 0.05
        em = torch.ones(A, device=device)
        theta = torch.arange(A, device=device, dtype=torch.float32) * (3.14159265 / A)
        I = torch.rand((A, B, W), generator=g, device=device) + 0.5
        C0 = torch.ones((B, H, W), device=device)
        lines = 3 if layout == "slice_major" else 1
        cands = [(fa, br) for fa in FWD_ANGLES for br in (BWD_ROWS if layout == "reference" else BWD_ROWS[:1])]
        best, best_ms = None, float("inf")
        for fa, br in cands:
            probs = [Problem(atten, em, theta, I, C0.clone()) for _ in range(lines)]
            run = BatchMLEM(probs, shapes=[(fa, br)] * lines, layout=layout)
            run.run(1, graph=False)
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record(run.stream)
            run.run(2 * self.reps, graph=True)
            stop.record(run.stream)
            stop.synchronize()
            ms = start.elapsed_time(stop) / (2 * self.reps)
            run.close()
            self.timings[(layout, H, W, fa, br)] = ms
            if ms < best_ms:
                best, best_ms = (fa, br), ms
        return best
