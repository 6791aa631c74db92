"""Differentiable attenuated XRF forward model on the GPU.

For concentrations C (K, B, H, W) -- one map per element, each element carrying
its own emission line -- and angles theta_a:

    F_j   = Rot_a(C_j)                                (torch_rotate_image_nd, FL/core.py:43)
    A_x,e = prod_{m>r} (1 - coef_x[a,e] F_e[m])       (torch_xray_atten_3D, FL/torch_functions.py:83-106)
    A_f,e = exp(-Cone(sum_j coef[a,j,e] F_j))         (torch_xrf_atten_3D, FL/torch_functions.py:110-176)
    A_e   = A_f,e * A_x,e                             (cal_total_atten_at_angles, FL/core.py:84-101)
    Pf_e  = em[a,e] * sum_h A_e * S_a(C_e)            (forward_emission_batch, FL/cuda_lib/atten_cuda.cpp:142)

``forward`` returns Pf (K, A, B, W).  ``vjp`` returns dL/dC for a cotangent
G = dL/dPf, including the dependence of the attenuation on C that the
reference's autograd wrapper (FL/cuda_lib/autograd.py:40-74) leaves out.
Attenuation exponents are kept for the backward pass when they fit the memory
budget and recomputed per angle otherwise.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .atten_util import coef_xray, coef_xrf, extract_single_value_param
from .core import angle_entries
from .detector_mask import Cone


def _f32(x, dev):
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32), device=dev)


class XRFModel:
    """Attenuated emission model for a fixed geometry.

    param:      the parameter dictionary (atten_util.make_param / load_param layout)
    angle_list: angles in degrees
    detector:   a dense mask (Mz, Mx, My) or a Cone
    em:         (A, K) emission factors; default em_cs[elem] * pix per element
    store:      'auto' | True | False -- keep the fluorescence exponents between forward and vjp
    """

    def __init__(self, param, angle_list, detector, em=None, position_det="r", store="auto",
                 memory_fraction=0.4, device="cuda"):
        self.device = torch.device(device)
        self.angles = np.asarray(angle_list, np.float64).reshape(-1)
        A = len(self.angles)
        elem = param["elem_type"]
        K = len(elem)
        self.A, self.K = A, K
        cx, cf = [], []
        for a in range(A):
            p = extract_single_value_param(param, a)
            cx.append(coef_xray(p))
            cf.append(coef_xrf(p))
        self.coef_x = _f32(np.stack(cx), self.device)       # (A, K)
        self.coef_xrf = _f32(np.stack(cf), self.device)     # (A, Kc, K)
        if em is None:
            em = np.zeros((A, K))
            for e, el in enumerate(elem):
                v = np.atleast_1d(np.asarray(param["em_cs"][el], np.float64))
                em[:, e] = (v if v.size == A else np.full(A, v[0])) * param["pix"]
        em = np.asarray(em, np.float32).reshape(A, K)
        self.em_T = _f32(em.T, self.device)                # (K, A): per-line, contiguous
        self.theta = _f32(self.angles / 180.0 * np.pi, self.device)
        self.trig = [angle_entries(a) for a in self.angles]
        self.side = {"r": 0, "l": 1}[position_det]
        self.detector = detector
        if isinstance(detector, Cone):
            self._dirs = [np.ascontiguousarray(v) for v in (detector.pz, detector.py, detector.qd, detector.weight)]
        else:
            self.mask = _f32(detector, self.device)
        self.store = store
        self.memory_fraction = memory_fraction
        self._T_store = None
        self._sweep_ws = {}

    # ------------------------------------------------------------------
    def _st(self):
        return _lib.stream_of(self.device)

    def _dir_ptrs(self):
        return [d.ctypes.data_as(_lib.ctypes.c_void_p) for d in self._dirs]

    def _sweep_workspace(self, shape):
        """Workspace of the all-directions cone sweep for the current stream (one per
        stream: concurrent angles must not share its ring), or None when the sweep
        does not apply to this direction set."""
        K, B, H, W = shape
        st = self._st()
        key = (int(getattr(st, "value", st) or 0), K, B, H, W)
        ws = self._sweep_ws.get(key)
        if ws is None:
            nb = int(_lib.load().fl_cone_sweep_workspace_bytes(K, B, W, H, *self._dir_ptrs()[:3],
                                                               self.detector.ndir))
            ws = False if nb == 0 else torch.empty((nb + 3) // 4, dtype=torch.float32, device=self.device)
            self._sweep_ws[key] = ws
        return None if ws is False else ws

    def _cone(self, x, out, shape, adjoint):
        K, B, H, W = shape
        ws = self._sweep_workspace(shape)
        if ws is not None:
            _lib.call("fl_cone_sweep", _lib.ptr(x), _lib.ptr(out), K, B, W, H, *self._dir_ptrs(), self.detector.ndir,
                      int(adjoint), _lib.ptr(ws), ws.numel() * 4, self._st())
        else:
            fn = "fl_cone_atten_adjoint" if adjoint else "fl_cone_atten"
            _lib.call(fn, _lib.ptr(x), _lib.ptr(out), K, B, W, H, *self._dir_ptrs(), self.detector.ndir, self._st())

    def _exponent(self, mu_rot, T_rot, shape):
        K, B, H, W = shape
        if isinstance(self.detector, Cone):
            self._cone(mu_rot, T_rot, shape, adjoint=False)
        else:
            Mz, Mx, My = self.mask.shape
            _lib.call("fl_cal_atten", _lib.ptr(mu_rot), _lib.ptr(self.mask), _lib.ptr(T_rot), K, B, W, H, Mz, Mx, My,
                      self._st())

    def _exponent_adjoint(self, g_rot, out, shape):
        K, B, H, W = shape
        if isinstance(self.detector, Cone):
            self._cone(g_rot, out, shape, adjoint=True)
        else:
            Mz, Mx, My = self.mask.shape
            _lib.call("fl_cal_atten_adjoint", _lib.ptr(g_rot), _lib.ptr(self.mask), _lib.ptr(out), K, B, W, H, Mz, Mx,
                      My, self._st())

    def _angle_maps(self, C, a, F, Ax, mu_rot, T_rot, recompute_T=True):
        """F = Rot(C), Ax = incident, T_rot = exponent (mask frame) for angle index a."""
        K, B, H, W = C.shape
        c, s = self.trig[a]
        st = self._st()
        _lib.call("fl_rotate", _lib.ptr(C), _lib.ptr(F), K * B, H, W, c, s, st)
        _lib.call("fl_incident_atten", _lib.ptr(F), _lib.ptr(self.coef_x[a]), _lib.ptr(Ax), K, B, H, W, B * H * W,
                  B * H * W, st)
        if recompute_T:
            _lib.call("fl_mu_xrf_rot", _lib.ptr(F), _lib.ptr(self.coef_xrf[a]), _lib.ptr(mu_rot), K, K, B, H, W,
                      self.side, st)
            self._exponent(mu_rot, T_rot, (K, B, H, W))

    def _use_store(self, C, streams=2, angle_chunk=16):
        """Whether to keep the exponents: with 'auto', when they fit memory_fraction of the
        free memory left after the vjp's own scratch (per stream ten maps of C's size and
        one accumulator, plus one chunk of attenuation maps)."""
        if self.store == "auto":
            vol = 4 * C.numel()
            need = self.A * vol
            n_st = max(1, min(int(streams), self.A))
            scratch = n_st * 11 * vol + min(angle_chunk, self.A) * vol
            free, _ = torch.cuda.mem_get_info(self.device)
            return need <= self.memory_fraction * max(free - scratch, 0)
        return bool(self.store)

    @staticmethod
    def _join(cur, side):
        """Make the caller's stream `cur` wait for everything queued on the side streams."""
        for st in side:
            done = torch.cuda.Event()
            done.record(st)
            cur.wait_event(done)

    # ------------------------------------------------------------------
    def attenuation(self, C, out=None, streams=2, lines=None):
        """Total attenuation (K, A, B, H, W) for concentrations C (K, B, H, W), or
        (len(lines), A, B, H, W) for the emission lines listed in `lines` (every
        line's exponent is still computed: the lines share the cone sweep; only
        the listed lines' maps are written).

        The angles are independent: they are dealt round-robin to `streams`
        streams (the caller's current stream and default-priority side streams),
        each with its own scratch maps, so that one angle's kernels overlap
        another's.  Every angle is computed exactly as with one stream.  The
        caller's stream waits for the side streams before this returns."""
        C = C.to(self.device, torch.float32).contiguous()
        K, B, H, W = C.shape
        sel = None if lines is None else [int(e) for e in lines]
        if sel is not None and (not sel or any(e < 0 or e >= K for e in sel)):
            raise ValueError(f"lines must be indices of the {K} emission lines")
        if out is None:
            out = torch.empty((K if sel is None else len(sel), self.A, B, H, W), dtype=torch.float32,
                              device=self.device)
        n_st = max(1, min(int(streams), self.A))
        with torch.cuda.device(self.device):
            cur = torch.cuda.current_stream(self.device)
            sts = [cur] + self._side_streams(n_st - 1)
            # scratch per stream, allocated on the caller's stream (released after it
            # has waited for the side streams, so no reuse can overtake them)
            bufs = []
            for _ in range(n_st):
                mu_rot = torch.empty((K, B, W, H), dtype=torch.float32, device=self.device)
                bufs.append((torch.empty_like(C), torch.empty_like(C), mu_rot, torch.empty_like(mu_rot)))
            ready = torch.cuda.Event()
            ready.record(cur)
            for st in sts[1:]:
                st.wait_event(ready)
            try:
                for a in range(self.A):
                    i = a % n_st
                    with torch.cuda.stream(sts[i]):
                        F, Ax, mu_rot, T_rot = bufs[i]
                        self._angle_maps(C, a, F, Ax, mu_rot, T_rot)
                        if sel is None:
                            _lib.call("fl_xrf_combine", _lib.ptr(T_rot), _lib.ptr(Ax), _lib.ptr(out[0, a]), K, B, H,
                                      W, B * H * W, out.stride(0), self.side, self._st())
                        else:
                            for j, e in enumerate(sel):
                                _lib.call("fl_xrf_combine", _lib.ptr(T_rot[e]), _lib.ptr(Ax[e]), _lib.ptr(out[j, a]),
                                          1, B, H, W, B * H * W, out.stride(0), self.side, self._st())
            finally:
                # also on an error: the scratch returns to the caller's stream's pool only
                # after the side streams are done with it
                self._join(cur, sts[1:])
        return out

    def _side_streams(self, n):
        """n default-priority streams of this model's device, created once."""
        if not hasattr(self, "_streams"):
            self._streams = []
        while len(self._streams) < n:
            self._streams.append(torch.cuda.Stream(device=self.device))
        return self._streams[:n]

    def forward(self, C, angle_chunk=16, streams=2):
        """Pf (K, A, B, W).  Keeps the fluorescence exponents for vjp when storing.

        Within each chunk of angles the attenuation maps are dealt round-robin to
        `streams` streams (as in `attenuation`); the chunk's projections then run
        on the caller's stream after it has waited for them.  Results equal the
        one-stream ones bit for bit."""
        C = C.to(self.device, torch.float32).contiguous()
        K, B, H, W = C.shape
        if K != self.K:
            raise RuntimeError(f"C has {K} element maps, the model {self.K}")
        self._T_store = None
        store = self._use_store(C, streams, angle_chunk)
        self._T_store = torch.empty((self.A, K, B, W, H), dtype=torch.float32, device=self.device) if store else None
        Pf = torch.empty((K, self.A, B, W), dtype=torch.float32, device=self.device)
        nA = min(angle_chunk, self.A)
        Atot = torch.empty((K, nA, B, H, W), dtype=torch.float32, device=self.device)
        n_st = max(1, min(int(streams), nA))
        bufs = []
        for _ in range(n_st):
            mu_rot = torch.empty((K, B, W, H), dtype=torch.float32, device=self.device)
            bufs.append((torch.empty_like(C), torch.empty_like(C), mu_rot, torch.empty_like(mu_rot)))
        st = self._st()
        with torch.cuda.device(self.device):
            cur = torch.cuda.current_stream(self.device)
            sts = [cur] + self._side_streams(n_st - 1)
            for a0 in range(0, self.A, nA):
                a1 = min(a0 + nA, self.A)
                # the side streams start after C is ready and the previous chunk's
                # projections (which read Atot) are done
                ready = torch.cuda.Event()
                ready.record(cur)
                for sx in sts[1:]:
                    sx.wait_event(ready)
                try:
                    for a in range(a0, a1):
                        i = (a - a0) % n_st
                        with torch.cuda.stream(sts[i]):
                            F, Ax, mu_rot, T_tmp = bufs[i]
                            T_rot = self._T_store[a] if store else T_tmp
                            self._angle_maps(C, a, F, Ax, mu_rot, T_rot)
                            _lib.call("fl_xrf_combine", _lib.ptr(T_rot), _lib.ptr(Ax), _lib.ptr(Atot[0, a - a0]), K,
                                      B, H, W, B * H * W, Atot.stride(0), self.side, self._st())
                finally:
                    self._join(cur, sts[1:])
                for e in range(K):
                    _lib.call("fl_forward_emission_batch", _lib.ptr(Atot[e]), _lib.ptr(C[e]), _lib.ptr(self.em_T[e, a0:]),
                              _lib.ptr(self.theta[a0:]), _lib.ptr(Pf[e, a0]), a1 - a0, B, 1, H, W, st)
        return Pf

    def vjp(self, C, G, angle_chunk=16, streams=2):
        """dL/dC (K, B, H, W) for G = dL/dPf (K, A, B, W).

        Within each chunk of angles the per-angle adjoints are dealt round-robin
        to `streams` streams, each with its own scratch and its own dL/dC
        accumulator; the side accumulators are added at the end.  (The sum over
        angles is then grouped by stream, so results match the one-stream ones
        to rounding, not bit for bit.)"""
        C = C.to(self.device, torch.float32).contiguous()
        G = G.to(self.device, torch.float32).contiguous()
        K, B, H, W = C.shape
        if tuple(G.shape) != (K, self.A, B, W):
            raise RuntimeError(f"G must be {(K, self.A, B, W)}")
        stored = self._T_store is not None
        nA = min(angle_chunk, self.A)
        Atot = torch.empty((K, nA, B, H, W), dtype=torch.float32, device=self.device)
        n_st = max(1, min(int(streams), nA))
        acc = [torch.zeros_like(C) for _ in range(n_st)]  # acc[0] is the result
        bufs = []
        for _ in range(n_st):
            F, Ax, gA, gx, dmux, dF = (torch.empty_like(C) for _ in range(6))
            mu_rot = torch.empty((K, B, W, H), dtype=torch.float32, device=self.device)
            bufs.append((F, Ax, gA, gx, dmux, dF, mu_rot, torch.empty_like(mu_rot), torch.empty_like(mu_rot),
                         torch.empty_like(mu_rot)))
        st = self._st()
        dC = acc[0]
        with torch.cuda.device(self.device):
            cur = torch.cuda.current_stream(self.device)
            sts = [cur] + self._side_streams(n_st - 1)
            for a0 in range(0, self.A, nA):
                a1 = min(a0 + nA, self.A)
                # side streams start after the inputs are ready and the previous
                # chunk's projections (which read Atot) are done
                ready = torch.cuda.Event()
                ready.record(cur)
                for sx in sts[1:]:
                    sx.wait_event(ready)
                try:
                    self._vjp_angles(C, G, a0, a1, n_st, sts, bufs, acc, Atot, self._T_store if stored else None, 0)
                finally:
                    self._join(cur, sts[1:])
                for e in range(K):
                    _lib.call("fl_backward_emission_accumulate", _lib.ptr(Atot[e]), _lib.ptr(G[e, a0]),
                              _lib.ptr(self.em_T[e, a0:]), _lib.ptr(self.theta[a0:]), _lib.ptr(dC[e]), a1 - a0, B, 1,
                              H, W, st)
            for extra in acc[1:]:
                dC.add_(extra)  # on the caller's stream, after it has waited for the side streams
        return dC

    def _vjp_angles(self, C, G, a0, a1, n_st, sts, bufs, acc, Atot, T_keep, t0):
        """The per-angle adjoints of angles a0..a1-1, dealt round-robin to the streams.
        Angle a's exponent is T_keep[a - t0] when T_keep is given, else recomputed."""
        K, B, H, W = C.shape
        for a in range(a0, a1):
            i = (a - a0) % n_st
            with torch.cuda.stream(sts[i]):
                sa = self._st()
                F, Ax, gA, gx, dmux, dF, mu_rot, T_tmp, dT_rot, dmu_rot = bufs[i]
                T_rot = T_tmp if T_keep is None else T_keep[a - t0]
                self._angle_maps(C, a, F, Ax, mu_rot, T_rot, recompute_T=T_keep is None)
                for e in range(K):
                    _lib.call("fl_sample_grad", _lib.ptr(C[e]), _lib.ptr(self.em_T[e, a:]),
                              _lib.ptr(self.theta[a:]), _lib.ptr(G[e, a]), _lib.ptr(gA[e]), 1, B, 1, H, W, 0, sa)
                _lib.call("fl_xrf_grad_prep", _lib.ptr(T_rot), _lib.ptr(Ax), _lib.ptr(gA), _lib.ptr(dT_rot),
                          _lib.ptr(gx), _lib.ptr(Atot[0, a - a0]), Atot.stride(0), K, B, H, W, self.side, sa)
                self._exponent_adjoint(dT_rot, dmu_rot, (K, B, H, W))
                _lib.call("fl_incident_grad", _lib.ptr(F), _lib.ptr(self.coef_x[a]), _lib.ptr(Ax),
                          _lib.ptr(gx), _lib.ptr(dmux), K, B, H, W, sa)
                _lib.call("fl_mu_adjoint", _lib.ptr(dmu_rot), _lib.ptr(dmux), _lib.ptr(self.coef_xrf[a]),
                          _lib.ptr(self.coef_x[a]), _lib.ptr(dF), K, K, B, H, W, self.side, sa)
                c, s = self.trig[a]
                _lib.call("fl_rotate_adjoint", _lib.ptr(dF), _lib.ptr(acc[i]), K * B, H, W, c, s, 1, sa)

    def loss_and_grad(self, C, I, eps=1e-6, angle_chunk=16, streams=2):
        """Poisson loss mean(P - I log P), P = max(Pf, eps) (FL/ml_lib/ml_loss.py:6-8),
        its gradient dL/dC and the projections Pf, in one pass over the angles.

        The loss's cotangent dL/dPf at angle a depends on angle a's projections
        only, so each chunk of angles is projected, its cotangent formed and its
        adjoint run while the chunk's fluorescence exponents are still resident:
        the cone runs twice per angle (exponent, then its adjoint), where forward
        followed by vjp runs it three times unless every angle's exponent is kept.
        Memory: one chunk of exponents and attenuation maps, not one per angle.
        Returns (loss as a 0-d tensor, dL/dC (K, B, H, W), Pf (K, A, B, W)); they
        equal poisson_loss(forward(C), I) and vjp's result to rounding."""
        C = C.detach().to(self.device, torch.float32).contiguous()
        I = I.to(self.device, torch.float32).contiguous()
        K, B, H, W = C.shape
        if K != self.K:
            raise RuntimeError(f"C has {K} element maps, the model {self.K}")
        if tuple(I.shape) != (K, self.A, B, W):
            raise RuntimeError(f"I must be {(K, self.A, B, W)}")
        nA = min(angle_chunk, self.A)
        n_st = max(1, min(int(streams), nA))
        Pf = torch.empty((K, self.A, B, W), dtype=torch.float32, device=self.device)
        G = torch.empty_like(Pf)
        Atot = torch.empty((K, nA, B, H, W), dtype=torch.float32, device=self.device)
        T_chunk = torch.empty((nA, K, B, W, H), dtype=torch.float32, device=self.device)
        acc = [torch.zeros_like(C) for _ in range(n_st)]  # acc[0] is the result
        bufs = []
        for _ in range(n_st):
            F, Ax, gA, gx, dmux, dF = (torch.empty_like(C) for _ in range(6))
            mu_rot = torch.empty((K, B, W, H), dtype=torch.float32, device=self.device)
            bufs.append((F, Ax, gA, gx, dmux, dF, mu_rot, torch.empty_like(mu_rot), torch.empty_like(mu_rot),
                         torch.empty_like(mu_rot)))
        inv_n = 1.0 / Pf.numel()
        st = self._st()
        dC = acc[0]
        with torch.cuda.device(self.device):
            cur = torch.cuda.current_stream(self.device)
            sts = [cur] + self._side_streams(n_st - 1)

            def fork():
                ready = torch.cuda.Event()
                ready.record(cur)
                for sx in sts[1:]:
                    sx.wait_event(ready)

            for a0 in range(0, self.A, nA):
                a1 = min(a0 + nA, self.A)
                fork()  # the previous chunk's emission adjoint (reads Atot) is done
                try:
                    for a in range(a0, a1):
                        i = (a - a0) % n_st
                        with torch.cuda.stream(sts[i]):
                            F, Ax, mu_rot = bufs[i][0], bufs[i][1], bufs[i][6]
                            T_rot = T_chunk[a - a0]
                            self._angle_maps(C, a, F, Ax, mu_rot, T_rot)
                            _lib.call("fl_xrf_combine", _lib.ptr(T_rot), _lib.ptr(Ax), _lib.ptr(Atot[0, a - a0]), K,
                                      B, H, W, B * H * W, Atot.stride(0), self.side, self._st())
                finally:
                    self._join(cur, sts[1:])
                for e in range(K):
                    _lib.call("fl_forward_emission_batch", _lib.ptr(Atot[e]), _lib.ptr(C[e]), _lib.ptr(self.em_T[e, a0:]),
                              _lib.ptr(self.theta[a0:]), _lib.ptr(Pf[e, a0]), a1 - a0, B, 1, H, W, st)
                # dL/dPf of the chunk, as autograd forms it through clamp_min, log and mean
                Pc, Ic = Pf[:, a0:a1], I[:, a0:a1]
                P = Pc.clamp_min(eps)
                G[:, a0:a1] = torch.where(Pc >= eps, inv_n - (Ic * inv_n) / P, torch.zeros((), device=self.device))
                fork()
                try:
                    self._vjp_angles(C, G, a0, a1, n_st, sts, bufs, acc, Atot, T_chunk, a0)
                finally:
                    self._join(cur, sts[1:])
                for e in range(K):
                    _lib.call("fl_backward_emission_accumulate", _lib.ptr(Atot[e]), _lib.ptr(G[e, a0]),
                              _lib.ptr(self.em_T[e, a0:]), _lib.ptr(self.theta[a0:]), _lib.ptr(dC[e]), a1 - a0, B, 1,
                              H, W, st)
            for extra in acc[1:]:
                dC.add_(extra)
            P = Pf.clamp_min(eps)
            loss = (P - I * torch.log(P)).mean()
        return loss, dC, Pf

    def release(self):
        self._T_store = None


class XRFModelFunction(torch.autograd.Function):
    """Pf = model(C) with the hand-written adjoint as its backward.  `streams` is passed
    to the model's forward and vjp (1 keeps every launch on the caller's stream)."""

    @staticmethod
    def forward(ctx, C, model, streams=2):
        ctx.model = model
        ctx.streams = streams
        ctx.save_for_backward(C)
        return model.forward(C.detach(), streams=streams)

    @staticmethod
    def backward(ctx, G):
        (C,) = ctx.saved_tensors
        return ctx.model.vjp(C.detach(), G, streams=ctx.streams), None, None


def xrf_forward_autograd(C, model, streams=2):
    return XRFModelFunction.apply(C, model, streams)
