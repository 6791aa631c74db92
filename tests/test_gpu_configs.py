"""Parity at the workloads' own sizes (fl_correction_amd.workloads), against the oracle.

* plain_64 (64^3, 2 lines, 60 angles, 3x3 cone over 10 deg, 20 MLEM iterations): the whole
  problem.  Attenuation maps on every 6th angle against the FP64 oracle chain, two angles
  against the reference's own op on the cone's splat mask (FP32 same order); MLEM of each
  line against the FP32 oracle (FP64 error printed).
* loop_128 (128^3, 3 lines, 180 angles, 5x5 cone over 15 deg, 50 iterations, attenuation
  recomputed every 5): the full loop on the GPU; the oracle checks the first two blocks and
  the last one on 4 slices (MLEM is slice-local) and the recomputed attenuation on a slab
  of slices at 4 angles; full-size properties: the cone's adjoint identity, determinism.
* model_256: the gradient on the 32^3 crop (4 lines, 5x5 cone over 15 deg) against FP64
  central differences at 64 voxels (tests/golden/model_crop_fd.npz) and FP64 torch
  autograd; at 256^3 x 4 lines on a few angles: stored and recomputed exponents give the
  same results bit for bit, the vjp agrees with a central difference along a random
  direction, attenuation and projections on a slab against the oracle.
* batch_48: the 48 uneven samples, 20 iterations under the hipGraph, one slice of every
  problem against the FP32 oracle (1e-5), or, for the few problems where the FP32 oracle's
  own serial order is already ~1e-5 from the FP64 oracle, no farther than 2x that from FP64.
* joint_512 (128 x 512^2, 5 lines, 7x7 cone over 20 deg): two angles at full size, the
  attenuation on a strip of one slice and the projections of that slice against the oracle,
  the vjp against a central difference along a random direction.
"""
import os

import numpy as np
import pytest
import torch

import helpers
from helpers import RTOL, oracle, rel_l2

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def _free():
    yield
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _theta(angles):
    return (np.asarray(angles) / 180.0 * np.pi).astype(np.float32)


# ---------------------------------------------------------------------------
# plain_64: the whole problem
# ---------------------------------------------------------------------------
def test_plain_64_whole_problem():
    from fl_correction_amd import workloads
    from fl_correction_amd.cuda_lib import atten_cuda
    W = workloads.PLAIN_64
    _say(f"{W.name}: {W.shape} x {W.K} lines x {W.A} angles, cone {W.cone_n}x{W.cone_n} over {W.cone_half} deg")
    C = W.phantom()
    ang, th = W.angles(), _theta(W.angles())
    cx, cf = W.coefs()
    cone = W.cone()
    model = W.model()
    atten = model.attenuation(_dev(C)).cpu().numpy()  # (K, A, B, H, W)
    worst = 0.0
    for a in range(0, W.A, 6):
        ref = oracle.total_atten_cone_f64(C, ang[a], cx, cf, cone)
        e = rel_l2(atten[:, a], ref)
        worst = max(worst, e)
        assert e <= RTOL, (a, e)
    # the reference's op (atten_kernel) on the cone's splat mask, FP32 same order
    M = cone.splat_mask(W.shape[2])
    for a in (7, 41):
        rot = oracle.rotate_torch(C, th[a])
        ref32 = oracle.xrf_atten_from_mask(oracle.mu_xrf(rot, cf), M) * oracle.incident_f32(oracle.mu_xray(rot, cx))
        assert rel_l2(atten[:, a], ref32) <= RTOL
    _say(f"  attenuation: worst rel-L2 vs FP64 {worst:.2e} over {len(range(0, W.A, 6))} angles")
    em = np.full((W.A, 1), workloads.PIX, np.float32)
    for e in range(W.K):
        Pf = oracle.forward_batch_f32(atten[e], C[e:e + 1], em, th)
        I = workloads.noisy(Pf, seed=W.seed + 10 + e)
        C0 = np.ones((1,) + W.shape, np.float32)
        got = atten_cuda.mlem_cuda_batch(_dev(C0), _dev(atten[e]), _dev(em), _dev(th), _dev(I), 20, 0.0,
                                         0.01).cpu().numpy()
        ref = oracle.mlem_batch_f32(C0, atten[e], em, th, I, 20, 0.0, 0.01)
        ref64 = oracle.mlem_batch_f64(C0, atten[e], em, th, I, 20, 0.0, 0.01)
        e32, e64 = rel_l2(got, ref), rel_l2(got, ref64)
        _say(f"  line {e}: 20 MLEM iterations, rel-L2 vs FP32 oracle {e32:.2e}, vs FP64 {e64:.2e}")
        assert e32 <= RTOL


# ---------------------------------------------------------------------------
# loop_128: attenuation recomputed every 5 iterations, 50 iterations
# ---------------------------------------------------------------------------
def _cone_ptrs(cone):
    from fl_correction_amd import _lib
    d = [np.ascontiguousarray(v) for v in (cone.pz, cone.py, cone.qd, cone.weight)]
    return d, [v.ctypes.data_as(_lib.ctypes.c_void_p) for v in d]


def test_loop_128_blocks_against_oracle():
    from fl_correction_amd import joint, workloads
    from fl_correction_amd.cuda_lib import atten_cuda
    W = workloads.LOOP_128
    _say(f"{W.name}: {W.shape} x {W.K} lines x {W.A} angles, 50 iterations, recompute every 5")
    C_true = W.phantom()
    model = W.model()
    Cd = _dev(C_true)
    At = model.attenuation(Cd)
    I = np.stack([workloads.noisy(atten_cuda.forward_emission_batch(At[e], Cd[e:e + 1], model.em_T[e][:, None]
                                                                    .contiguous(), model.theta).cpu().numpy(),
                                  seed=W.seed + 20 + e) for e in range(W.K)])
    del At
    snaps = {}

    def keep(done, C):
        if done in (5, 10, 45, 50):
            snaps[done] = C.detach().cpu().numpy().copy()

    C0 = torch.ones_like(Cd)
    joint.self_consistent_mlem(model, _dev(I), C0, n_iter=50, recompute_every=5, callback=keep,
                               start_unattenuated=True)
    torch.cuda.synchronize()
    S = [9, 50, 77, 121]  # slices checked
    th = _theta(W.angles())
    em = np.full((W.A, 1), workloads.PIX, np.float32)
    ones = np.ones((W.A, len(S)) + W.shape[1:], np.float32)
    # block 1: unattenuated
    for e in range(W.K):
        ref = oracle.mlem_batch_f32(np.ones((1, len(S)) + W.shape[1:], np.float32), ones, em, th,
                                    np.ascontiguousarray(I[e][:, S]), 5, 0.0, 0.01)
        err = rel_l2(snaps[5][e][S], ref[0])
        _say(f"  block 1 line {e}: rel-L2 {err:.2e}")
        assert err <= RTOL
    # blocks 2 and 10: attenuation recomputed from the estimate, then 5 iterations
    cx, cf = W.coefs()
    ang = W.angles()
    for start, end in ((5, 10), (45, 50)):
        Cs = snaps[start]
        A_s = model.attenuation(_dev(Cs))
        for a in (0, 44, 97, 151):
            ref = oracle.total_atten_cone_f64(Cs, ang[a], cx, cf, model.detector, 63, 65)
            got = A_s[:, a, 63:65].cpu().numpy()
            err = rel_l2(got, ref)
            _say(f"  attenuation of the estimate after {start}, angle {ang[a]:.0f}: rel-L2 vs FP64 {err:.2e}")
            assert err <= RTOL
        A_S = A_s[:, :, S].cpu().numpy()  # (K, A, |S|, H, W)
        for e in range(W.K):
            ref = oracle.mlem_batch_f32(np.ascontiguousarray(Cs[e][S])[None], np.ascontiguousarray(A_S[e]), em, th,
                                        np.ascontiguousarray(I[e][:, S]), end - start, 0.0, 0.01)
            err = rel_l2(snaps[end][e][S], ref[0])
            _say(f"  iterations {start + 1}-{end} line {e}: rel-L2 {err:.2e}")
            assert err <= RTOL
        del A_s
    assert np.isfinite(snaps[50]).all() and snaps[50].min() >= 0


def test_loop_128_full_size_properties():
    """The cone march's adjoint identity and bitwise determinism at 128^3 x 3 lines."""
    from fl_correction_amd import _lib, workloads
    W = workloads.LOOP_128
    K, (B, H, Wd) = W.K, W.shape
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.rand((K, B, Wd, H), device="cuda", generator=g) * 0.05
    y = torch.rand((K, B, Wd, H), device="cuda", generator=g) - 0.5
    _, ptrs = _cone_ptrs(W.cone())
    st = _lib.stream_of(x.device)
    Kx, Kx2, Kty = torch.empty_like(x), torch.empty_like(x), torch.empty_like(y)
    _lib.call("fl_cone_atten", _lib.ptr(x), _lib.ptr(Kx), K, B, Wd, H, *ptrs, W.cone().ndir, st)
    _lib.call("fl_cone_atten", _lib.ptr(x), _lib.ptr(Kx2), K, B, Wd, H, *ptrs, W.cone().ndir, st)
    _lib.call("fl_cone_atten_adjoint", _lib.ptr(y), _lib.ptr(Kty), K, B, Wd, H, *ptrs, W.cone().ndir, st)
    assert torch.equal(Kx, Kx2)
    lhs, rhs = float((Kx.double() * y.double()).sum()), float((x.double() * Kty.double()).sum())
    _say(f"  cone adjoint identity at {B}x{Wd}x{H} x {K}: {lhs:.6e} vs {rhs:.6e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    model = W.model()
    C = W.phantom_device()
    a1 = model.attenuation(C, streams=2)
    a2 = model.attenuation(C, streams=1)
    assert torch.equal(a1, a2)
    assert float(a1.min()) > 0 and float(a1.max()) <= 1.0


# ---------------------------------------------------------------------------
# model_256: the gradient
# ---------------------------------------------------------------------------
def test_model_crop_gradient_against_fd_and_autograd():
    import torch_model as tm
    from fl_correction_amd.model import XRFModel
    case = helpers.model_crop_case()
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "model_crop_fd.npz"))
    model = XRFModel(case["param"], case["angles"], case["cone"], em=case["em"])
    C = _dev(case["C_eval"])
    Pf = model.forward(C)
    I = _dev(case["I"])
    n = Pf.numel()
    G = (1.0 - I / Pf) / n  # dL/dPf of mean(Pf - I log Pf)
    g = model.vjp(C, G).cpu().numpy()
    loss = float((Pf.double() - I.double() * torch.log(Pf.double())).mean())
    _say(f"model crop 32^3 x 4 lines x {len(case['angles'])} angles: loss {loss:.6e} (FP64 oracle {float(fx['loss']):.6e})")
    assert abs(loss - float(fx["loss"])) <= 1e-5 * abs(float(fx["loss"]))
    gi = np.array([g[tuple(ix)] for ix in fx["idx"]])
    fd = fx["fd"]
    err = np.abs(gi - fd).max() / np.abs(fd).max()
    _say(f"  vjp vs FP64 central differences at {len(fd)} voxels: max error {err:.2e} of max |fd|, "
         f"rel-L2 {rel_l2(gi, fd):.2e}")
    assert err <= 1e-4 and rel_l2(gi, fd) <= 1e-4
    # FP64 torch autograd of the whole gradient
    Ct = torch.tensor(case["C_eval"], dtype=torch.float64, requires_grad=True)
    mask = case["cone"].splat_mask(case["C_eval"].shape[3])
    P = tm.forward(Ct, case["angles"], case["coef_x"], case["coef_xrf"], case["em"], mask)
    It = torch.tensor(case["I"], dtype=torch.float64)
    (P - It * torch.log(P)).mean().backward()
    e_ag = rel_l2(g, Ct.grad.numpy())
    _say(f"  vjp vs FP64 torch autograd: rel-L2 {e_ag:.2e}")
    assert e_ag <= 1e-4


def _directional_check(model, C, seed, eps=1e-2):
    """<vjp(G), v> against the central difference of <G, Pf(C + t v)> at t = +-eps,
    v = C * u with u uniform in [-1, 1] (a relative perturbation)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    u = torch.rand(C.shape, device="cuda", generator=g) * 2 - 1
    v = C * u
    Pf = model.forward(C, streams=1)
    G = torch.rand(Pf.shape, device="cuda", generator=g) - 0.5
    d = float((model.vjp(C, G, streams=1).double() * v.double()).sum())
    lp = float((model.forward(C + eps * v, streams=1).double() * G.double()).sum())
    lm = float((model.forward(C - eps * v, streams=1).double() * G.double()).sum())
    fd = (lp - lm) / (2 * eps)
    return d, fd


def test_model_256_full_size_angles():
    from fl_correction_amd import workloads
    from fl_correction_amd.model import XRFModel
    W = workloads.MODEL_256
    angles = W.angles()[[0, 91, 180, 287]]
    _say(f"{W.name}: {W.shape} x {W.K} lines, angles {angles.tolist()}")
    C = W.phantom_device() * 0.5 + 0.05
    stored = XRFModel(W.param(), angles, W.cone(), store=True)
    again = XRFModel(W.param(), angles, W.cone(), store=False)
    P1, P2 = stored.forward(C), again.forward(C)
    assert torch.equal(P1, P2)
    G = torch.rand(P1.shape, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3)) - 0.5
    g1, g2 = stored.vjp(C, G), again.vjp(C, G)
    assert torch.equal(g1, g2)
    stored.release()
    d, fd = _directional_check(again, C, seed=11)
    _say(f"  vjp . v = {d:.8e}, central difference {fd:.8e}, rel {abs(d - fd) / abs(fd):.2e}")
    assert abs(d - fd) <= 1e-3 * abs(fd)
    # attenuation and projections of two slices against the oracle at one angle
    Cn = C.cpu().numpy()
    cx, cf = W.coefs()
    At = again.attenuation(C)
    b0, b1 = 127, 129
    ref = oracle.total_atten_cone_f64(Cn, angles[1], cx, cf, W.cone(), b0, b1)
    got = At[:, 1, b0:b1].cpu().numpy()
    e_at = rel_l2(got, ref)
    th = _theta(angles[1:2])
    em = np.full((1, 1), workloads.PIX, np.float32)
    e_pf = max(rel_l2(P2[e, 1, b0:b1].cpu().numpy(),
                      oracle.forward_batch_f32(got[e][None], np.ascontiguousarray(Cn[e, b0:b1])[None], em, th)[0])
               for e in range(W.K))
    _say(f"  slices {b0}-{b1 - 1}: attenuation rel-L2 vs FP64 {e_at:.2e}, projections vs FP32 oracle {e_pf:.2e}")
    assert e_at <= RTOL and e_pf <= RTOL


# ---------------------------------------------------------------------------
# batch_48: 48 uneven samples under the hipGraph
# ---------------------------------------------------------------------------
def test_batch_48_samples_against_oracle():
    from fl_correction_amd import workloads
    from fl_correction_amd.batch import BatchMLEM, Problem
    from fl_correction_amd.cuda_lib import atten_cuda
    samples = workloads.batch_samples()
    probs, host = [], []
    for s in samples:
        model = s.model()
        C_true = s.phantom_device()
        At = model.attenuation(C_true)
        for e in range(s.K):
            em = model.em_T[e].contiguous()
            Pf = atten_cuda.forward_emission_batch(At[e], C_true[e:e + 1], em[:, None].contiguous(), model.theta)
            I = _dev(workloads.noisy(Pf.cpu().numpy(), seed=s.seed + e))
            b = s.shape[0] // 2
            host.append((At[e][:, b:b + 1].cpu().numpy(), em.cpu().numpy(), model.theta.cpu().numpy(),
                         I[:, b:b + 1].cpu().numpy(), b))
            probs.append(Problem(At[e], em, model.theta, I, torch.ones(s.shape, device="cuda")))
        del At
    n_vox = sum(p.C.numel() for p in probs)
    _say(f"batch_48: {len(samples)} samples, {len(probs)} problems, {n_vox / 1e6:.0f} M voxels")
    run = BatchMLEM(probs, beta=0.0)
    run.run(20, graph=True)
    torch.cuda.synchronize()
    worst, floor = 0.0, []
    for i, (p, (a, em, th, I, b)) in enumerate(zip(probs, host)):
        got = p.C[b].cpu().numpy()
        C0 = np.ones((1, 1) + got.shape, np.float32)
        ref = oracle.mlem_batch_f32(C0, a, em[:, None], th, I, 20, 0.0, 0.01)[0, 0]
        err = rel_l2(got, ref)
        worst = max(worst, err)
        if err > RTOL:
            # 20 iterations on noisy data, mostly zero background: here two FP32 summation
            # orders can differ by ~1e-5.  Then the result must be as close to the FP64
            # oracle as the FP32 oracle's own serial order is (within 2x).
            ref64 = oracle.mlem_batch_f64(C0, a, em[:, None], th, I, 20, 0.0, 0.01)[0, 0]
            e_hip, e_ref = rel_l2(got, ref64), rel_l2(ref, ref64)
            floor.append((i, err, e_hip, e_ref))
            _say(f"  problem {i}: rel-L2 vs FP32 oracle {err:.2e}; vs FP64: this {e_hip:.2e}, FP32 oracle {e_ref:.2e}")
            assert e_ref >= 0.5 * RTOL and e_hip <= 2 * e_ref, (i, err, e_hip, e_ref)
    _say(f"  one slice of each of {len(probs)} problems after 20 iterations: worst rel-L2 vs FP32 oracle {worst:.2e}; "
         f"{len(floor)} above 1e-5, each within 2x of the FP32 oracle's own distance to FP64")
    assert len(floor) <= len(probs) // 20
    run.close()


# ---------------------------------------------------------------------------
# joint_512: two angles at full size
# ---------------------------------------------------------------------------
def test_joint_512_two_angles():
    from fl_correction_amd import workloads
    from fl_correction_amd.model import XRFModel
    W = workloads.JOINT_512
    angles = W.angles()[[0, 275]]
    _say(f"{W.name}: {W.shape} x {W.K} lines, cone {W.cone_n}x{W.cone_n} over {W.cone_half} deg, angles "
         f"{angles.tolist()}")
    C = W.phantom_device() * 0.5 + 0.05
    model = XRFModel(W.param(), angles, W.cone(), store=False)
    At = model.attenuation(C)
    assert float(At.min()) > 0 and float(At.max()) <= 1.0
    b, r0, r1 = 64, 192, 320
    Cn = C.cpu().numpy()
    cx, cf = W.coefs()
    ref = oracle.total_atten_cone_f64(Cn, angles[1], cx, cf, W.cone(), b, b + 1, r0, r1)
    got = At[:, 1, b:b + 1, r0:r1].cpu().numpy()
    e_at = rel_l2(got, ref)
    Pf = model.forward(C)
    th = _theta(angles[1:2])
    em = np.full((1, 1), workloads.PIX, np.float32)
    full = At[:, 1, b:b + 1].cpu().numpy()
    e_pf = max(rel_l2(Pf[e, 1, b:b + 1].cpu().numpy(),
                      oracle.forward_batch_f32(full[e][None], np.ascontiguousarray(Cn[e, b:b + 1])[None], em, th)[0])
               for e in range(W.K))
    _say(f"  slice {b}, rows {r0}-{r1 - 1}: attenuation rel-L2 vs FP64 {e_at:.2e}; projections vs FP32 oracle "
         f"{e_pf:.2e}")
    assert e_at <= RTOL and e_pf <= RTOL
    del At
    d, fd = _directional_check(model, C, seed=13)
    _say(f"  vjp . v = {d:.8e}, central difference {fd:.8e}, rel {abs(d - fd) / abs(fd):.2e}")
    assert abs(d - fd) <= 1e-3 * abs(fd)
