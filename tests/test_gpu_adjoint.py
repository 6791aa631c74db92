"""GPU checks of the hand-written adjoint of the attenuated forward model.

* each adjoint kernel against its forward by the dot-product identity <K x, y> = <x, K^T y>;
* the full model: HIP forward against the pure-torch FP64 model (oracle/torch_model.py), and
  the HIP vector-Jacobian product against torch autograd through that model
  (tolerance: relative L2 <= 1e-4 for an FP32 gradient against FP64);
* stored and recomputed attenuation give the same gradient.
"""
import numpy as np
import pytest
import torch

from helpers import oracle, rel_l2
import torch_model as tm

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def dot(a, b):
    return float((a.double() * b.double()).sum())


def test_rotate_adjoint_dot():
    from fl_correction_amd import _lib
    rng = np.random.default_rng(1)
    for (N, H, W) in [(3, 32, 32), (2, 20, 36)]:
        x, y = _dev(rng.uniform(-1, 1, (N, H, W))), _dev(rng.uniform(-1, 1, (N, H, W)))
        for th in (0.0, 0.3, 1.2, np.pi / 2, 2.9):
            c, s = float(np.float32(np.cos(np.float32(th)))), float(np.float32(np.sin(np.float32(th))))
            Rx, Rty = torch.empty_like(x), torch.empty_like(y)
            st = _lib.stream_of(x.device)
            _lib.call("fl_rotate", _lib.ptr(x), _lib.ptr(Rx), N, H, W, c, s, st)
            _lib.call("fl_rotate_adjoint", _lib.ptr(y), _lib.ptr(Rty), N, H, W, c, s, 0, st)
            l, r = dot(Rx, y), dot(x, Rty)
            assert abs(l - r) <= 1e-5 * max(abs(l), 1.0), (N, H, W, th, l, r)


def test_cone_and_mask_adjoint_dot():
    from fl_correction_amd import _lib
    from fl_correction_amd.detector_mask import Cone
    rng = np.random.default_rng(2)
    K, Z, X, Y = 2, 9, 21, 17
    x, y = _dev(rng.uniform(-1, 1, (K, Z, X, Y))), _dev(rng.uniform(-1, 1, (K, Z, X, Y)))
    cone = Cone(n=3, half_angle_deg=12, q=8)
    d = [np.ascontiguousarray(v) for v in (cone.pz, cone.py, cone.qd, cone.weight)]
    dp = [v.ctypes.data_as(_lib.ctypes.c_void_p) for v in d]
    Kx, Kty = torch.empty_like(x), torch.empty_like(y)
    st = _lib.stream_of(x.device)
    _lib.call("fl_cone_atten", _lib.ptr(x), _lib.ptr(Kx), K, Z, X, Y, *dp, cone.ndir, st)
    _lib.call("fl_cone_atten_adjoint", _lib.ptr(y), _lib.ptr(Kty), K, Z, X, Y, *dp, cone.ndir, st)
    assert abs(dot(Kx, y) - dot(x, Kty)) <= 1e-5 * abs(dot(Kx, y))
    mask = _dev(rng.uniform(0, 1, (5, 8, 3)))
    _lib.call("fl_cal_atten", _lib.ptr(x), _lib.ptr(mask), _lib.ptr(Kx), K, Z, X, Y, 5, 8, 3, st)
    _lib.call("fl_cal_atten_adjoint", _lib.ptr(y), _lib.ptr(mask), _lib.ptr(Kty), K, Z, X, Y, 5, 8, 3, st)
    assert abs(dot(Kx, y) - dot(x, Kty)) <= 1e-5 * abs(dot(Kx, y))


def test_sample_grad_is_projector_derivative():
    from fl_correction_amd.cuda_lib import autograd
    rng = np.random.default_rng(3)
    A, B, R, H, W = 5, 2, 2, 24, 30
    atten = _dev(rng.uniform(0.2, 1, (A, B, H, W))).requires_grad_(True)
    C = _dev(rng.uniform(0, 1, (R, B, H, W))).requires_grad_(True)
    em, theta = _dev(rng.uniform(0.5, 1.5, (A, R))), _dev(rng.uniform(0, np.pi, A))
    G = _dev(rng.uniform(-1, 1, (A, B, W)))
    Pf = autograd.forward_emission_batch_autograd(atten, C, em, theta)
    (Pf * G).sum().backward()
    # the projector is bilinear in (atten, C): <dL/datten, atten> = <dL/dC, C> = <Pf, G>
    ref = dot(Pf.detach(), G)
    assert abs(dot(atten.grad, atten.detach()) - ref) <= 1e-5 * abs(ref)
    assert abs(dot(C.grad, C.detach()) - ref) <= 1e-5 * abs(ref)


def _model_case(detector_kind, seed=4, K=2, B=5, N=24, A=5):
    from fl_correction_amd.atten_util import make_param
    from fl_correction_amd.detector_mask import Cone, generate_detector_mask
    rng = np.random.default_rng(seed)
    C = rng.uniform(0, 1, (K, B, N, N)).astype(np.float32)
    # cross-sections scaled so that the attenuation is strong (exercises dA/dC)
    cs_x, cs_line = rng.uniform(100, 400, K), rng.uniform(50, 300, (K, K))
    param = make_param([f"E{i}" for i in range(K)], [f"C{i}" for i in range(K)], cs_x, cs_line, rho=1.0, pix=1e-4)
    angles = np.array([0.0, 27.0, 90.0, 133.0, 171.0])[:A]
    if detector_kind == "cone":
        det = Cone(n=3, half_angle_deg=15, q=8)
        mask = det.splat_mask(N)
    else:
        det = mask = generate_detector_mask(20, 20, N).astype(np.float32)
    em = rng.uniform(0.5, 1.5, (A, K)).astype(np.float32)
    coef_x = np.tile((cs_x * 1e-4).astype(np.float32), (A, 1))
    coef_xrf = np.tile((cs_line * 1e-4).astype(np.float32), (A, 1, 1))
    return C, param, angles, det, mask, em, coef_x, coef_xrf


@pytest.mark.parametrize("kind", ["mask", "cone"])
@pytest.mark.parametrize("store", [True, False])
def test_model_forward_and_vjp_vs_torch_autograd(kind, store):
    from fl_correction_amd.model import XRFModel
    C, param, angles, det, mask, em, coef_x, coef_xrf = _model_case(kind)
    model = XRFModel(param, angles, det, em=em, store=store)
    Pf = model.forward(_dev(C))
    Ct = torch.tensor(C, dtype=torch.float64, requires_grad=True)
    Pref = tm.forward(Ct, angles, coef_x, coef_xrf, em, mask)
    e_f = rel_l2(Pf.cpu().numpy(), Pref.detach().numpy())
    G = np.random.default_rng(5).uniform(-1, 1, Pf.shape)
    (Pref * torch.tensor(G)).sum().backward()
    dC = model.vjp(_dev(C), _dev(G)).cpu().numpy()
    e_g = rel_l2(dC, Ct.grad.numpy())
    print(f"model {kind} store={store}: forward vs torch f64 {e_f:.2e}, vjp vs autograd f64 {e_g:.2e}")
    assert e_f <= 1e-5
    assert e_g <= 1e-4


def test_autograd_function_and_attenuation_term_matters():
    from fl_correction_amd.model import XRFModel, xrf_forward_autograd
    C, param, angles, det, mask, em, coef_x, coef_xrf = _model_case("mask", seed=6)
    model = XRFModel(param, angles, det, em=em)
    Cd = _dev(C).requires_grad_(True)
    Pf = xrf_forward_autograd(Cd, model)
    I = Pf.detach() * 1.1
    loss = (Pf - I * torch.log(Pf.clamp_min(1e-6))).mean()  # Poisson NLL (ml_loss.py:6-8)
    loss.backward()
    g_full = Cd.grad.clone()
    # gradient with the attenuation held fixed (the reference's autograd) differs
    from fl_correction_amd.cuda_lib import autograd
    At = model.attenuation(_dev(C))
    Cd2 = _dev(C).requires_grad_(True)
    Pf2 = torch.stack([autograd.forward_emission_batch_autograd(At[e], Cd2[e:e + 1], _dev(em[:, e:e + 1]), model.theta)
                       for e in range(C.shape[0])])
    assert rel_l2(Pf2.detach().cpu().numpy(), Pf.detach().cpu().numpy()) < 1e-5
    (Pf2 - I * torch.log(Pf2.clamp_min(1e-6))).mean().backward()
    assert rel_l2(Cd2.grad.cpu().numpy(), g_full.cpu().numpy()) > 1e-3


@pytest.mark.parametrize("kind", ["cone", "mask"])
def test_attenuation_streams_agree(kind):
    """The attenuation maps and the forward, dealt over 2 or 3 streams, equal the
    one-stream results bit for bit (every angle is computed the same way), also
    when the caller's stream has pending work that produced C; the vjp equals
    the one-stream vjp to rounding (its sum over angles is grouped by stream)."""
    from fl_correction_amd.model import XRFModel
    C, param, angles, det, mask, em, coef_x, coef_xrf = _model_case(kind, seed=21)
    model = XRFModel(param, angles, det, em=em)
    Cd = _dev(C)
    one = model.attenuation(Cd, streams=1)
    Pf1 = model.forward(Cd, angle_chunk=2, streams=1)
    for n in (2, 3):
        C2 = (Cd * 2.0) * 0.5  # produced on the current stream just before the call
        got = model.attenuation(C2, streams=n)
        assert torch.equal(got, one), n
        # the forward's chunks of 2 angles over 5 angles: a partial last chunk
        assert torch.equal(model.forward(C2, angle_chunk=2, streams=n), Pf1), n
    # vjp: the sum over angles is grouped by stream, so equal to rounding
    G = torch.randn_like(Pf1)
    model.forward(Cd, angle_chunk=2, streams=1)
    g1 = model.vjp(Cd, G, angle_chunk=2, streams=1).cpu().numpy()
    for n in (2, 3):
        gn = model.vjp(Cd, G, angle_chunk=2, streams=n).cpu().numpy()
        assert rel_l2(gn, g1) <= 1e-6, n


def test_stored_exponents_over_streams():
    """forward with stored exponents over 2 streams, then vjp over 2 streams, against the
    one-stream forward and vjp; and the autograd function with streams=1."""
    from fl_correction_amd.model import XRFModel, xrf_forward_autograd
    C, param, angles, det, mask, em, coef_x, coef_xrf = _model_case("cone", seed=23)
    one = XRFModel(param, angles, det, em=em, store=True)
    two = XRFModel(param, angles, det, em=em, store=True)
    Cd = _dev(C)
    P1 = one.forward(Cd, angle_chunk=2, streams=1)
    P2 = two.forward(Cd, angle_chunk=2, streams=2)
    assert two._T_store is not None and torch.equal(P1, P2)
    G = torch.randn_like(P1)
    g1 = one.vjp(Cd, G, angle_chunk=2, streams=1)
    g2 = two.vjp(Cd, G, angle_chunk=2, streams=2)
    assert rel_l2(g2.cpu().numpy(), g1.cpu().numpy()) <= 1e-6
    Cg = Cd.clone().requires_grad_(True)
    Pf = xrf_forward_autograd(Cg, one, streams=1)
    assert torch.equal(Pf.detach(), P1)
    (Pf * G).sum().backward()
    assert torch.equal(Cg.grad, one.vjp(Cd, G, streams=1))


@pytest.mark.parametrize("kind", ["cone", "mask"])
@pytest.mark.parametrize("streams", [1, 2])
def test_loss_and_grad_matches_forward_then_vjp(kind, streams):
    """The one-pass loss, gradient and projections (chunks of 2 of 5 angles, the
    last one partial) against the Poisson loss of the forward and the gradient
    through the autograd function (forward, then vjp with the exponents
    recomputed), and the loss and gradient against torch autograd through the
    FP64 torch model.  Some projections lie below eps, so the clamp's zero
    derivative is exercised."""
    from fl_correction_amd.joint import poisson_loss
    from fl_correction_amd.model import XRFModel, xrf_forward_autograd
    C, param, angles, det, mask, em, coef_x, coef_xrf = _model_case(kind, seed=31)
    model = XRFModel(param, angles, det, em=em, store=False)
    Cd = _dev(C)
    Pf0 = model.forward(Cd)
    # 10% of the projections clamped, eps half-way between two projections so that
    # the FP64 model clamps the same ones
    srt = Pf0.flatten().sort().values
    k = srt.numel() // 10
    eps = 0.5 * float(srt[k - 1] + srt[k])
    I = Pf0 * torch.from_numpy(np.random.default_rng(32).uniform(0.7, 1.3, Pf0.shape)).float().cuda()
    loss, g, Pf = model.loss_and_grad(Cd, I, eps=eps, angle_chunk=2, streams=streams)
    assert torch.equal(Pf, Pf0)
    Cg = Cd.clone().requires_grad_(True)
    ref_loss = poisson_loss(xrf_forward_autograd(Cg, model, streams=streams), I, eps)
    ref_loss.backward()
    e_l = abs(float(loss) - float(ref_loss)) / abs(float(ref_loss))
    e_g = rel_l2(g.cpu().numpy(), Cg.grad.cpu().numpy())
    # FP64 torch model and autograd
    Ct = torch.tensor(C, dtype=torch.float64, requires_grad=True)
    P64 = tm.forward(Ct, angles, coef_x, coef_xrf, em, mask).clamp_min(eps)
    I64 = I.double().cpu()
    L64 = (P64 - I64 * torch.log(P64)).mean()
    L64.backward()
    e_l64 = abs(float(loss) - float(L64)) / abs(float(L64))
    e_g64 = rel_l2(g.cpu().numpy(), Ct.grad.numpy())
    print(f"loss_and_grad {kind} streams={streams}: vs forward+vjp loss {e_l:.1e} grad {e_g:.1e}; "
          f"vs torch f64 loss {e_l64:.1e} grad {e_g64:.1e}")
    assert e_l <= 1e-6
    assert e_g <= 1e-5
    assert e_l64 <= 1e-5
    assert e_g64 <= 1e-4
