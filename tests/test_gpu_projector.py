"""GPU parity of the emission projector, its transpose and MLEM against the oracle.

Bar: relative L2 <= 1e-5 against the FP32 same-order restatement of the reference
kernels (oracle/fl_oracle.c); the error against the FP64 oracle is printed.
"""
import numpy as np
import pytest
import torch

from helpers import RTOL, oracle, projector_inputs, rel_l2

pytestmark = pytest.mark.gpu

SPECIAL = [0.0, np.pi / 2, np.pi / 4, np.pi, 3 * np.pi / 4, 1e-3]

SHAPES = [
    # A, B, R, H, W
    (6, 3, 1, 32, 32),
    (7, 5, 1, 64, 64),
    (5, 2, 2, 40, 70),     # ragged, H != W, two references
    (4, 1, 3, 97, 33),     # R larger than one reference chunk
    (9, 6, 1, 128, 128),
    (3, 4, 1, 130, 66),    # sizes just past tile multiples
]


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.fixture(scope="module")
def ops():
    from fl_correction_amd.cuda_lib import atten_cuda
    return atten_cuda


@pytest.mark.parametrize("shape", SHAPES)
def test_forward_batch(ops, shape):
    A, B, R, H, W = shape
    th = SPECIAL[:A] + list(np.random.default_rng(1).uniform(0, np.pi, max(0, A - len(SPECIAL))))
    atten, C, em, theta, _ = projector_inputs(A, B, R, H, W, seed=11, thetas=th[:A])
    got = ops.forward_emission_batch(_dev(atten), _dev(C), _dev(em), _dev(theta)).cpu().numpy()
    ref = oracle.forward_batch_f32(atten, C, em, theta)
    ref64 = oracle.forward_batch_f64(atten, C, em, theta)
    e32, e64 = rel_l2(got, ref), rel_l2(got, ref64)
    print(f"forward {shape}: rel-L2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}, oracle-f32 vs f64 {rel_l2(ref, ref64):.2e}")
    assert got.shape == (A, B, W)
    assert e32 <= RTOL


@pytest.mark.parametrize("shape", SHAPES)
def test_backward_batch(ops, shape):
    A, B, R, H, W = shape
    th = SPECIAL[:A] + list(np.random.default_rng(2).uniform(0, np.pi, max(0, A - len(SPECIAL))))
    atten, C, em, theta, ratio = projector_inputs(A, B, R, H, W, seed=12, thetas=th[:A])
    got = ops.backward_emission_batch(_dev(atten), _dev(ratio), _dev(em), _dev(theta)).cpu().numpy()
    ref = oracle.backward_batch_f32(atten, ratio, em, theta)
    ref64 = oracle.backward_batch_f64(atten, ratio, em, theta)
    e32, e64 = rel_l2(got, ref), rel_l2(got, ref64)
    print(f"backward {shape}: rel-L2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}")
    assert got.shape == (R, B, H, W)
    assert e32 <= RTOL


def test_single_slice_forms(ops):
    A, R, H, W = 8, 2, 48, 48
    atten, C, em, theta, ratio = projector_inputs(A, 1, R, H, W, seed=5)
    atten, C, ratio = atten[:, 0], C[:, 0], ratio[:, 0]
    got = ops.forward_emission(_dev(atten), _dev(C), _dev(em), _dev(theta)).cpu().numpy()
    assert rel_l2(got, oracle.forward_single_f32(atten, C, em, theta)) <= RTOL
    gb = ops.backward_emission(_dev(atten), _dev(ratio), _dev(em), _dev(theta)).cpu().numpy()
    assert rel_l2(gb, oracle.backward_single_f32(atten, ratio, em, theta)) <= RTOL


def test_deterministic(ops):
    A, B, R, H, W = 12, 4, 1, 96, 96
    atten, C, em, theta, ratio = projector_inputs(A, B, R, H, W, seed=7)
    a, c, e, t, r = (_dev(x) for x in (atten, C, em, theta, ratio))
    f1 = ops.forward_emission_batch(a, c, e, t)
    f2 = ops.forward_emission_batch(a, c, e, t)
    b1 = ops.backward_emission_batch(a, r, e, t)
    b2 = ops.backward_emission_batch(a, r, e, t)
    assert torch.equal(f1, f2)
    assert torch.equal(b1, b2)


def test_adjoint_dot_product(ops):
    A, B, R, H, W = 10, 3, 2, 64, 80
    atten, C, em, theta, ratio = projector_inputs(A, B, R, H, W, seed=9)
    Pf = ops.forward_emission_batch(_dev(atten), _dev(C), _dev(em), _dev(theta)).cpu().double().numpy()
    BtR = ops.backward_emission_batch(_dev(atten), _dev(ratio), _dev(em), _dev(theta)).cpu().double().numpy()
    lhs = float((Pf * ratio).sum())
    rhs = float((C * BtR).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


@pytest.mark.parametrize("shape", [(16, 2, 1, 256, 256), (12, 2, 3, 200, 150), (20, 2, 1, 24, 160), (20, 2, 2, 160, 24)])
def test_spread_angles_wide_images(ops, shape):
    """Angles in random order over [0, 2 pi) on images wider than two 64-column
    strips: the forward merges the footprints of a group's angles, and a merge
    of far-apart angles must not outgrow the staged width.  The elongated
    shapes put many clamped samples into the image corners (the back-projector's
    border gather walks strips that reach into the corner regions).  Results
    must match the oracle and repeat bit for bit."""
    A, B, R, H, W = shape
    th = np.random.default_rng(31).uniform(0, 2 * np.pi, A)
    atten, C, em, theta, ratio = projector_inputs(A, B, R, H, W, seed=13, thetas=th)
    print(f"spread angles {shape}", flush=True)
    a, c, e, t, r = (_dev(x) for x in (atten, C, em, theta, ratio))
    f1 = ops.forward_emission_batch(a, c, e, t)
    f2 = ops.forward_emission_batch(a, c, e, t)
    assert torch.equal(f1, f2)
    assert rel_l2(f1.cpu().numpy(), oracle.forward_batch_f32(atten, C, em, theta)) <= RTOL
    b1 = ops.backward_emission_batch(a, r, e, t)
    b2 = ops.backward_emission_batch(a, r, e, t)
    assert torch.equal(b1, b2)
    assert rel_l2(b1.cpu().numpy(), oracle.backward_batch_f32(atten, ratio, em, theta)) <= RTOL


@pytest.mark.parametrize("beta", [0.0, 1e-3])
def test_mlem_batch(ops, beta):
    # reduced config-0 shape: 2 references would share a line; here one line, 20 iterations
    from fl_correction_amd import phantom
    B = H = W = 48
    A = 30
    C_true = phantom.ellipsoid_phantom((B, H, W), 1, n_ell=(6, 12), semi_axes=(4, 12), seed=20260000)
    theta = np.deg2rad(phantom.angle_list(A)).astype(np.float32)
    rng = np.random.default_rng(3)
    atten = rng.uniform(0.3, 1.0, size=(A, B, H, W)).astype(np.float32)
    em = np.full((A, 1), 1e-4 * 1.0, np.float32) * 1e4
    I = oracle.forward_batch_f32(atten, C_true, em, theta)
    I = phantom.poisson_noise(I, 1e4, seed=4)
    C0 = np.ones_like(C_true)
    got = ops.mlem_cuda_batch(_dev(C0), _dev(atten), _dev(em), _dev(theta), _dev(I), 20, beta, 0.01).cpu().numpy()
    ref = oracle.mlem_batch_f32(C0, atten, em, theta, I, 20, beta, 0.01)
    ref64 = oracle.mlem_batch_f64(C0, atten, em, theta, I, 20, beta, 0.01)
    e32, e64 = rel_l2(got, ref), rel_l2(got, ref64)
    print(f"mlem beta={beta}: rel-L2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}, oracle-f32 vs f64 {rel_l2(ref, ref64):.2e}")
    assert e32 <= RTOL


def test_mlem_single(ops):
    A, H, W = 16, 40, 40
    rng = np.random.default_rng(8)
    C_true = rng.uniform(0, 1, (1, H, W)).astype(np.float32)
    atten = rng.uniform(0.3, 1.0, (A, H, W)).astype(np.float32)
    em = np.ones((A, 1), np.float32)
    theta = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    I = oracle.forward_single_f32(atten, C_true, em, theta)
    C0 = np.ones_like(C_true)
    got = ops.mlem_cuda(_dev(C0), _dev(atten), _dev(em), _dev(theta), _dev(I), 10).cpu().numpy()
    ref = oracle.mlem_batch_f32(C0, atten, em, theta, I, 10)
    assert rel_l2(got, ref) <= RTOL


def test_errors(ops):
    x = torch.zeros(2, 2, 8, 8)
    with pytest.raises(RuntimeError):
        ops.forward_emission_batch(x, x, torch.zeros(2, 2), torch.zeros(2))  # host tensors
    a = torch.zeros(2, 1, 8, 8, device="cuda")
    with pytest.raises(RuntimeError):
        ops.forward_emission_batch(a, torch.zeros(1, 2, 8, 8, device="cuda"), torch.ones(2, 1, device="cuda"),
                                   torch.zeros(2, device="cuda"))  # batch mismatch
    with pytest.raises(RuntimeError):
        ops.forward_emission_batch(a[0], a, torch.ones(2, 1, device="cuda"), torch.zeros(2, device="cuda"))


@pytest.mark.parametrize("shape,delta", [((2, 3, 40, 70), 0.01), ((1, 1, 5, 2), 0.05), ((3, 2, 64, 64), 1e-3)])
def test_huber_grad_batch_against_oracle(shape, delta):
    """fl_huber_grad_batch (the reference's internal huber_grad_batch, atten_cuda.cpp:229-238)
    against the FP32 oracle: the same four clipped differences summed in the same order."""
    import torch
    from fl_correction_amd import _lib
    rng = np.random.default_rng(61)
    C = rng.uniform(0, 1, shape).astype(np.float32)
    C[..., ::3, ::2] = C[..., ::3, ::2] * 0.001  # steps both below and above delta
    Cd = torch.from_numpy(C).cuda()
    g = torch.full_like(Cd, 99.0)
    R, B, H, W = shape
    _lib.call("fl_huber_grad_batch", _lib.ptr(Cd), _lib.ptr(g), R, B, H, W, float(delta), _lib.stream_of(Cd.device))
    ref = oracle.huber_batch_f32(C, delta)
    assert np.array_equal(g.cpu().numpy(), ref)
