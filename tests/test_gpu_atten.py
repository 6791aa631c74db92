"""GPU parity of the attenuation kernels against the oracle.

Bar: relative L2 <= 1e-5 against the FP32 same-order restatement (or, for the
rotation, against the reference's own torch calls run on the CPU); the error
against the FP64 statement is printed.
"""
import numpy as np
import pytest
import torch

from helpers import RTOL, oracle, rel_l2

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


@pytest.mark.parametrize("shape,mask_args", [((2, 9, 24, 20), (20, 30, 24)), ((1, 16, 32, 40), (14, 14, 32)),
                                             ((3, 5, 17, 70), (10, 40, 17))])
def test_cal_atten_pyramid_mask(shape, mask_args):
    from fl_correction_amd.cuda_lib import atten_cuda
    from fl_correction_amd.detector_mask import generate_detector_mask
    rng = np.random.default_rng(1)
    mu = rng.uniform(0, 0.05, shape).astype(np.float32)
    mask = generate_detector_mask(*mask_args).astype(np.float32)
    assert np.allclose(mask, oracle.generate_detector_mask(*mask_args))  # product vs oracle generator
    got = atten_cuda.cal_atten_cuda(_dev(mu), _dev(mask)).cpu().numpy()
    ref = oracle.atten_f32(mu, mask)
    ref64 = oracle.atten_f64(mu, mask)
    print(f"cal_atten {shape} mask {mask.shape}: vs f32 {rel_l2(got, ref):.2e} vs f64 {rel_l2(got, ref64):.2e}")
    assert rel_l2(got, ref) <= RTOL


def test_cal_atten_random_mask_wide_rows():
    from fl_correction_amd.cuda_lib import atten_cuda
    rng = np.random.default_rng(2)
    mu = rng.uniform(0, 1, (1, 3, 9, 1100)).astype(np.float32)   # Y > 1024 (the reference's limit)
    mask = rng.uniform(0, 1, (3, 4, 5)).astype(np.float32)
    mask[1, 2, :] = 0
    got = atten_cuda.cal_atten_cuda(_dev(mu), _dev(mask)).cpu().numpy()
    assert rel_l2(got, oracle.atten_f32(mu, mask)) <= RTOL


@pytest.mark.parametrize("shape", [(3, 32, 32), (2, 40, 24), (1, 64, 64)])
def test_rotation(shape):
    from fl_correction_amd.torch_image_util import torch_rotate_image_nd
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, shape).astype(np.float32)
    for th in (0.0, 0.4, np.pi / 2, 2.2, np.float32(np.deg2rad(137.5))):
        got = torch_rotate_image_nd(_dev(x), float(th)).cpu().numpy()
        ref = oracle.rotate_torch(x, th)
        ref64 = oracle.rotate_f64(x, th)
        assert rel_l2(got, ref) <= RTOL, (th, rel_l2(got, ref))
        assert np.abs(got - ref).max() < 1e-4
        assert rel_l2(got, ref64) <= RTOL


def test_incident():
    from fl_correction_amd.torch_functions import incident_atten
    rng = np.random.default_rng(4)
    frac = rng.uniform(0, 1, (3, 4, 50, 33)).astype(np.float32)
    coef = np.array([0.01, 0.03, 0.002], np.float32)
    got = incident_atten(_dev(frac), coef).cpu().numpy()
    mux = oracle.mu_xray(frac, coef)
    assert rel_l2(got, oracle.incident_f32(mux)) <= RTOL
    assert rel_l2(got, oracle.incident_f64(mux.astype(np.float64))) <= RTOL


@pytest.mark.parametrize("side", ["r", "l"])
def test_xrf_atten_mask(side):
    from fl_correction_amd.atten_util import make_param
    from fl_correction_amd.detector_mask import generate_detector_mask
    from fl_correction_amd.torch_functions import torch_xrf_atten_3D
    rng = np.random.default_rng(5)
    K, B, H, W = 3, 6, 28, 20
    frac = rng.uniform(0, 1, (K, B, H, W)).astype(np.float32)
    cs_line = rng.uniform(20, 400, (K, K))
    param = make_param(["Ni", "Co", "Mn"], ["LiNiO2", "LiCoO2", "LiMnO2"], rng.uniform(20, 400, K), cs_line,
                       rho=1.0, pix=1e-4)
    mask = generate_detector_mask(20, 24, W).astype(np.float32)
    got = torch_xrf_atten_3D(_dev(frac), param, mask, position_det=side)
    coef = (cs_line * 1e-4).astype(np.float32)
    mu = oracle.mu_xrf(frac, coef)
    if side == "r":
        ref = oracle.xrf_atten_from_mask(mu, mask)
    else:  # mirror: rot90 c-clock in, clock out
        mu_r = np.ascontiguousarray(oracle.rot90_cclock(mu))
        ref = np.exp(-np.ascontiguousarray(oracle.rot90_clock(oracle.atten_f32(mu_r, mask)))).astype(np.float32)
    print(f"xrf mask side {side}: {rel_l2(got, ref):.2e}")
    assert got.shape == (K, B, H, W)
    assert rel_l2(got, ref) <= RTOL


def test_cone_equals_splat_mask_and_march():
    from fl_correction_amd import _lib
    from fl_correction_amd.cuda_lib import atten_cuda
    from fl_correction_amd.detector_mask import Cone
    rng = np.random.default_rng(6)
    K, Z, X, Y = 2, 11, 30, 26
    mu = rng.uniform(0, 0.05, (K, Z, X, Y)).astype(np.float32)
    cone = Cone(n=3, half_angle_deg=15.0, q=8)
    mu_d = _dev(mu)
    T = torch.empty_like(mu_d)
    args = [np.ascontiguousarray(a) for a in (cone.pz, cone.py, cone.qd, cone.weight)]
    _lib.call("fl_cone_atten", _lib.ptr(mu_d), _lib.ptr(T), K, Z, X, Y,
              *[a.ctypes.data_as(_lib.ctypes.c_void_p) for a in args], cone.ndir, _lib.stream_of(mu_d.device))
    got = T.cpu().numpy()
    ref64 = oracle.cone_march_f64(mu, cone.pz, cone.py, cone.qd, cone.weight.astype(np.float64))
    M = cone.splat_mask(X)
    via_mask = atten_cuda.cal_atten_cuda(mu_d, _dev(M)).cpu().numpy()
    print(f"cone: vs f64 march {rel_l2(got, ref64):.2e}, mask path vs f64 {rel_l2(via_mask, ref64):.2e}")
    assert rel_l2(got, ref64) <= RTOL
    assert rel_l2(via_mask, oracle.atten_f32(mu, M)) <= RTOL


@pytest.mark.parametrize("n,half,q,shape", [
    (5, 15.0, 8, (2, 20, 40, 150)),    # q = 8, 4, 1 stencils; y wider than one wave (unchecked path)
    (3, 40.0, 4, (1, 13, 33, 200)),    # slopes near 1: chains leave through the z and y faces
    (3, 20.0, 16, (1, 9, 50, 130)),    # 64-tap stencils
    (5, 15.0, 8, (3, 37, 24, 131)),    # three lines, odd sizes: partial y blocks and z groups
])
def test_cone_wide_volumes(n, half, q, shape):
    from fl_correction_amd import _lib
    from fl_correction_amd.detector_mask import Cone
    rng = np.random.default_rng(16)
    mu = rng.uniform(0, 0.05, shape).astype(np.float32)
    K, Z, X, Y = shape
    cone = Cone(n=n, half_angle_deg=half, q=q)
    args = [np.ascontiguousarray(a) for a in (cone.pz, cone.py, cone.qd, cone.weight)]
    ptrs = [a.ctypes.data_as(_lib.ctypes.c_void_p) for a in args]
    mu_d = _dev(mu)
    T = torch.empty_like(mu_d)
    print(f"cone n={n} half={half} q={q} shape={shape}", flush=True)
    _lib.call("fl_cone_atten", _lib.ptr(mu_d), _lib.ptr(T), K, Z, X, Y, *ptrs, cone.ndir, _lib.stream_of(mu_d.device))
    got = T.cpu().numpy()
    ref64 = oracle.cone_march_f64(mu, cone.pz, cone.py, cone.qd, cone.weight.astype(np.float64))
    print(f"  vs f64 march {rel_l2(got, ref64):.2e}")
    assert rel_l2(got, ref64) <= RTOL
    # the transpose: <K x, y> = <x, K^T y>
    y = _dev(rng.uniform(-1, 1, shape).astype(np.float32))
    Kty = torch.empty_like(y)
    _lib.call("fl_cone_atten_adjoint", _lib.ptr(y), _lib.ptr(Kty), K, Z, X, Y, *ptrs, cone.ndir,
              _lib.stream_of(mu_d.device))
    lhs = float((T.double() * y.double()).sum())
    rhs = float((mu_d.double() * Kty.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_total_atten_pipeline_against_reference_composition():
    from fl_correction_amd import core
    from fl_correction_amd.atten_util import make_param
    from fl_correction_amd.detector_mask import generate_detector_mask
    rng = np.random.default_rng(7)
    K, B, N = 2, 5, 24
    frac = rng.uniform(0, 1, (K, B, N, N)).astype(np.float32)
    cs_x, cs_line = rng.uniform(20, 400, K), rng.uniform(20, 400, (K, K))
    param = make_param(["Fe", "Cu"], ["Fe2O3", "CuO"], cs_x, cs_line, rho=1.0, pix=1e-4)
    mask = generate_detector_mask(20, 20, N).astype(np.float32)
    angles = np.array([0.0, 33.0, 90.0, 150.0])
    got = core.cal_atten_at_angles(_dev(frac), param, angles, mask).cpu().numpy()
    xrf = core.cal_xrf_atten_at_angles(_dev(frac), param, angles, mask).cpu().numpy()
    xray = core.cal_xray_atten_at_angles(_dev(frac), param, angles).cpu().numpy()
    for a, ang in enumerate(angles):
        th = np.float32(ang / 180 * np.pi)
        rot = oracle.rotate_torch(frac, th)
        a_x = oracle.incident_f32(oracle.mu_xray(rot, (cs_x * 1e-4).astype(np.float32)))
        a_f = oracle.xrf_atten_from_mask(oracle.mu_xrf(rot, (cs_line * 1e-4).astype(np.float32)), mask)
        assert rel_l2(xray[:, a], a_x) <= RTOL
        assert rel_l2(xrf[:, a], a_f) <= RTOL
        assert rel_l2(got[:, a], a_f * a_x) <= RTOL
