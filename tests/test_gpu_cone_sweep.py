"""fl_cone_sweep: the cone march for all directions in one sweep over blocks of Q
planes, against the FP64 march (oracle.cone_march_f64), the per-direction chain
form (fl_cone_atten / fl_cone_atten_adjoint) and the adjoint identity.

Bars: relative L2 <= 1e-5 against the FP64 march and <= 2e-6 against the chain
form (the two split each ray sum differently); <K x, y> = <x, K^T y> to 1e-5.
"""
import numpy as np
import pytest
import torch

from helpers import RTOL, oracle, rel_l2

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _run(name, x, cone, K, Z, X, Y, adjoint=None):
    from fl_correction_amd import _lib
    args = [np.ascontiguousarray(a) for a in (cone.pz, cone.py, cone.qd, cone.weight)]
    ptrs = [a.ctypes.data_as(_lib.ctypes.c_void_p) for a in args]
    out = torch.empty_like(x)
    st = _lib.stream_of(x.device)
    if name == "sweep":
        nb = _lib.load().fl_cone_sweep_workspace_bytes(K, Z, X, Y, *ptrs[:3], cone.ndir)
        assert nb > 0
        ws = torch.empty((nb + 3) // 4, dtype=torch.float32, device=x.device)
        _lib.call("fl_cone_sweep", _lib.ptr(x), _lib.ptr(out), K, Z, X, Y, *ptrs, cone.ndir, int(adjoint),
                  _lib.ptr(ws), ws.numel() * 4, st)
    else:
        fn = "fl_cone_atten_adjoint" if adjoint else "fl_cone_atten"
        _lib.call(fn, _lib.ptr(x), _lib.ptr(out), K, Z, X, Y, *ptrs, cone.ndir, st)
    return out


CASES = [
    (3, 10.0, 8, (2, 11, 30, 26)),     # one partial x block, y within one wave
    (5, 15.0, 8, (2, 20, 40, 150)),    # q = 8, 4, 1 in reduced form; y over three waves
    (3, 40.0, 4, (1, 13, 33, 200)),    # slopes near 1: rays leave through the z and y faces
    (3, 20.0, 16, (1, 9, 50, 130)),    # 64-tap stencils, Q = 16
    (5, 15.0, 8, (3, 37, 24, 131)),    # three lines, odd sizes
    (7, 20.0, 8, (2, 16, 64, 70)),     # 49 directions
    (1, 10.0, 8, (1, 5, 17, 9)),       # one straight ray
]


@pytest.mark.parametrize("n,half,q,shape", CASES)
def test_sweep_equals_march_and_chain(n, half, q, shape):
    from fl_correction_amd.detector_mask import Cone
    rng = np.random.default_rng(21 + n)
    K, Z, X, Y = shape
    mu = rng.uniform(0, 0.05, shape).astype(np.float32)
    cone = Cone(n=n, half_angle_deg=half, q=q)
    print(f"sweep n={n} half={half} q={q} shape={shape}", flush=True)
    mu_d = _dev(mu)
    got = _run("sweep", mu_d, cone, K, Z, X, Y, adjoint=False)
    chain = _run("chain", mu_d, cone, K, Z, X, Y, adjoint=False)
    ref64 = oracle.cone_march_f64(mu, cone.pz, cone.py, cone.qd, cone.weight.astype(np.float64))
    g = got.cpu().numpy()
    print(f"  vs f64 march {rel_l2(g, ref64):.2e}, vs chain {rel_l2(g, chain.cpu().numpy()):.2e}")
    assert rel_l2(g, ref64) <= RTOL
    assert rel_l2(g, chain.cpu().numpy()) <= 2e-6
    # the adjoint sweep against the chain adjoint, and the identity <K x, y> = <x, K^T y>
    y = _dev(rng.uniform(-1, 1, shape).astype(np.float32))
    kty = _run("sweep", y, cone, K, Z, X, Y, adjoint=True)
    kty_chain = _run("chain", y, cone, K, Z, X, Y, adjoint=True)
    assert rel_l2(kty.cpu().numpy(), kty_chain.cpu().numpy()) <= 2e-6
    lhs = float((got.double() * y.double()).sum())
    rhs = float((mu_d.double() * kty.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_sweep_is_deterministic_and_checks_arguments():
    from fl_correction_amd import _lib
    from fl_correction_amd.detector_mask import Cone
    K, Z, X, Y = 2, 12, 40, 70
    cone = Cone(n=5, half_angle_deg=15.0, q=8)
    x = torch.rand((K, Z, X, Y), device="cuda") * 0.05
    a = _run("sweep", x, cone, K, Z, X, Y, adjoint=False)
    b = _run("sweep", x, cone, K, Z, X, Y, adjoint=False)
    assert torch.equal(a, b)
    # denominators that divide no common Q in {4, 8, 16}: unsupported, no workspace size
    args = [np.ascontiguousarray(v) for v in (np.array([1, 0], np.int32), np.array([0, 1], np.int32),
                                              np.array([3, 3], np.int32))]
    ptrs = [v.ctypes.data_as(_lib.ctypes.c_void_p) for v in args]
    assert _lib.load().fl_cone_sweep_workspace_bytes(K, Z, X, Y, *ptrs, 2) == 0
    # a workspace one byte short is refused
    args = [np.ascontiguousarray(v) for v in (cone.pz, cone.py, cone.qd, cone.weight)]
    ptrs = [v.ctypes.data_as(_lib.ctypes.c_void_p) for v in args]
    nb = _lib.load().fl_cone_sweep_workspace_bytes(K, Z, X, Y, *ptrs[:3], cone.ndir)
    ws = torch.empty((nb + 3) // 4, dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError):
        _lib.call("fl_cone_sweep", _lib.ptr(x), _lib.ptr(a), K, Z, X, Y, *ptrs, cone.ndir, 0, _lib.ptr(ws), nb - 1,
                  _lib.stream_of(x.device))
