"""Pins the CPU oracle against independent statements of the same mathematics.

The reference ships no tests, golden vectors or fixtures, so each oracle
function is checked against something that fixes its answer independently:
  * projector / transpose: a dense system matrix built the way the reference's
    own legacy CPU path builds it (FL/legacy/FL_correction_core.py:1198-1245,
    2203-2279), the theta = 0 closed form (including the -1.001 clamp of the
    last row/column), and the adjoint dot-product identity;
  * MLEM: textbook MLEM / One-Step-Late iterations on that dense matrix, and
    the fixed-point property of the noiseless solution;
  * Huber gradient: a vectorised restatement with array shifts;
  * mask correlation: torch.nn.functional.conv3d;
  * incident product: torch.cumprod and the uniform-mu closed form;
  * rotation: torch's affine_grid + grid_sample (the reference's own calls)
    against the closed-form FP64 statement;
  * cone: the ray-splat mask correlated by the mask kernel against the direct march.
"""
import os
import sys

import numpy as np
import pytest
import torch

from helpers import oracle, projector_inputs, rel_l2


# ---------------------------------------------------------------------------
# dense system matrix (legacy CPU formulation)
# ---------------------------------------------------------------------------


def dense_H(atten2d_list, thetas, em):
    """H (A*W, H*W) for one slice: row (a, q) sums, over detector rows p, atten[a,p,q]
    times the bilinear weights of the sample point, as generate_H_from_atten_coef
    (FL_correction_core.py:1198-1245) writes them (samples outside [0, n-1] skipped)."""
    A = len(thetas)
    _, Hh, Ww = atten2d_list.shape
    cx = (Ww - 1) / 2.0
    cy = (Hh - 1) / 2.0
    M = np.zeros((A * Ww, Hh * Ww))
    for i, th in enumerate(thetas):
        T00, T01 = np.cos(-th), -np.sin(-th)
        T10, T11 = np.sin(-th), np.cos(-th)
        for q in range(Ww):
            for p in range(Hh):
                r = T00 * (p - cy) + T01 * (q - cx) + cy
                c = T10 * (p - cy) + T11 * (q - cx) + cx
                if r > Hh - 1 or r < 0 or c > Ww - 1 or c < 0:
                    continue
                r0, c0 = int(np.floor(r)), int(np.floor(c))
                r1, c1 = int(np.ceil(r)), int(np.ceil(c))
                fr, fc = r - r0, c - c0
                w_ = atten2d_list[i, p, q] * em[i]
                M[i * Ww + q, r0 * Ww + c0] += w_ * (1 - fr) * (1 - fc)
                M[i * Ww + q, r1 * Ww + c0] += w_ * fr * (1 - fc)
                M[i * Ww + q, r0 * Ww + c1] += w_ * (1 - fr) * fc
                M[i * Ww + q, r1 * Ww + c1] += w_ * fr * fc
    return M


def zero_border(C, k=2):
    C = C.copy()
    C[..., :k, :] = 0
    C[..., -k:, :] = 0
    C[..., :, :k] = 0
    C[..., :, -k:] = 0
    return C


@pytest.fixture(scope="module")
def small():
    A, B, R, H, W = 5, 2, 1, 12, 12
    atten, C, em, theta, ratio = projector_inputs(A, B, R, H, W, seed=3)
    theta = theta.astype(np.float32)
    return A, B, R, H, W, atten, zero_border(C), em, theta, ratio


def test_forward_matches_dense_matrix(small):
    A, B, R, H, W, atten, C, em, theta, _ = small
    th64 = theta.astype(np.float64)
    ref = oracle.forward_batch_f64(atten, C, em, theta)
    for b in range(B):
        M = dense_H(atten[:, b].astype(np.float64), th64, em[:, 0].astype(np.float64))
        Pf = (M @ C[0, b].astype(np.float64).ravel()).reshape(A, W)
        assert np.allclose(Pf, ref[:, b], rtol=1e-12, atol=1e-12)
    f32 = oracle.forward_batch_f32(atten, C, em, theta)
    assert rel_l2(f32, ref) < 1e-6


def test_backward_matches_dense_transpose(small):
    A, B, R, H, W, atten, _, em, theta, ratio = small
    ref = oracle.backward_batch_f64(atten, ratio, em, theta)
    for b in range(B):
        M = dense_H(atten[:, b].astype(np.float64), theta.astype(np.float64), em[:, 0].astype(np.float64))
        bt = (M.T @ ratio[:, b].astype(np.float64).ravel()).reshape(H, W)
        # the dense (skip) and kernel (clamp) boundaries differ only on the outer ring
        assert np.allclose(bt[2:-2, 2:-2], ref[0, b, 2:-2, 2:-2], rtol=1e-10, atol=1e-10)
    assert rel_l2(oracle.backward_batch_f32(atten, ratio, em, theta), ref) < 1e-6


def test_theta_zero_closed_form():
    A, B, R, H, W = 1, 2, 2, 9, 7
    atten, C, em, _, _ = projector_inputs(A, B, R, H, W, seed=4)
    theta = np.zeros(1, np.float32)
    # sample points: row min(h, H-1.001), col min(w, W-1.001)
    x = np.minimum(np.arange(H, dtype=np.float64), H - 1.001)
    y = np.minimum(np.arange(W, dtype=np.float64), W - 1.001)
    x0, y0 = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - x0, y - y0
    Cd = C.astype(np.float64)
    samp = ((1 - fx)[:, None] * (1 - fy)[None, :] * Cd[..., x0[:, None], y0[None, :]]
            + (1 - fx)[:, None] * fy[None, :] * Cd[..., x0[:, None], y0[None, :] + 1]
            + fx[:, None] * (1 - fy)[None, :] * Cd[..., x0[:, None] + 1, y0[None, :]]
            + fx[:, None] * fy[None, :] * Cd[..., x0[:, None] + 1, y0[None, :] + 1])
    expect = np.einsum("r,rbhw,bhw->bw", em[0].astype(np.float64), samp, atten[0].astype(np.float64))
    assert np.allclose(oracle.forward_batch_f64(atten, C, em, theta)[0], expect, rtol=1e-12)
    assert rel_l2(oracle.forward_batch_f32(atten, C, em, theta)[0], expect) < 1e-6
    assert rel_l2(oracle.forward_single_f32(atten[:, 0], C[:, 0], em, theta), expect[0][None]) < 1e-6


@pytest.mark.parametrize("R", [1, 3])
def test_adjoint_identity(R):
    A, B, H, W = 7, 3, 20, 26
    atten, C, em, theta, ratio = projector_inputs(A, B, R, H, W, seed=5)
    lhs = (oracle.forward_batch_f64(atten, C, em, theta) * ratio).sum()
    rhs = (C.astype(np.float64) * oracle.backward_batch_f64(atten, ratio, em, theta)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    lhs32 = (oracle.forward_batch_f32(atten, C, em, theta).astype(np.float64) * ratio).sum()
    rhs32 = (C.astype(np.float64) * oracle.backward_batch_f32(atten, ratio, em, theta)).sum()
    assert abs(lhs32 - rhs32) <= 1e-5 * abs(lhs)


def test_single_slice_equals_batch_b1():
    A, R, H, W = 6, 2, 15, 18
    atten, C, em, theta, ratio = projector_inputs(A, 1, R, H, W, seed=6)
    f = oracle.forward_single_f32(atten[:, 0], C[:, 0], em, theta)
    fb = oracle.forward_batch_f32(atten, C, em, theta)[:, 0]
    assert rel_l2(f, fb) < 1e-6
    b = oracle.backward_single_f32(atten[:, 0], ratio[:, 0], em, theta)
    bb = oracle.backward_batch_f32(atten, ratio, em, theta)[:, 0]
    assert rel_l2(b, bb) < 1e-6


# ---------------------------------------------------------------------------
# Huber and MLEM
# ---------------------------------------------------------------------------


def huber_shift(C, delta):
    g = np.zeros_like(C, dtype=np.float64)
    C = C.astype(np.float64)
    for axis, step in ((-2, 1), (-2, -1), (-1, 1), (-1, -1)):
        nb = np.roll(C, step, axis=axis)
        d = np.clip(C - nb, -delta, delta)
        idx = [slice(None)] * C.ndim
        idx[axis] = 0 if step == 1 else -1  # wrapped neighbour does not exist
        d[tuple(idx)] = 0
        g += d
    return g


def test_huber():
    rng = np.random.default_rng(7)
    C = rng.uniform(0, 0.05, (2, 3, 11, 9)).astype(np.float32)
    assert np.allclose(oracle.huber_batch_f64(C, 0.01), huber_shift(C, 0.01), atol=1e-15)
    assert np.allclose(oracle.huber_batch_f32(C, 0.01), huber_shift(C, 0.01), atol=1e-7)


@pytest.mark.parametrize("beta", [0.0, 0.05])
def test_mlem_matches_textbook_dense(small, beta):
    A, B, R, H, W, atten, C, em, theta, _ = small
    n_iter = 6
    I = oracle.forward_batch_f64(atten, C, em, theta) + 0.01
    x0 = np.ones((1, B, H, W))
    got = oracle.mlem_batch_f64(x0, atten, em, theta, I, n_iter, beta, 0.01)
    for b in range(B):
        M = dense_H(atten[:, b].astype(np.float64), theta.astype(np.float64), em[:, 0].astype(np.float64))
        # the kernels clamp at the border; the dense matrix skips: compare on a clamped-consistent operator
        Fk = lambda v: oracle.forward_batch_f64(atten[:, b:b + 1], v.reshape(1, 1, H, W), em, theta).ravel()
        Bk = lambda r: oracle.backward_batch_f64(atten[:, b:b + 1], r.reshape(A, 1, W), em, theta).ravel()
        x = np.ones(H * W)
        sens = np.maximum(Bk(np.ones(A * W)), 1e-6)
        y = I[:, b].ravel()
        for _ in range(n_iter):
            back = Bk(y / np.maximum(Fk(x), 1e-6))
            if beta > 0:
                den = sens + beta * huber_shift(x.reshape(1, H, W), 0.01).ravel()
            else:
                den = sens
            x = np.maximum(x * back / den, 0)
        assert np.allclose(got[0, b].ravel(), x, rtol=1e-10, atol=1e-12)
        # and the operator pair equals the dense matrix away from the clamped ring
        xin = zero_border(np.random.default_rng(b).uniform(0, 1, (H, W)))
        assert np.allclose(M @ xin.ravel(), Fk(xin), rtol=1e-12, atol=1e-12)


def test_mlem_fixed_point(small):
    A, B, R, H, W, atten, C, em, theta, _ = small
    Cpos = C + 0.1
    I = oracle.forward_batch_f64(atten, Cpos, em, theta)
    out = oracle.mlem_batch_f64(Cpos, atten, em, theta, I, 3, 0.0)
    # voxels no detector sample reaches (sensitivity 0, clamped to 1e-6) go to 0
    sens = oracle.backward_batch_f64(atten, np.ones((A, B, W)), em, theta)
    seen = sens > 1e-6
    assert seen.mean() > 0.9
    assert np.allclose(out[seen], Cpos[seen], rtol=1e-10)
    assert np.all(out[~seen] == 0)


def test_mlem_f32_tracks_f64(small):
    A, B, R, H, W, atten, C, em, theta, _ = small
    I = oracle.forward_batch_f64(atten, C, em, theta).astype(np.float32) + 0.01
    x0 = np.ones((1, B, H, W), np.float32)
    a = oracle.mlem_batch_f32(x0, atten, em, theta, I, 20, 1e-3, 0.01)
    b = oracle.mlem_batch_f64(x0, atten, em, theta, I, 20, 1e-3, 0.01)
    assert rel_l2(a, b) < 1e-5


# ---------------------------------------------------------------------------
# attenuation pieces
# ---------------------------------------------------------------------------


def test_mask_correlation_is_conv3d():
    rng = np.random.default_rng(8)
    mu = rng.uniform(0, 1, (2, 6, 9, 8))
    mask = rng.uniform(0, 1, (3, 5, 5))
    K, Z, X, Y = mu.shape
    Mz, Mx, My = mask.shape
    hz, hy = Mz // 2, My // 2
    t = torch.from_numpy(mu)[:, None]
    t = torch.nn.functional.pad(t, (hy, My - 1 - hy, 0, Mx - 1, hz, Mz - 1 - hz))
    ref = torch.nn.functional.conv3d(t, torch.from_numpy(mask)[None, None])[:, 0].numpy()
    assert np.allclose(oracle.atten_f64(mu, mask), ref, rtol=1e-12)
    assert rel_l2(oracle.atten_f32(mu, mask), ref) < 1e-6


def test_incident_closed_form_and_cumprod():
    H, W = 10, 4
    mu = np.full((1, 2, H, W), 0.03, np.float64)
    expect = (1 - 0.03) ** (H - 1 - np.arange(H))
    got = oracle.incident_f64(mu)
    assert np.allclose(got[0, 0, :, 0], expect, rtol=1e-14)
    rng = np.random.default_rng(9)
    mu = rng.uniform(0, 0.05, (2, 3, H, W)).astype(np.float32)
    f = torch.from_numpy(1 - mu)
    cp = torch.flip(torch.cumprod(torch.flip(f, dims=(2,)), dim=2), dims=(2,))
    ref = torch.ones_like(f)
    ref[:, :, :-1] = cp[:, :, 1:]
    # torch's CPU cumprod accumulates float32 in double; the device scan in float32:
    # neither order is reproducible bit for bit, so agreement is to rounding
    assert np.allclose(oracle.incident_f32(mu), ref.numpy(), rtol=2e-6, atol=0)


@pytest.mark.parametrize("shape", [(3, 16, 16), (2, 12, 20)])
def test_rotation_torch_vs_closed_form(shape):
    rng = np.random.default_rng(10)
    x = rng.uniform(0, 1, shape)
    for th in (0.0, 0.3, np.pi / 2, 2.5):
        t64 = oracle.rotate_torch(x, th, dtype="float64")
        c64 = oracle.rotate_f64(x, th)
        assert np.allclose(t64, c64, atol=1e-12)
        t32 = oracle.rotate_torch(x.astype(np.float32), th)
        assert np.abs(t32 - c64).max() < 1e-4


def test_rot90_frames():
    img = np.arange(2 * 3 * 4).reshape(2, 3, 4)
    r = oracle.rot90_clock(img)
    assert r.shape == (2, 4, 3)
    for b in range(2):
        for i in range(4):
            for j in range(3):
                assert r[b, i, j] == img[b, 3 - 1 - j, i]
    back = oracle.rot90_cclock(r)
    assert np.array_equal(back, img)


def test_cone_march_equals_splat_mask_correlation():
    rng = np.random.default_rng(11)
    mu = rng.uniform(0, 1, (1, 7, 12, 9))
    pz, py, q = [1, -1, 0, 1], [0, 2, -1, -3], [4, 3, 1, 8]
    wgt = [0.3, 0.2, 0.25, 0.25]
    X = mu.shape[2]
    M = oracle.ray_splat_mask(pz, py, q, wgt, X)
    a = oracle.atten_f64(mu, M)
    b = oracle.cone_march_f64(mu, pz, py, q, wgt)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-13)


def test_detector_mask_invariants():
    m = oracle.generate_detector_mask(20, 20, 24)
    assert m.ndim == 3 and m.shape[1] == 24
    assert np.all(m >= 0)
    assert np.allclose(m, m[::-1, :, :])          # symmetric in slice
    assert np.allclose(m, m[:, :, ::-1])          # symmetric in column
    assert np.allclose(m, np.transpose(m, (2, 1, 0)))  # equal angles: slice <-> column
    assert np.all(m[:, 0, :] == 0)               # the voxel's own plane is excluded
    # support inside the cone: |offset| <= depth * tan(half angle) + 1
    t = np.tan(np.deg2rad(10))
    zc = m.shape[0] // 2
    for d in range(1, 24):
        nz = np.nonzero(m[:, d, zc])[0]
        if len(nz):
            assert np.abs(nz - zc).max() <= d * t + 1


# ---------------------------------------------------------------------------
# Legacy dense-matrix path (FL/legacy/FL_correction_core.py generate_H_jit,
# mlem_matrix): the legacy-faithful CPU baseline
# ---------------------------------------------------------------------------
def _disc_slice(N, margin, seed):
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:N, 0:N]
    ctr = (N - 1) / 2
    inside = (r - ctr) ** 2 + (c - ctr) ** 2 <= (ctr - margin) ** 2
    return np.where(inside, rng.uniform(0.2, 1.0, (N, N)), 0.0)


def test_legacy_H_theta_zero_is_column_sum():
    """At theta = 0 every detector point samples its own pixel with weight 1."""
    rng = np.random.default_rng(41)
    N = 12
    atten = rng.uniform(0.3, 1.0, (1, N, N))
    C = rng.uniform(0.0, 1.0, (N, N))
    Hm = oracle.legacy_H_f64(atten, np.zeros(1))
    np.testing.assert_allclose(Hm @ C.ravel(), (atten[0] * C).sum(axis=0), rtol=1e-13, atol=0)


def test_legacy_H_matches_projector_away_from_the_border():
    """Where the slice is zero near its border, skipping outside points (legacy)
    and clamping them (the projector) give the same projections."""
    rng = np.random.default_rng(42)
    N, A = 20, 9
    theta = rng.uniform(0, np.pi, A).astype(np.float32)
    atten = rng.uniform(0.3, 1.0, (A, N, N))
    C = _disc_slice(N, 2.0, 43)
    Hm = oracle.legacy_H_f64(atten, theta.astype(np.float64))
    got = (Hm @ C.ravel()).reshape(A, N)
    ref = oracle.forward_batch_f64(atten[:, None], C[None, None], np.ones((A, 1)), theta)[:, 0]
    assert rel_l2(got, ref) <= 1e-12


def test_legacy_mlem_fixed_point_and_unsampled_pixels():
    """y = H x*: x* is a fixed point on sampled pixels; pixels no row samples
    (b_sum = 0) become 0 (the reference's NaN -> 0)."""
    rng = np.random.default_rng(44)
    N, A = 16, 7
    theta = rng.uniform(0.2, 1.3, A)  # no angle near 0: the corners go unsampled
    atten = rng.uniform(0.3, 1.0, (A, N, N))
    Hm = oracle.legacy_H_f64(atten, theta)
    x = rng.uniform(0.5, 1.0, (N, N)).astype(np.float32)
    y = Hm @ x.astype(np.float64).ravel()
    got = oracle.legacy_mlem_f64(Hm, y, x, 3)
    sampled = (Hm.sum(axis=0) > 0).reshape(N, N)
    assert (~sampled).any() and sampled.any()
    np.testing.assert_allclose(got[sampled], x[sampled], rtol=1e-10)
    assert np.all(got[~sampled] == 0)


def test_legacy_mlem_raises_the_poisson_likelihood():
    rng = np.random.default_rng(45)
    N, A = 16, 12
    theta = np.linspace(0, np.pi, A, endpoint=False)
    atten = rng.uniform(0.3, 1.0, (A, N, N))
    Hm = oracle.legacy_H_f64(atten, theta)
    y = rng.poisson(Hm @ _disc_slice(N, 1.0, 46).ravel() * 50) / 50.0
    x = np.ones((N, N), np.float32)
    ll = []
    for n in range(1, 6):
        est = oracle.legacy_mlem_f64(Hm, y, x, n).ravel()
        pf = np.maximum(Hm @ est, 1e-12)
        ll.append(float(np.sum(y * np.log(pf) - pf)))
    assert all(b >= a - 1e-9 * abs(a) for a, b in zip(ll, ll[1:]))


# ---------------------------------------------------------------------------
# FP64 chain of the cone attenuation and of the whole model (used at the
# workloads' own sizes), and the crop case's central-difference fixture
# ---------------------------------------------------------------------------
def test_cone_chain_f64_equals_mask_chain_and_boxes():
    from fl_correction_amd.detector_mask import Cone
    rng = np.random.default_rng(31)
    frac = rng.uniform(0, 1, (2, 6, 20, 24)).astype(np.float32)
    cone = Cone(n=3, half_angle_deg=15, q=8)
    cx = np.array([0.02, 0.03], np.float32)
    cf = rng.uniform(0.01, 0.03, (2, 2)).astype(np.float32)
    for ang in (0.0, 33.0, 121.0):
        full = oracle.total_atten_cone_f64(frac, ang, cx, cf, cone)
        part = oracle.total_atten_cone_f64(frac, ang, cx, cf, cone, 2, 4, 5, 13)
        assert np.array_equal(full[:, 2:4, 5:13], part)
        rot = oracle.rotate_torch(frac, np.float32(ang / 180 * np.pi))
        ref = oracle.xrf_atten_from_mask(oracle.mu_xrf(rot, cf), cone.splat_mask(24)) * \
            oracle.incident_f32(oracle.mu_xray(rot, cx))
        assert rel_l2(ref, full) < 1e-6


def test_model_forward_f64_equals_torch_model():
    import torch_model as tm
    from fl_correction_amd.detector_mask import Cone
    rng = np.random.default_rng(32)
    K, B, N, A = 2, 4, 14, 3
    C = rng.uniform(0, 1, (K, B, N, N))
    angles = np.array([0.0, 40.0, 115.0])
    cx = rng.uniform(0.01, 0.04, (A, K)).astype(np.float32)
    cf = rng.uniform(0.005, 0.03, (A, K, K)).astype(np.float32)
    em = rng.uniform(0.5, 1.5, (A, K)).astype(np.float32)
    cone = Cone(n=3, half_angle_deg=15, q=8)
    a = oracle.model_forward_f64(C, angles, cx, cf, em, cone)
    b = tm.forward(torch.tensor(C), angles, cx, cf, em, cone.splat_mask(N)).numpy()
    # the torch model rounds the rotation matrix to float32 as the reference's float32 callers do
    assert rel_l2(a, b) < 1e-6


def test_torch_model_sparse_taps_equal_conv3d():
    import torch_model as tm
    from fl_correction_amd.detector_mask import Cone
    rng = np.random.default_rng(33)
    mu = torch.tensor(rng.uniform(0, 1, (2, 7, 16, 9)))
    M = torch.tensor(Cone(n=3, half_angle_deg=20, q=8).splat_mask(16).astype(np.float64))
    assert (M != 0).sum() < 0.25 * M.numel()
    sparse = tm.mask_correlation(mu, M)
    import torch.nn.functional as F
    Mz, Mx, My = M.shape
    t = F.pad(mu[:, None], (My // 2, My - 1 - My // 2, 0, Mx - 1, Mz // 2, Mz - 1 - Mz // 2))
    dense = F.conv3d(t, M[None, None])[:, 0]
    assert torch.allclose(sparse, dense, rtol=1e-12, atol=1e-14)


def test_crop_fixture_central_differences():
    """Two of the fixture's 64 FP64 central differences, recomputed from the oracle."""
    import helpers
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_model_crop_fd as mk
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "model_crop_fd.npz"))
    case = helpers.model_crop_case()
    assert np.array_equal(fx["idx"], mk.voxel_indices(case["C_eval"].shape))
    for k in (0, 37):
        fd = mk.central_difference(case, fx["idx"][k])
        assert abs(fd - fx["fd"][k]) <= 1e-9 * np.abs(fx["fd"]).max()
