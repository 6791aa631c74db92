"""CPU oracle for the XRF self-absorption hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only to check results (or, for the
baseline, as the timed CPU port).  The product package never imports it.

Contents
  * ctypes bindings of ``fl_oracle.c`` (FP32 same-order restatements of the
    reference kernels and FP64 yardsticks), built by ``oracle/Makefile``;
  * NumPy restatements of the reference's host-side Python: the linear
    attenuation assembly, the rot90 frame change, the detector-mask generator;
  * the reference's own torch rotation (affine_grid + grid_sample), run on CPU;
  * the direction-sampled cone: its ray-splat mask and direct FP64 march.

Every function cites the reference file:line it follows (paths relative to
the reference root).  Functions whose only check is a property test say so
in their docstring.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libfl_oracle.so")
_lib = None


def build() -> str:
    """Compile fl_oracle.c with the committed Makefile (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


_i = ctypes.c_int
_d = ctypes.c_double
_fl = ctypes.c_float

# ----------------------------------------------------------------------------
# A1  atten_kernel  (FL/cuda_lib/atten_cuda_kernel.cu:10-57)
# ----------------------------------------------------------------------------


def atten_f32(mu, mask):
    """FP32 same-order ``cal_atten_cuda`` (atten_cuda.cpp:12-28 -> .cu:10-57)."""
    mu, mask = _f32(mu), _f32(mask)
    out = np.empty_like(mu)
    K, Z, X, Y = mu.shape
    Mz, Mx, My = mask.shape
    lib().or_atten_f32(_p(mu), _p(mask), _p(out), _i(K), _i(Z), _i(X), _i(Y), _i(Mz), _i(Mx), _i(My))
    return out


def atten_f64(mu, mask):
    mu, mask = _f64(mu), _f64(mask)
    out = np.empty_like(mu)
    K, Z, X, Y = mu.shape
    Mz, Mx, My = mask.shape
    lib().or_atten_f64(_p(mu), _p(mask), _p(out), _i(K), _i(Z), _i(X), _i(Y), _i(Mz), _i(Mx), _i(My))
    return out


# ----------------------------------------------------------------------------
# F1 / B1  (atten_cuda_kernel.cu:91-265 single slice, :298-487 batch)
# ----------------------------------------------------------------------------


def forward_single_f32(atten, C, em_cs, theta):
    """``forward_emission`` (atten_cuda.cpp:82-108, kernel .cu:91-142)."""
    atten, C, em_cs, theta = _f32(atten), _f32(C), _f32(em_cs), _f32(theta)
    A, H, W = atten.shape
    R = C.shape[0]
    Pf = np.empty((A, W), np.float32)
    lib().or_forward_single_f32(_p(atten), _p(C), _p(em_cs), _p(theta), _p(Pf), _i(A), _i(H), _i(W), _i(R))
    return Pf


def backward_single_f32(atten, ratio, em_cs, theta):
    """``backward_emission`` (atten_cuda.cpp:113-135, kernel .cu:177-231)."""
    atten, ratio, em_cs, theta = _f32(atten), _f32(ratio), _f32(em_cs), _f32(theta)
    A, H, W = atten.shape
    R = em_cs.shape[1]
    out = np.empty((R, H, W), np.float32)
    lib().or_backward_single_f32(_p(atten), _p(ratio), _p(em_cs), _p(theta), _p(out), _i(A), _i(H), _i(W), _i(R))
    return out


def forward_batch_f32(atten, C, em_cs, theta):
    """``forward_emission_batch`` (atten_cuda.cpp:142-182, kernel .cu:298-337)."""
    atten, C, em_cs, theta = _f32(atten), _f32(C), _f32(em_cs), _f32(theta)
    A, B, H, W = atten.shape
    R = C.shape[0]
    Pf = np.empty((A, B, W), np.float32)
    lib().or_forward_batch_f32(_p(atten), _p(C), _p(em_cs), _p(theta), _p(Pf), _i(A), _i(B), _i(R), _i(H), _i(W))
    return Pf


def backward_batch_f32(atten, ratio, em_cs, theta):
    """``backward_emission_batch`` (atten_cuda.cpp:188-214, kernel .cu:390-446)."""
    atten, ratio, em_cs, theta = _f32(atten), _f32(ratio), _f32(em_cs), _f32(theta)
    A, B, H, W = atten.shape
    R = em_cs.shape[1]
    out = np.empty((R, B, H, W), np.float32)
    lib().or_backward_batch_f32(_p(atten), _p(ratio), _p(em_cs), _p(theta), _p(out), _i(A), _i(B), _i(R), _i(H), _i(W))
    return out


def forward_batch_f64(atten, C, em_cs, theta):
    atten, C, em_cs, theta = _f64(atten), _f64(C), _f64(em_cs), _f64(np.asarray(theta, np.float32))
    A, B, H, W = atten.shape
    R = C.shape[0]
    Pf = np.empty((A, B, W), np.float64)
    lib().or_forward_batch_f64(_p(atten), _p(C), _p(em_cs), _p(theta), _p(Pf), _i(A), _i(B), _i(R), _i(H), _i(W))
    return Pf


def backward_batch_f64(atten, ratio, em_cs, theta):
    atten, ratio, em_cs, theta = _f64(atten), _f64(ratio), _f64(em_cs), _f64(np.asarray(theta, np.float32))
    A, B, H, W = atten.shape
    R = em_cs.shape[1]
    out = np.empty((R, B, H, W), np.float64)
    lib().or_backward_batch_f64(_p(atten), _p(ratio), _p(em_cs), _p(theta), _p(out), _i(A), _i(B), _i(R), _i(H), _i(W))
    return out


# ----------------------------------------------------------------------------
# U1 / U2  (atten_cuda_kernel.cu:495-616, atten_cuda.cpp:244-345)
# ----------------------------------------------------------------------------


def huber_batch_f32(C, delta):
    """huber_grad_batch_kernel (.cu:564-616); 3-D input = the single kernel (.cu:495-534)."""
    C = _f32(C)
    squeeze = C.ndim == 3
    if squeeze:
        C = C[:, None]
    R, B, H, W = C.shape
    g = np.empty_like(C)
    lib().or_huber_batch_f32(_p(C), _p(g), _i(R), _i(B), _i(H), _i(W), _fl(delta))
    return g[:, 0] if squeeze else g


def huber_batch_f64(C, delta):
    C = _f64(C)
    squeeze = C.ndim == 3
    if squeeze:
        C = C[:, None]
    R, B, H, W = C.shape
    g = np.empty_like(C)
    lib().or_huber_batch_f64(_p(C), _p(g), _i(R), _i(B), _i(H), _i(W), _d(delta))
    return g[:, 0] if squeeze else g


def mlem_batch_f32(C_init, atten, em_cs, theta, I, n_iter, beta=1e-3, delta=0.01):
    """``mlem_cuda_batch`` (atten_cuda.cpp:283-345); 3-D inputs = ``mlem_cuda`` (:244-281)."""
    C_init, atten, em_cs, theta, I = _f32(C_init), _f32(atten), _f32(em_cs), _f32(theta), _f32(I)
    squeeze = C_init.ndim == 3
    if squeeze:
        C_init, atten, I = C_init[:, None], atten[:, None], I[:, None]
        C_init, atten, I = _f32(C_init), _f32(atten), _f32(I)
    R, B, H, W = C_init.shape
    A = atten.shape[0]
    out = np.empty_like(C_init)
    lib().or_mlem_batch_f32(_p(C_init), _p(atten), _p(em_cs), _p(theta), _p(I), _p(out), _i(R), _i(A), _i(B), _i(H), _i(W),
                            _i(n_iter), _d(beta), _d(delta))
    return out[:, 0] if squeeze else out


def mlem_batch_f64(C_init, atten, em_cs, theta, I, n_iter, beta=1e-3, delta=0.01):
    C_init, atten, em_cs, I = _f64(C_init), _f64(atten), _f64(em_cs), _f64(I)
    theta = _f64(np.asarray(theta, np.float32))
    squeeze = C_init.ndim == 3
    if squeeze:
        C_init, atten, I = _f64(C_init[:, None]), _f64(atten[:, None]), _f64(I[:, None])
    R, B, H, W = C_init.shape
    A = atten.shape[0]
    out = np.empty_like(C_init)
    lib().or_mlem_batch_f64(_p(C_init), _p(atten), _p(em_cs), _p(theta), _p(I), _p(out), _i(R), _i(A), _i(B), _i(H), _i(W),
                            _i(n_iter), _d(beta), _d(delta))
    return out[:, 0] if squeeze else out


# ----------------------------------------------------------------------------
# A2 / A4  linear attenuation assembly and the incident-beam product
# ----------------------------------------------------------------------------


def mu_xray(frac, coef_x):
    """``torch_extract_3D_mu_xray`` (torch_functions.py:39-52): mu_x[j] = frac[j]*coef_j,
    coef_j = cs[comp_j-x]*rho[comp_j]*pix rounded to float32 (:48-50). Per element (Q2)."""
    frac = _f32(frac)
    coef = np.asarray(coef_x, np.float64).astype(np.float32)
    return (frac * coef[:, None, None, None]).astype(np.float32)


def mu_xrf(frac, coef_xrf):
    """``torch_extract_3D_mu_xrf`` (torch_functions.py:71-81): mu[i] = sum_j frac[j]*coef[j,i],
    coef[j,i] = cs[comp_j-elem_i]*rho*pix rounded to float32, accumulated in j order from 0."""
    frac = _f32(frac)
    coef = np.asarray(coef_xrf, np.float64).astype(np.float32)  # (K_comp, K_line)
    K = coef.shape[1]
    out = np.zeros((K,) + frac.shape[1:], np.float32)
    for i in range(K):
        for j in range(frac.shape[0]):
            out[i] = out[i] + frac[j] * coef[j, i]
    return out


def incident_f32(mux):
    """``torch_xray_atten_3D`` (torch_functions.py:95-102), FP32 cumprod order."""
    mux = _f32(mux)
    H, W = mux.shape[-2:]
    N = int(np.prod(mux.shape[:-2]))
    out = np.empty_like(mux)
    lib().or_incident_f32(_p(mux), _p(out), _i(N), _i(H), _i(W))
    return out


def incident_f64(mux):
    mux = _f64(mux)
    H, W = mux.shape[-2:]
    N = int(np.prod(mux.shape[:-2]))
    out = np.empty_like(mux)
    lib().or_incident_f64(_p(mux), _p(out), _i(N), _i(H), _i(W))
    return out


def rot90_clock(img):
    """``torch_rot90_3D(img, ax=0, mode='clock')`` (torch_image_util.py:117-147, :130-132) on (..., B, H, W):
    transpose(1,2) then flip the last axis -> out[b,i,j] = img[b, H-1-j, i]."""
    return np.flip(np.swapaxes(img, -1, -2), axis=-1)


def rot90_cclock(img):
    """``torch_rot90_3D(img, ax=0, mode='c-clock')``: transpose(1,2) then flip axis 1
    -> out[b,i,j] = img[b, j, W-1-i] for img (B, H, W)."""
    return np.flip(np.swapaxes(img, -1, -2), axis=-2)


def xrf_atten_from_mask(mu_line, mask):
    """``torch_xrf_atten_3D`` (torch_functions.py:110-176), offset 0, detector 'r':
    rot90 clock (:143), cal_atten_cuda (:152), rot90 c-clock (:171), exp(-res) (:174).
    mu_line: (K, B, H, W) linear attenuation at each line.  FP32 same-order."""
    mu_r = np.ascontiguousarray(rot90_clock(_f32(mu_line)))
    res = atten_f32(mu_r, mask)
    res = np.ascontiguousarray(rot90_cclock(res))
    return np.exp(-res).astype(np.float32)


# ----------------------------------------------------------------------------
# A5  rotation: the reference's own torch calls, run on CPU
# ----------------------------------------------------------------------------


def rotate_torch(x, theta, dtype="float32"):
    """``torch_rotate_image_nd(x, theta)`` (torch_image_util.py:13-114) with its defaults
    (enable_crop=True, align_corners=False, bilinear, zeros), executed with the container's
    torch on CPU.  x: (..., H, W) numpy; theta in radians (float32 tensor in the reference's
    callers, core.py:35,67)."""
    import torch
    import torch.nn.functional as F
    tdt = torch.float32 if dtype == "float32" else torch.float64
    xt = torch.as_tensor(np.ascontiguousarray(x)).to(tdt)
    shape = xt.shape
    H, W = shape[-2:]
    xt = xt.reshape(-1, 1, H, W)
    N = xt.shape[0]
    th = float(np.float32(theta))  # math.cos(theta) of a float32 tensor, as at :77-78
    cos_t, sin_t = math.cos(th), math.sin(th)
    mat = torch.tensor([[cos_t, -sin_t, 0.0], [sin_t, cos_t, 0.0]], dtype=tdt).unsqueeze(0).repeat(N, 1, 1)
    grid = F.affine_grid(mat, size=(N, 1, H, W), align_corners=False)
    y = F.grid_sample(xt, grid, mode="bilinear", padding_mode="zeros", align_corners=False)
    return y.reshape(shape).numpy()


def rotate_f64(x, theta):
    """Closed-form FP64 statement of the same rotation (no torch)."""
    x = _f64(x)
    shape = x.shape
    H, W = shape[-2:]
    N = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
    out = np.empty_like(x)
    lib().or_rotate_f64(_p(x), _d(float(np.float32(theta))), _p(out), _i(N), _i(H), _i(W))
    return out


# ----------------------------------------------------------------------------
# A7  pyramid detector mask  (FL/detector_mask.py:22-186)
# ----------------------------------------------------------------------------


def g_mask_loops(M_shape, alfa, N):
    """Scalar-loop restatement of ``g_mask`` (detector_mask.py:109-151).  The reference
    is numba-compiled: tan and cos of the float32 angle are float32, and every product
    with the int64 loop indices is float64 (numba's int64 x float32 -> float64), so the
    loops here take the float32 values as float64 numbers.  Pinned against the wedge
    areas of each cell in tests/test_mask_pin.py."""
    tan_alfa = float(np.tan(np.float32(alfa)))
    cos_alfa = float(np.cos(np.float32(alfa)))
    M1 = np.zeros((M_shape[0], M_shape[1]))
    for I in range(M_shape[0]):
        for J in range(M_shape[1]):
            i = I + 1
            j = J + 1
            a = abs(N - i)
            if a >= j * tan_alfa:
                M1[I, J] = 0
            elif a < (j - 1) * tan_alfa and (a + 1) > j * tan_alfa:
                d1 = (j - 1) * tan_alfa - np.floor((j - 1) * tan_alfa)
                d2 = j * tan_alfa - np.floor(j * tan_alfa)
                M1[I, J] = 0.5 * (d1 + d2) / cos_alfa
            elif a < j * tan_alfa and (a + 1) > j * tan_alfa and a > (j - 1) * tan_alfa:
                d1 = j * tan_alfa - np.floor(j * tan_alfa)
                M1[I, J] = 0.5 * d1 * (d1 / tan_alfa) / cos_alfa
            elif (a + 1) < j * tan_alfa and a < (j - 1) * tan_alfa and (a + 1) > (j - 1) * tan_alfa:
                d1 = np.ceil((j - 1) * tan_alfa) - (j - 1) * tan_alfa
                M1[I, J] = (1 - 0.5 * d1 * (d1 / tan_alfa)) / cos_alfa
            else:
                tmp = np.arctan(1.0 * (N - i) / j)
                M1[I, J] = 1 / np.cos(tmp)
    return M1


def generate_detector_mask(alfa0, theta0, leng0):
    """Restatement of ``generate_detector_mask`` (detector_mask.py:22-106) with
    ``g_radial_mask_approximate`` (:179-186).  Returns (slice, row, col) = (Mz, Mx, My).
    Pinned by the wedge-area construction, the per-depth sums and the prefix property
    (tests/test_mask_pin.py)."""
    alfa = np.float32(alfa0) / 2 / 180 * np.pi
    theta = np.float32(theta0) / 2 / 180 * np.pi
    N1_0 = np.int16(np.ceil(leng0 * np.tan(alfa)))
    N2_0 = np.int16(np.ceil(leng0 * np.tan(theta)))
    leng = leng0 + 30
    N1 = np.int16(np.ceil(leng * np.tan(alfa)))
    N2 = np.int16(np.ceil(leng * np.tan(theta)))
    s = (2 * N1 - 1, 2 * N2 - 1, leng)
    s0 = (2 * N1_0 - 1, 2 * N2_0 - 1, leng0)
    M1 = g_mask_loops((s[0], s[2]), alfa, N1)
    M2 = g_mask_loops((s[1], s[2]), theta, N2)
    M1[N1 - 1, :] = 1
    M1[N1 - 1, 0] = 0
    M2[N2 - 1, :] = 1
    M2[N2 - 1, 0] = 0
    Mask = M1[:, None, :] * M2[None, :, :]
    dis_sum = Mask.sum(axis=(0, 1))
    with np.errstate(divide="ignore", invalid="ignore"):
        M_normal = Mask / dis_sum[None, None, :]
    M_normal[~np.isfinite(M_normal)] = 0
    Mask3D = M_normal * Mask
    cent = (np.floor(s[0] / 2), np.floor(s[1] / 2))
    delt = (np.floor(s0[0] / 2), np.floor(s0[1] / 2))
    xs, xe = int(cent[0] - delt[0]), int(cent[0] + delt[0] + 1)
    ys, ye = int(cent[1] - delt[1]), int(cent[1] + delt[1] + 1)
    cut = Mask3D[xs:xe, ys:ye, 0:leng0]
    cut = np.transpose(cut, [1, 2, 0])
    cut[np.isnan(cut)] = 0
    return cut


# ----------------------------------------------------------------------------
# Direction-sampled cone
# ----------------------------------------------------------------------------


def ray_splat_mask(pz, py, q, wgt, L):
    """Dense mask M (Mz, L, My) such that ``atten_kernel(mu, M)`` equals the direction-sampled
    cone march with taps t = 1..L-1: each ray point (t*pz/q, t, t*py/q) is splatted
    bilinearly onto its four integer neighbours (exact rational offsets)."""
    pz, py, q = (np.asarray(v, np.int64) for v in (pz, py, q))
    wgt = np.asarray(wgt, np.float64)
    ez = int(max(1, max(int(math.ceil(abs(int(a)) * (L - 1) / int(b))) + 1 for a, b in zip(pz, q))))
    ey = int(max(1, max(int(math.ceil(abs(int(a)) * (L - 1) / int(b))) + 1 for a, b in zip(py, q))))
    M = np.zeros((2 * ez + 1, L, 2 * ey + 1), np.float64)
    for d in range(len(pz)):
        for t in range(1, L):
            nz, ny = t * int(pz[d]), t * int(py[d])
            z0, y0 = nz // int(q[d]), ny // int(q[d])
            fz, fy = (nz - z0 * q[d]) / q[d], (ny - y0 * q[d]) / q[d]
            for a, wz in ((0, 1 - fz), (1, fz)):
                for b, wy in ((0, 1 - fy), (1, fy)):
                    if wz == 0 or wy == 0:
                        continue
                    M[ez + z0 + a, t, ey + y0 + b] += wgt[d] * wz * wy
    return M


def cone_march_f64(mu, pz, py, q, wgt, z0=0, z1=None, y0=0, y1=None):
    """Direct FP64 march of the direction-sampled cone on mu (K, Z, X, Y), cone along +X.
    With z0/z1 (y0/y1), only outputs z0 <= z < z1 (y0 <= y < y1); all of mu is read:
    (K, z1-z0, X, y1-y0)."""
    mu = _f64(mu)
    K, Z, X, Y = mu.shape
    z1 = Z if z1 is None else int(z1)
    y1 = Y if y1 is None else int(y1)
    if not (0 <= z0 < z1 <= Z and 0 <= y0 < y1 <= Y):
        raise ValueError("need 0 <= z0 < z1 <= Z and 0 <= y0 < y1 <= Y")
    pz, py, q = (np.ascontiguousarray(v, dtype=np.int32) for v in (pz, py, q))
    wgt = _f64(wgt)
    out = np.empty((K, z1 - z0, X, y1 - y0))
    for k in range(K):
        lib().or_cone_march_box_f64(_p(mu[k]), _p(out[k]), _i(Z), _i(X), _i(Y), _i(z0), _i(z1), _i(y0), _i(y1),
                                    _p(pz), _p(py), _p(q), _p(wgt), _i(len(pz)))
    return out


def total_atten_cone_f64(frac, angle_deg, coef_x, coef_xrf, cone, b0=0, b1=None, r0=0, r1=None):
    """Total attenuation A[e] = exp(-Cone(mu_e)) * incident_e at one angle, detector 'r',
    for output slices b0 <= b < b1 and rows r0 <= r < r1: (K, b1-b0, r1-r0, W).  The
    reference's driver chain (FL/core.py:19-101: rotate, torch_extract_3D_mu_xray / _xrf,
    torch_xray_atten_3D, torch_xrf_atten_3D with the cone's splat mask, product) with the
    rotation run by the container's torch in float32 (the reference's own op, rotate_torch)
    and every later step in float64 (incident_f64, the direct cone march).  In the mask
    frame of detector 'r' row r is y = H-1-r."""
    frac = _f32(frac)
    K, B, H, W = frac.shape
    b1 = B if b1 is None else int(b1)
    r1 = H if r1 is None else int(r1)
    rot = rotate_torch(frac, np.float32(angle_deg / 180.0 * np.pi))
    ax = incident_f64(mu_xray(rot[:, b0:b1], coef_x))[:, :, r0:r1]
    mu_line = mu_xrf(rot, coef_xrf).astype(np.float64)
    T = cone_march_f64(np.ascontiguousarray(rot90_clock(mu_line)), cone.pz, cone.py, cone.qd,
                       np.asarray(cone.weight, np.float64), b0, b1, H - r1, H - r0)
    # T[k, b, x=col, y'] with y' = (H-1-r) - (H-r1) = r1-1-r: row r is y' reversed
    T = np.swapaxes(T, -1, -2)[:, :, ::-1, :]
    return np.exp(-T) * ax


def model_forward_f64(C, angles_deg, coef_x, coef_xrf, em, cone):
    """Pf (K, A, B, W) of the attenuated model in float64 (detector 'r'): per angle the
    total attenuation of C (total_atten_cone_f64, rotation in float64 by the container's
    torch) and the projector oracle forward_batch_f64 per line with em[a, e]."""
    C = _f64(C)
    K, B, H, W = C.shape
    A = len(angles_deg)
    out = np.empty((K, A, B, W))
    for a, ang in enumerate(angles_deg):
        th = np.float32(ang / 180.0 * np.pi)
        rot = rotate_torch(C, th, dtype="float64")
        ax = incident_f64(rot * np.asarray(coef_x[a], np.float32).astype(np.float64)[:, None, None, None])
        cf = np.asarray(coef_xrf[a], np.float32).astype(np.float64)
        mu_line = np.einsum("jbhw,je->ebhw", rot, cf)
        T = cone_march_f64(np.ascontiguousarray(rot90_clock(mu_line)), cone.pz, cone.py, cone.qd,
                           np.asarray(cone.weight, np.float64))
        At = np.exp(-np.ascontiguousarray(rot90_cclock(T))) * ax
        for e in range(K):
            out[e, a] = forward_batch_f64(At[e][None], C[e][None], np.array([[em[a, e]]], np.float64),
                                          np.array([th], np.float64))[0]
    return out


# ----------------------------------------------------------------------------
# Legacy CPU path  (FL/legacy/FL_correction_core.py: generate_H_jit :2203-2279,
# mlem_matrix :99-119), FP64 -- the legacy-faithful CPU baseline
# ----------------------------------------------------------------------------
def legacy_H_f64(atten, theta):
    """Dense system matrix of one slice, (A*W, H*W): row i*W + q is detector
    column q at angle i, samples outside the image skipped (not clamped).
    atten (A, H, W), theta (A) radians."""
    atten = _f64(atten)
    A, H, W = atten.shape
    th = _f64(theta)
    Hm = np.empty((A * W, H * W), np.float64)
    lib().or_legacy_H_f64(_p(th), _p(atten), _p(Hm), _i(A), _i(H), _i(W))
    return Hm


def legacy_mlem_f64(Hm, y, C0, n_iter):
    """mlem_matrix on one slice: Hm (A*W, H*W), y (A*W) projections, C0 (H, W)
    start (taken as float32); returns the (H, W) FP64 estimate."""
    Hm = _f64(Hm)
    nrow, npix = Hm.shape
    y = _f64(np.ravel(y))
    c0 = _f32(np.ravel(C0))
    if y.size != nrow or c0.size != npix:
        raise ValueError("y must have one value per row of Hm and C0 one per column")
    out = np.empty(npix, np.float64)
    lib().or_legacy_mlem_f64(_p(Hm), _p(y), _p(c0), _p(out), ctypes.c_int64(nrow), ctypes.c_int64(npix),
                             _i(int(n_iter)))
    return out.reshape(np.shape(C0))


def rel_l2(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = np.linalg.norm(y.ravel())
    return float(np.linalg.norm((x - y).ravel()) / (n if n > 0 else 1.0))
