"""Detector-cone masks: the reference's pyramid mask and the direction-sampled cone.

``generate_detector_mask`` follows FL/detector_mask.py:22-106 (pyramid mask with
partial-volume edge weights, per-depth normalisation Mask^2/sum(Mask), cropped
to (2*N1_0-1, L, 2*N2_0-1) and transposed to (slice, row, col)); it is host-side
data preparation, vectorised with NumPy instead of the reference's numba loops.

``Cone`` is the direction-sampled form: n x n ray directions inside the cone,
each with rational slopes p/q so that a ray's bilinear samples repeat exactly
every q planes.  ``fl_cone_atten`` marches these rays with an exact recursion;
``Cone.splat_mask`` gives the dense mask that makes the reference's
``cal_atten_cuda`` compute the same line integral.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def g_mask(M_shape, alfa, N):
    """Partial-volume weights of one cone half-plane (FL/detector_mask.py:109-151)."""
    tan_a = np.tan(alfa)
    cos_a = np.cos(alfa)
    I = np.arange(M_shape[0])[:, None] + 1
    J = np.arange(M_shape[1])[None, :] + 1
    a = np.abs(N - I).astype(np.float64)
    jt, j1t = J * tan_a, (J - 1) * tan_a
    M = np.zeros((M_shape[0], M_shape[1]))
    c0 = a >= jt
    c1 = ~c0 & (a < j1t) & (a + 1 > jt)
    c2 = ~c0 & ~c1 & (a < jt) & (a + 1 > jt) & (a > j1t)
    c3 = ~c0 & ~c1 & ~c2 & (a + 1 < jt) & (a < j1t) & (a + 1 > j1t)
    c4 = ~(c0 | c1 | c2 | c3)
    d1 = j1t - np.floor(j1t)
    d2 = jt - np.floor(jt)
    with np.errstate(divide="ignore", invalid="ignore"):
        M = np.where(c1, 0.5 * (d1 + d2) / cos_a, M)
        M = np.where(c2, 0.5 * d2 * (d2 / tan_a) / cos_a, M)
        d3 = np.ceil(j1t) - j1t
        M = np.where(c3, (1 - 0.5 * d3 * (d3 / tan_a)) / cos_a, M)
        M = np.where(c4, 1.0 / np.cos(np.arctan(1.0 * (N - I) / J)), M)
    return M


def g_radial_mask_approximate(Mask):
    """Per-depth normalisation Mask[:, :, k] / sum(Mask[:, :, k]) (FL/detector_mask.py:179-186)."""
    dis_sum = Mask.sum(axis=(0, 1))
    with np.errstate(divide="ignore", invalid="ignore"):
        M = Mask / dis_sum[None, None, :]
    M[~np.isfinite(M)] = 0
    return M


def generate_detector_mask(alfa0, theta0, leng0):
    """Pyramid detector mask, shape (2*N1_0-1, leng0, 2*N2_0-1) = (slice, row, col), as
    generate_detector_mask (FL/detector_mask.py:22-106).  alfa0/theta0: full horizontal and
    vertical dispersion angles in degrees; leng0: radial length in voxels."""
    alfa = np.float32(alfa0) / 2 / 180 * np.pi
    theta = np.float32(theta0) / 2 / 180 * np.pi
    N1_0 = np.int16(np.ceil(leng0 * np.tan(alfa)))
    N2_0 = np.int16(np.ceil(leng0 * np.tan(theta)))
    leng = leng0 + 30
    N1 = np.int16(np.ceil(leng * np.tan(alfa)))
    N2 = np.int16(np.ceil(leng * np.tan(theta)))
    s = (2 * N1 - 1, 2 * N2 - 1, leng)
    s0 = (2 * N1_0 - 1, 2 * N2_0 - 1, leng0)
    M1 = g_mask((s[0], s[2]), alfa, N1)
    M2 = g_mask((s[1], s[2]), theta, N2)
    M1[N1 - 1, :] = 1
    M1[N1 - 1, 0] = 0
    M2[N2 - 1, :] = 1
    M2[N2 - 1, 0] = 0
    Mask = M1[:, None, :] * M2[None, :, :]
    Mask3D = g_radial_mask_approximate(Mask) * Mask
    cx, cy = np.floor(s[0] / 2), np.floor(s[1] / 2)
    dx, dy = np.floor(s0[0] / 2), np.floor(s0[1] / 2)
    cut = Mask3D[int(cx - dx):int(cx + dx + 1), int(cy - dy):int(cy + dy + 1), 0:leng0]
    cut = np.transpose(cut, [1, 2, 0])
    cut[np.isnan(cut)] = 0
    return np.ascontiguousarray(cut)


def prep_detector_mask3D(alfa=15, beta=60, length_maximum=200, fn_save=None):
    """Masks for every length from length_maximum down to 7 (FL/detector_mask.py:12-19).
    fn_save, if given, must end in .npz (HDF5 is not available here)."""
    mask = {f"{i}": generate_detector_mask(alfa, beta, i) for i in range(length_maximum, 6, -1)}
    if fn_save:
        if not str(fn_save).endswith(".npz"):
            raise ValueError("fn_save must be a .npz path")
        np.savez(fn_save, **mask)
    return mask


def load_mask3D(fn):
    with np.load(fn, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ---------------------------------------------------------------------------
# Direction-sampled cone
# ---------------------------------------------------------------------------


@dataclass
class Cone:
    """n x n ray directions inside a detector cone.

    Directions (in the mask frame, cone along +X) have slopes
    (dZ/dX, dY/dX) = (pz/q, py/q): for i, j in 0..n-1 the nominal slopes
    tan(half_z) * (2i/(n-1) - 1) and tan(half_y) * (2j/(n-1) - 1), each rounded
    to the nearest multiple of 1/q.  Each ray is weighted by its path length per
    unit step, sqrt(1 + sz^2 + sy^2), divided by n^2, so the line integral is the
    cone-averaged attenuation exponent (the quantity the reference's per-depth
    normalised mask approximates).
    """

    n: int = 3
    half_angle_deg: float = 10.0
    half_angle_z_deg: float | None = None
    q: int = 16
    pz: np.ndarray = field(init=False)
    py: np.ndarray = field(init=False)
    qd: np.ndarray = field(init=False)
    weight: np.ndarray = field(init=False)

    def __post_init__(self):
        n, q = int(self.n), int(self.q)
        if n < 1 or q < 1:
            raise ValueError("n and q must be >= 1")
        hz = self.half_angle_deg if self.half_angle_z_deg is None else self.half_angle_z_deg
        tz, ty = math.tan(math.radians(hz)), math.tan(math.radians(self.half_angle_deg))
        lin = np.zeros(1) if n == 1 else np.linspace(-1.0, 1.0, n)
        pz, py, w = [], [], []
        for a in lin:
            for b in lin:
                iz, iy = int(round(a * tz * q)), int(round(b * ty * q))
                pz.append(iz)
                py.append(iy)
                w.append(math.sqrt(1.0 + (iz / q) ** 2 + (iy / q) ** 2) / (n * n))
        self.pz = np.asarray(pz, np.int32)
        self.py = np.asarray(py, np.int32)
        self.qd = np.full(len(pz), q, np.int32)
        self.weight = np.asarray(w, np.float32)

    @property
    def ndir(self):
        return len(self.pz)

    def splat_mask(self, L):
        """Dense mask (Mz, L, My) for which atten_kernel computes this cone's march with
        taps t = 1..L-1: each ray sample splatted bilinearly onto its 4 voxels."""
        ez = max(1, max(int(math.ceil(abs(int(p)) * (L - 1) / int(q))) + 1 for p, q in zip(self.pz, self.qd)))
        ey = max(1, max(int(math.ceil(abs(int(p)) * (L - 1) / int(q))) + 1 for p, q in zip(self.py, self.qd)))
        M = np.zeros((2 * ez + 1, L, 2 * ey + 1), np.float64)
        t = np.arange(1, L)
        for pz, py, q, w in zip(self.pz, self.py, self.qd, self.weight):
            nz, ny = t * int(pz), t * int(py)
            z0, y0 = nz // int(q), ny // int(q)
            fz, fy = (nz - z0 * q) / q, (ny - y0 * q) / q
            for a, wz in ((0, 1 - fz), (1, fz)):
                for b, wy in ((0, 1 - fy), (1, fy)):
                    np.add.at(M, (ez + z0 + a, t, ey + y0 + b), float(w) * wz * wy)
        return M.astype(np.float32)
