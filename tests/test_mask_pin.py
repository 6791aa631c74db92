"""Independent pin of the pyramid detector mask (FL/detector_mask.py:22-186).

The reference's g_mask (:109-151) splits each (lateral, depth) cell of one cone
half-plane into cases and gives each a hand-derived weight.  Here the weights come
from geometry instead: the area of the cell [a, a+1] x [j-1, j] that lies inside the
wedge u < v*tan(alpha) is the integral over v of clip(v*tan - a, 0, 1), which has the
closed form (R(j*tan - a) - R((j-1)*tan - a)) / tan with R the antiderivative of the
clipped ramp.  The reference divides a partial cell's area by cos(alpha) and gives a
fully covered cell the ray length 1/cos(atan((N-i)/j)) = sqrt(1 + ((N-i)/j)^2).
A wrong triangle or trapezoid weight, Mask/sum(Mask) in place of Mask^2/sum(Mask), a
swapped slice/column axis or a shifted crop fails one of these checks.
"""
import numpy as np
import pytest

from helpers import oracle


def _ramp_antiderivative(x):
    x = np.asarray(x, np.float64)
    return np.where(x <= 0, 0.0, np.where(x >= 1, x - 0.5, 0.5 * x * x))


def _wedge_area(a, j, t):
    """Area of [a, a+1] x [j-1, j] below the line u = v*t (t > 0)."""
    return (_ramp_antiderivative(j * t - a) - _ramp_antiderivative((j - 1) * t - a)) / t


def _supersampled_area(a, j, t, n=400):
    u = a + (np.arange(n) + 0.5) / n
    v = j - 1 + (np.arange(n) + 0.5) / n
    return float((u[:, None] < v[None, :] * t).mean())


def _half_plane_from_geometry(rows, cols, alfa, N):
    """One half-plane of weights (rows, cols) from the wedge areas."""
    t = np.float64(np.tan(alfa))
    cos_a = np.float64(np.cos(alfa))
    i = np.arange(rows)[:, None] + 1
    j = np.arange(cols)[None, :] + 1
    a = np.abs(N - i).astype(np.float64)
    area = _wedge_area(a, j, t)
    full = 1.0 / np.sqrt(1.0 / (1.0 + ((N - i) / j) ** 2))  # sqrt(1 + ((N-i)/j)^2)
    return np.where(area >= 1.0, full, area / cos_a)


def _mask_from_geometry(alfa0, theta0, L):
    """The whole mask from the half-plane areas: the reference's box sizes and crop
    (:47-104), the axis rows set to 1 except at depth 0 (:63-66), the per-depth
    normalisation Mask^2 / sum(Mask) (:77-90, :179-186), laid out (slice, row, col)."""
    alfa = np.float32(alfa0) / 2 / 180 * np.pi
    theta = np.float32(theta0) / 2 / 180 * np.pi
    n1_0, n2_0 = int(np.ceil(L * np.tan(alfa))), int(np.ceil(L * np.tan(theta)))
    Lb = L + 30
    n1, n2 = int(np.ceil(Lb * np.tan(alfa))), int(np.ceil(Lb * np.tan(theta)))
    m1 = _half_plane_from_geometry(2 * n1 - 1, Lb, alfa, n1)
    m2 = _half_plane_from_geometry(2 * n2 - 1, Lb, theta, n2)
    for m, n in ((m1, n1), (m2, n2)):
        m[n - 1, :] = 1.0
        m[n - 1, 0] = 0.0
    # depth j: mask(x, y) = m1[x, j] * m2[y, j]; its sum factorises
    s1, s2 = m1.sum(0), m2.sum(0)
    with np.errstate(divide="ignore", invalid="ignore"):
        norm = np.where(s1 * s2 > 0, 1.0 / (s1 * s2), 0.0)
    x = slice(n1 - n1_0, n1 + n1_0 - 1)
    y = slice(n2 - n2_0, n2 + n2_0 - 1)
    a1, a2 = m1[x, :L], m2[y, :L]
    # (slice, row, col) = (vertical y, depth, horizontal x)
    return (a2[:, :, None] ** 2) * (a1.T[None, :, :] ** 2) * norm[None, :L, None]


ANGLES = [(20, 20), (10, 10), (30, 30), (11, 28), (28, 11), (40, 16)]


@pytest.mark.parametrize("full_angle", [10, 11, 20, 28, 30, 40, 60])
def test_half_plane_weights_are_wedge_areas(full_angle):
    alfa = np.float32(full_angle) / 2 / 180 * np.pi
    L = 70
    N = int(np.ceil(L * np.tan(alfa)))
    rows, cols = 2 * N - 1, L
    got = oracle.g_mask_loops((rows, cols), alfa, N)
    exp = _half_plane_from_geometry(rows, cols, alfa, N)
    # the axis row (i = N) is overwritten by the generator (:63-66); the rest is geometry
    off = np.arange(rows) != N - 1
    assert np.allclose(got[off], exp[off], rtol=1e-12, atol=1e-13)
    # every case of the reference's split is exercised
    t = float(np.tan(alfa))
    i = np.arange(rows)[:, None] + 1
    j = np.arange(cols)[None, :] + 1
    area = _wedge_area(np.abs(N - i), j, t)
    partial = (area > 0) & (area < 1)
    assert partial.sum() > 10 and (area >= 1).sum() > 10 and (area == 0).sum() > 10
    # the closed form against a supersampled count on partial cells of each kind
    ii, jj = np.nonzero(partial)
    for k in np.linspace(0, len(ii) - 1, 12).astype(int):
        a_, j_ = abs(N - (ii[k] + 1)), jj[k] + 1
        assert abs(_wedge_area(a_, j_, t) - _supersampled_area(a_, j_, t)) < 5e-3


@pytest.mark.parametrize("full_angle", [10, 20, 30])
def test_product_half_plane_equals_geometry(full_angle):
    from fl_correction_amd.detector_mask import g_mask
    alfa = np.float32(full_angle) / 2 / 180 * np.pi
    N = int(np.ceil(60 * np.tan(alfa)))
    exp = _half_plane_from_geometry(2 * N - 1, 60, alfa, N)
    off = np.arange(2 * N - 1) != N - 1
    assert np.allclose(g_mask((2 * N - 1, 60), alfa, N)[off], exp[off], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("alfa0,theta0", ANGLES)
@pytest.mark.parametrize("L", [24, 64])
def test_mask_equals_geometry(alfa0, theta0, L):
    from fl_correction_amd.detector_mask import generate_detector_mask
    exp = _mask_from_geometry(alfa0, theta0, L)
    for got in (oracle.generate_detector_mask(alfa0, theta0, L), generate_detector_mask(alfa0, theta0, L)):
        assert got.shape == exp.shape
        assert np.allclose(got, exp, rtol=1e-10, atol=1e-14 * exp.max())


@pytest.mark.parametrize("alfa0,theta0", ANGLES)
def test_per_depth_sum_is_sum_of_squares_over_sum(alfa0, theta0):
    """sum over (slice, col) of depth d = sum(Mask_d^2) / sum(Mask_d): nothing of the cone
    is cropped away at depths below L."""
    L = 40
    m = oracle.generate_detector_mask(alfa0, theta0, L)
    alfa = np.float32(alfa0) / 2 / 180 * np.pi
    theta = np.float32(theta0) / 2 / 180 * np.pi
    Lb = L + 30
    n1, n2 = int(np.ceil(Lb * np.tan(alfa))), int(np.ceil(Lb * np.tan(theta)))
    m1 = _half_plane_from_geometry(2 * n1 - 1, Lb, alfa, n1)
    m2 = _half_plane_from_geometry(2 * n2 - 1, Lb, theta, n2)
    for mm, n in ((m1, n1), (m2, n2)):
        mm[n - 1, :] = 1.0
        mm[n - 1, 0] = 0.0
    for d in range(1, L):
        full = np.outer(m1[:, d], m2[:, d])
        assert np.isclose(m[:, d, :].sum(), (full ** 2).sum() / full.sum(), rtol=1e-12)
    assert m[:, 0, :].sum() == 0


@pytest.mark.parametrize("alfa0,theta0", ANGLES)
def test_prefix_property(alfa0, theta0):
    """mask(L) is the centre crop of mask(L') cut to its first L depths, for L <= L'."""
    big = oracle.generate_detector_mask(alfa0, theta0, 90)
    for L in (7, 20, 33, 61, 90):
        m = oracle.generate_detector_mask(alfa0, theta0, L)
        cz, cy = (big.shape[0] - m.shape[0]) // 2, (big.shape[2] - m.shape[2]) // 2
        crop = big[cz:cz + m.shape[0], :L, cy:cy + m.shape[2]]
        assert np.allclose(m, crop, rtol=1e-13, atol=0)


def test_planted_mistakes_are_caught(monkeypatch):
    """Mistakes planted in the oracle's generator fail the geometry comparison: a triangle
    weight without its 1/2, and the per-depth normalisation Mask/sum(Mask)."""
    exp = _mask_from_geometry(20, 20, 24)
    good = oracle.g_mask_loops

    def no_half(M_shape, alfa, N):
        M = good(M_shape, alfa, N)
        t, c = float(np.tan(np.float32(alfa))), float(np.cos(np.float32(alfa)))
        i = np.arange(M_shape[0])[:, None] + 1
        j = np.arange(M_shape[1])[None, :] + 1
        a = np.abs(N - i)
        tri = (a < j * t) & (a + 1 > j * t) & (a > (j - 1) * t)
        d = j * t - np.floor(j * t)
        return np.where(tri, d * d / t / c, M)

    monkeypatch.setattr(oracle, "g_mask_loops", no_half)
    assert not np.allclose(oracle.generate_detector_mask(20, 20, 24), exp, rtol=1e-3)
    monkeypatch.setattr(oracle, "g_mask_loops", good)
    assert np.allclose(oracle.generate_detector_mask(20, 20, 24), exp, rtol=1e-10, atol=1e-14 * exp.max())
    # Mask / sum(Mask): the square root of Mask^2 / sum(Mask) * sum(Mask) per depth
    s = exp.sum(axis=(0, 2), keepdims=True)
    unsquared = np.sqrt(exp * np.where(s > 0, s, 1.0))
    assert not np.allclose(unsquared, exp, rtol=1e-3)
