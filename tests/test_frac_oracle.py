"""The cal_frac oracle (oracle/frac_oracle.py) pinned on the CPU.

The multi-Otsu restatement (skimage is absent) against the textbook
between-class variance over every split in FP64, and against closed forms:
three distinct values give thresholds at their first two bins; a trimodal
image splits between its modes.  The hole fill, dilation and erosion are
scipy.ndimage's own (as the reference calls them); the known-answer cases
here check the oracle calls them as the reference does.
"""
import numpy as np
import pytest

import helpers  # noqa: F401  (puts oracle/ on sys.path)
import frac_oracle as fo


def _prob(img):
    hist, _ = fo.histogram_centers(img)
    return (hist / hist.sum()).astype(np.float32)


@pytest.mark.parametrize("seed", range(6))
def test_multiotsu_split_maximises_between_class_variance(seed):
    rng = np.random.default_rng(seed)
    kind = seed % 3
    if kind == 0:
        img = rng.gamma(2.0, 1.0, (40, 50))
    elif kind == 1:
        img = np.concatenate([rng.normal(m, s, 700) for m, s in ((0, 0.3), (2, 0.5), (5, 0.4))])
    else:
        img = rng.random((30, 30)) ** 3
    img = img.astype(np.float32)
    _, idx = fo.multiotsu_threshold(img)
    bf = fo.multiotsu_brute_f64(_prob(img))
    best = max(bf.values())
    # float32 moments: the chosen split's FP64 variance is the maximum to rounding
    assert bf[idx] >= best * (1 - 2e-6)
    # and no split with a clearly larger variance precedes or follows it
    assert all(v <= bf[idx] * (1 + 2e-6) for v in bf.values())


def test_multiotsu_three_values_and_failures():
    img = np.array([[1.0, 1.0, 5.0], [9.0, 9.0, 9.0]], np.float32)
    t, idx = fo.multiotsu_threshold(img)
    hist, centers = fo.histogram_centers(img)
    nz = np.flatnonzero(hist)
    assert idx == (nz[0], nz[1]) and t == centers[nz[0]]
    with pytest.raises(ValueError):
        fo.multiotsu_threshold(np.ones((4, 4), np.float32))          # one value
    with pytest.raises(ValueError):
        fo.multiotsu_threshold(np.array([[0, 1], [0, 1]], np.float32))  # two values
    with pytest.raises(ValueError):
        fo.multiotsu_threshold(np.array([[0, np.nan], [2, 3]], np.float32))


def test_multiotsu_trimodal_splits_between_modes():
    rng = np.random.default_rng(7)
    img = np.concatenate([rng.uniform(0, 1, 500), rng.uniform(4, 5, 500), rng.uniform(9, 10, 500)]).astype(np.float32)
    t, _ = fo.multiotsu_threshold(img)
    # the threshold is the centre of class 0's last bin: within a bin of the
    # first mode's top, below the second mode
    width = (img.max() - img.min()) / 256
    assert img[:500].max() - width <= t < img[500:1000].min()


def test_adaptive_threshold_known_masks():
    # a ring above background: Otsu keeps the ring, the fill closes its hole,
    # one cross dilation grows it by one pixel in the four directions
    H = W = 41
    yy, xx = np.mgrid[:H, :W]
    r = np.hypot(yy - 20, xx - 20)
    img = np.where((r > 8) & (r < 12), 10.0, 0.0) + np.where(r < 3, 5.0, 0.0)
    img = img + 0.001 * (yy % 3)  # background values within the first bin
    m0 = fo.adaptive_threshold(img)
    assert m0[20, 30] == 1 and m0[20, 20] == 1 and m0[20, 25] == 0 and m0[0, 0] == 0
    m1 = fo.adaptive_threshold(img, fill_hole=True)
    assert m1[20, 25] == 1 and m1[0, 0] == 0                       # the hole filled
    assert np.array_equal(m1, (r < 12).astype(float)) or m1.sum() >= (r < 12).sum() - 4
    m2 = fo.adaptive_threshold(img, fill_hole=True, dilation=1)
    assert m2.sum() > m1.sum() and np.all(m2 >= m1)
    m3 = fo.adaptive_threshold(img, fill_hole=True, dilation=1, erosion=1)
    assert np.all(m3 <= m2)
    # a constant image fails: zero mask; a stack thresholds slice by slice
    assert fo.adaptive_threshold(np.ones((8, 8))).max() == 0
    st = fo.adaptive_threshold(np.stack([img, np.ones_like(img)]))
    assert np.array_equal(st[0], m0) and st[1].max() == 0


def test_cal_frac_oracle_identities():
    rng = np.random.default_rng(3)
    H = W = 48
    yy, xx = np.mgrid[:H, :W]
    blob = (np.hypot(yy - 24, xx - 24) < 15).astype(np.float32)
    img = np.stack([blob * rng.uniform(0.5, 1.0, (H, W)), blob * rng.uniform(0.1, 0.3, (H, W))]).astype(np.float32)
    img += 1e-3 * rng.random(img.shape).astype(np.float32)
    r = fo.cal_frac(img, scale_range=[])
    m = r["img_mask"]
    assert m.dtype == np.float64 and m[24, 24] == 1 and m[0, 0] == 0
    # without clipping the fractions add up to ratio inside the mask
    ratio = r["img_sum"] / r["pix_max"]
    inside = m > 0
    np.testing.assert_allclose(r["frac"].sum(0)[inside], ratio[inside], rtol=1e-6)
    assert np.all(r["frac"][:, ~inside] == 0)
    # every threshold failing: the mask falls back to 1 and the float32 types stay
    flat = np.ones((2, 8, 8), np.float32)
    rf = fo.cal_frac(flat)
    assert rf["img_mask"] == 1 and rf["img_sum"].dtype == np.float32
