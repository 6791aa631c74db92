"""cal_frac on the GPU (fl_frac_* through fl_correction_amd.frac) against the oracle.

Bars: the multi-Otsu threshold and the masks (fill, dilation, erosion) equal the
oracle's exactly; img_sum, frac, pix_median and pix_max equal exactly (the same
IEEE operations in the same order); float32 and float64 stacks, 2-D and 3-D
sums, widths around the 32-pixel words, failing thresholds.
"""
import numpy as np
import pytest
import torch

import helpers  # noqa: F401  (puts oracle/ on sys.path)
import frac_oracle as fo

pytestmark = pytest.mark.gpu


def _images(seed, H, W, dtype):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[:H, :W]
    r = np.hypot(yy - H / 2, xx - W / 2) / max(H, W)
    base = np.where(r < 0.35, 1.0, 0.05) + np.where((r > 0.12) & (r < 0.16), -0.9, 0.0)
    img = base * rng.uniform(0.5, 1.5, (H, W)) + 0.02 * rng.random((H, W))
    return img.astype(dtype)


SHAPES = [(40, 50), (33, 31), (64, 32), (17, 97), (1, 40), (128, 129)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", SHAPES)
def test_adaptive_threshold_equals_oracle(dtype, shape):
    from fl_correction_amd import frac
    H, W = shape
    imgs = np.stack([_images(s, H, W, dtype) for s in range(3)])
    imgs[2] = np.random.default_rng(9).gamma(2.0, 1.0, (H, W)).astype(dtype)
    for fill, dil, ero in ((False, 0, 0), (True, 0, 0), (True, 2, 0), (True, 1, 1), (False, 0, 2)):
        print(f"adaptive_threshold {shape} {np.dtype(dtype).name} fill={fill} dilation={dil} erosion={ero}", flush=True)
        got = frac.adaptive_threshold(imgs, fill_hole=fill, dilation=dil, erosion=ero)
        ref = fo.adaptive_threshold(imgs, fill_hole=fill, dilation=dil, erosion=ero)
        assert got.dtype == np.float64
        assert np.array_equal(got, ref)
        assert np.array_equal(frac.adaptive_threshold(imgs[0], fill, dil, ero), ref[0])


def test_thresholds_equal_oracle():
    from fl_correction_amd import _lib, frac
    rng = np.random.default_rng(4)
    cases = [rng.gamma(2.0, 1.0, (50, 60)), rng.random((40, 40)) ** 4,
             np.concatenate([rng.normal(m, 0.2, 900) for m in (0, 3, 7)]).reshape(45, 60),
             np.array([[1.0, 1.0, 5.0, 9.0] * 8] * 8)]          # three distinct values
    for dtype in (np.float32, np.float64):
        for img in cases:
            img = img.astype(dtype)
            t = torch.from_numpy(img[None].copy()).cuda()
            _, status, thr = frac._adaptive_mask(t, False, 0, 0)
            ref, _ = fo.multiotsu_threshold(img)
            assert int(status[0]) == 0 and thr.cpu().numpy()[0] == ref


def test_failing_thresholds_zero_the_mask():
    from fl_correction_amd import frac
    imgs = np.stack([np.ones((20, 20)), np.tile([0.0, 1.0], (20, 10)), _images(1, 20, 20, np.float64)])
    bad = imgs.copy()
    bad[2, 3, 4] = np.nan
    for stack in (imgs, bad):
        got = frac.adaptive_threshold(stack, fill_hole=True, dilation=2)
        assert np.array_equal(got, fo.adaptive_threshold(stack, fill_hole=True, dilation=2))
    assert frac.adaptive_threshold(bad)[2].max() == 0 and frac.adaptive_threshold(imgs)[2].max() == 1


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout", ["stack", "volume", "image"])
def test_cal_frac_equals_oracle(dtype, layout):
    from fl_correction_amd import frac
    rng = np.random.default_rng(11)
    if layout == "stack":
        img = np.stack([_images(s, 48, 70, dtype) * w for s, w in ((1, 1.0), (2, 0.3), (3, 0.05))])
    elif layout == "volume":
        img = np.stack([np.stack([_images(10 * e + b, 40, 36, dtype) for b in range(4)]) for e in range(2)])
    else:
        img = _images(5, 30, 44, dtype)
    img = img + (1e-3 * rng.random(img.shape)).astype(dtype)
    print(f"cal_frac {layout} {np.dtype(dtype).name} {img.shape}", flush=True)
    got = frac.cal_frac(img)
    ref = fo.cal_frac(img)
    for k in ("img_sum", "frac", "img_mask"):
        assert np.asarray(got[k]).dtype == np.asarray(ref[k]).dtype, k
        assert np.array_equal(got[k], ref[k]), k
    assert got["pix_median"] == ref["pix_median"] and got["pix_max"] == ref["pix_max"]
    # keyword forms and torch tensors in, tensors out
    g2 = frac.cal_frac(torch.from_numpy(img).cuda(), scale_range=[], enable_scale=False)
    r2 = fo.cal_frac(img, scale_range=[], enable_scale=False)
    assert isinstance(g2["frac"], torch.Tensor) and np.array_equal(g2["frac"].cpu().numpy(), r2["frac"])


def test_cal_frac_mask_fallback_keeps_float32():
    from fl_correction_amd import frac
    img = np.stack([np.full((16, 16), 2.0, np.float32), np.full((16, 16), 1.0, np.float32)])
    img[0, 3, 3] = 5.0
    got, ref = frac.cal_frac(img), fo.cal_frac(img)
    assert ref["img_mask"] == 1 and got["img_mask"] == 1
    assert got["img_sum"].dtype == np.float32 and np.array_equal(got["img_sum"], ref["img_sum"])
    assert np.array_equal(got["frac"], ref["frac"])


def test_cal_frac_several_arrays_and_large_slice():
    from fl_correction_amd import frac
    a = [np.stack([_images(s + b, 64, 64, np.float32) for b in range(3)]) for s in (0, 7)]
    got = frac.cal_frac(*a)
    # pre_treat: stacked to float64 with NaN and inf zeroed
    ref = fo.cal_frac(np.nan_to_num(np.stack(a).astype(np.float64), nan=0.0, posinf=0.0, neginf=0.0))
    assert np.array_equal(got["frac"], ref["frac"]) and np.array_equal(got["img_mask"], ref["img_mask"])
    big = _images(3, 512, 512, np.float32)[None]
    print("adaptive_threshold 512x512", flush=True)
    assert np.array_equal(frac.adaptive_threshold(big, True, 2), fo.adaptive_threshold(big, True, 2))
