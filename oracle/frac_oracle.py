"""CPU oracle for cal_frac and its adaptive threshold -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, and only to check results.

A NumPy restatement of FL/core.py:159-211 (cal_frac) and FL/image_util.py:337-375
(adaptive_threshold), which the reference runs on numpy, scipy.ndimage and
skimage.filters.threshold_multiotsu.  numpy and scipy are importable here and
are called exactly as the reference calls them (np.histogram, np.sort,
np.median, ndimage.binary_fill_holes / binary_dilation / binary_erosion).
skimage is absent: ``multiotsu_threshold`` restates its published algorithm
(three classes, 256 bins: the normalised image histogram rounded to float32,
float32 cumulative zeroth and first moments, the between-class variance
sum_k S_k^2 / P_k of every split, the first maximum in split order).  That
restatement is pinned by ``multiotsu_brute_f64`` (the textbook between-class
variance sum_k w_k (mu_k - mu_T)^2 over every split, in FP64) and by
closed-form cases in tests/test_frac_oracle.py; its float32 tie-breaking is
"parity unpinned" against skimage itself.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

NBINS = 256


def histogram_centers(image):
    """skimage.exposure.histogram(image, 256, source_range='image') for a float
    image: np.histogram over [min, max], bin centres (e[:-1] + e[1:]) / 2."""
    hist, edges = np.histogram(image.reshape(-1), bins=NBINS)
    return hist, (edges[:-1] + edges[1:]) / 2.0


def _split_search_f32(prob):
    """First maximum over splits (i1, i2), 0 <= i1 < i2 <= nbins-2, of
    var(0, i1) + var(i1+1, i2) + var(i2+1, nbins-1) in float32, var(i, j) =
    S^2 / P over bins i..j (0 when P = 0)."""
    f32 = np.float32
    n = prob.size
    P = np.empty(n, f32)
    S = np.empty(n, f32)
    z, f = f32(prob[0]), f32(0.0) * f32(prob[0])
    P[0], S[0] = z, f
    for i in range(1, n):
        z = f32(z + prob[i])
        f = f32(f + f32(i) * prob[i])
        P[i], S[i] = z, f

    def ratio(ff, zz):  # S^2 / P, 0 where P = 0
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(zz > 0, ff * ff / np.where(zz > 0, zz, f32(1.0)), f32(0.0)).astype(f32)

    last = n - 1
    best, bi = f32(0.0), None
    for i1 in range(0, n - 2):
        v0 = ratio(S[i1:i1 + 1], P[i1:i1 + 1])[0]
        i2 = np.arange(i1 + 1, n - 1)
        # var(i1+1, i2) and var(i2+1, last) for each i2
        v1 = ratio(S[i2] - S[i1], P[i2] - P[i1])
        v2 = ratio(S[last] - S[i2], P[last] - P[i2])
        sig = ((v0 + v1).astype(f32) + v2).astype(f32)
        k = int(np.argmax(sig))
        if sig[k] > best:
            best, bi = sig[k], (i1, int(i2[k]))
    return bi


def multiotsu_threshold(image):
    """Lowest of the two three-class multi-Otsu thresholds of `image`
    (skimage.filters.threshold_multiotsu(image), classes=3, nbins=256), with the
    split indices.  Raises ValueError where skimage raises (fewer than three
    distinct histogram bins) or np.histogram raises (a non-finite value)."""
    hist, centers = histogram_centers(np.asarray(image))
    prob = (hist / hist.sum()).astype(np.float32)
    nz = np.count_nonzero(prob)
    if nz < 3:
        raise ValueError("fewer than three distinct histogram bins")
    if nz == 3:
        idx = tuple(np.flatnonzero(prob)[:-1])
    else:
        idx = _split_search_f32(prob)
        if idx is None:
            raise ValueError("no split with a positive between-class variance")
    return centers[idx[0]], idx


def multiotsu_brute_f64(prob):
    """Every split (i1, i2) of a histogram and its between-class variance
    sum_k w_k (mu_k - mu_T)^2 in FP64 (w_k the class weights, mu_k the class
    mean bin indices): the textbook definition, for pinning the search."""
    p = np.asarray(prob, np.float64)
    n = p.size
    idx = np.arange(n, dtype=np.float64)
    muT = float((p * idx).sum() / p.sum())
    out = {}
    for i1 in range(0, n - 2):
        for i2 in range(i1 + 1, n - 1):
            s = 0.0
            for a, b in ((0, i1), (i1 + 1, i2), (i2 + 1, n - 1)):
                w = p[a:b + 1].sum()
                if w > 0:
                    mu = (p[a:b + 1] * idx[a:b + 1]).sum() / w
                    s += w * (mu - muT) ** 2
            out[(i1, i2)] = s
    return out


def adaptive_threshold(img, fill_hole=False, dilation=0, erosion=0):
    """FL/image_util.py:337-375 on a 2-D image or a stack of 2-D images: mask =
    0 where the image lies below its lowest multi-Otsu threshold (np.digitize
    region 0), then binary_fill_holes with np.ones((5, 5)), binary_dilation and
    binary_erosion with the 4-neighbour cross; a failed threshold zeroes the
    image's mask.  float64 0/1, as the reference's np.ones mask."""
    img = np.asarray(img)
    struct = ndimage.generate_binary_structure(2, 1)

    def one(image):
        m = np.ones(image.shape)
        try:
            t, _ = multiotsu_threshold(image)
        except ValueError:
            return np.zeros(image.shape)
        regions = np.digitize(image, bins=[t])
        m[regions < 1] = 0
        if fill_hole:
            m = ndimage.binary_fill_holes(m, np.ones((5, 5)))
        if dilation > 0:
            m = ndimage.binary_dilation(m, structure=struct, iterations=dilation)
        if erosion > 0:
            m = ndimage.binary_erosion(m, structure=struct, iterations=erosion)
        return np.asarray(m, np.float64)

    if img.ndim == 3:
        return np.stack([one(img[i]) for i in range(img.shape[0])])
    if img.ndim == 2:
        return one(img)
    raise ValueError("image shape not recognized")


def cal_frac(img, scale_range=(0.02, 0.98), enable_scale=True, scale_limit=0.95):
    """FL/core.py:159-211 for one stack img (n_elem, H, W) or (n_elem, B, H, W)
    (a 2-D image is one element), in the reference's numpy types: the element
    sum and the threshold in img's type, float64 after the float64 mask (the
    integer 1 when every threshold fails keeps img's type)."""
    img = np.array(img, copy=True)
    if img.ndim == 2:
        img = img[None]
    n = img.shape[0]
    frac = np.zeros(img.shape)
    img_sum = np.sum(img, axis=0) + 1e-12
    mask = adaptive_threshold(img_sum, fill_hole=True, dilation=2)
    if np.max(mask) == 0:
        mask = 1
    img_sum = img_sum * mask
    sum_sort = np.sort(img_sum[img_sum > 0])
    l = len(sum_sort)
    pix_max = np.median(sum_sort[int(l * scale_limit)])
    ratio = img_sum / pix_max if enable_scale else 1
    pix_median = np.median(sum_sort)
    for i in range(n):
        with np.errstate(divide="ignore", invalid="ignore"):
            t = img[i] / img_sum
        t = np.where(np.isfinite(t) & (t >= 0), t, 0).astype(t.dtype)  # rm_abnormal (image_util.py:396-401)
        if len(scale_range):
            ts = np.sort(t[t > 0])
            m = len(ts)
            if m > 100:
                ts = ts[int(m * scale_range[0]):int(m * scale_range[1])]
                t[t >= ts[-1]] = ts[-1]
                t[t <= ts[0]] = ts[0]
        frac[i] = t * mask * ratio
    return {"img_sum": img_sum, "frac": frac, "pix_median": pix_median, "pix_max": pix_max, "img_mask": mask}
