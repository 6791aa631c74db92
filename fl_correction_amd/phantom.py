"""Seeded synthetic inputs: ellipsoid concentration phantoms, mass-attenuation
cross-sections, angle lists and Poisson noise.

The reference ships no data (its demo reads /data/...); these generators make
inputs of the workload shapes: per element, a number of axis-aligned
ellipsoids with uniform concentrations in [0.1, 1.0] g/cm^3, summed where they
overlap, on a zero background, centred inside the inscribed cylinder.
All randomness comes from numpy's PCG64 with an explicit seed.
"""
from __future__ import annotations

import numpy as np


def rng_for(seed):
    return np.random.Generator(np.random.PCG64(seed))


def ellipsoid_phantom(shape, n_elem, n_ell=(6, 12), semi_axes=(4.0, 16.0), conc=(0.1, 1.0), seed=0, box=None):
    """Concentration maps (n_elem, B, H, W) float32.

    shape: (B, H, W) = (slices, rows, cols).  n_ell: inclusive range of the number
    of ellipsoids per element.  semi_axes: range of each semi-axis in voxels.
    box: ((b0, b1), (h0, h1), (w0, w1)) evaluates only that sub-box of the volume
    (the same draws, so the values equal the full phantom's crop).
    """
    B, H, W = shape
    rng = rng_for(seed)
    (b0, b1), (h0, h1), (w0, w1) = box if box is not None else ((0, B), (0, H), (0, W))
    out = np.zeros((n_elem, b1 - b0, h1 - h0, w1 - w0), np.float32)
    zz, yy, xx = np.meshgrid(np.arange(b0, b1, dtype=np.float32), np.arange(h0, h1, dtype=np.float32),
                             np.arange(w0, w1, dtype=np.float32), indexing="ij")
    cy, cx = (H - 1) / 2.0, (W - 1) / 2.0
    r_cyl = min(H, W) / 2.0
    for e in range(n_elem):
        n = int(rng.integers(n_ell[0], n_ell[1] + 1))
        for _ in range(n):
            ax = rng.uniform(semi_axes[0], semi_axes[1], size=3)
            # centre inside the inscribed cylinder (in-plane radius), any slice
            rad = r_cyl * 0.75 * np.sqrt(rng.uniform(0.0, 1.0))
            phi = rng.uniform(0.0, 2 * np.pi)
            c = (rng.uniform(0, B - 1), cy + rad * np.sin(phi), cx + rad * np.cos(phi))
            val = np.float32(rng.uniform(conc[0], conc[1]))
            inside = (((zz - c[0]) / ax[0]) ** 2 + ((yy - c[1]) / ax[1]) ** 2 + ((xx - c[2]) / ax[2]) ** 2) <= 1.0
            out[e][inside] += val
    return out


def cross_sections(n_elem, seed=0, lo=20.0, hi=400.0):
    """Mass attenuation cross-sections in cm^2/g, uniform in [lo, hi].

    Returns (cs_x, cs_line): cs_x[k] at the incident energy for compound k (n_elem,),
    cs_line[k, e] of compound k at the emission line of element e (n_elem, n_elem).
    """
    rng = rng_for(seed)
    cs_x = rng.uniform(lo, hi, size=n_elem)
    cs_line = rng.uniform(lo, hi, size=(n_elem, n_elem))
    return cs_x, cs_line


def angle_list(n_angle, span_deg=180.0):
    """Evenly spaced angles in degrees over [0, span)."""
    return np.linspace(0.0, span_deg, n_angle, endpoint=False)


def poisson_noise(Pf, peak=1.0e4, seed=0):
    """I = Poisson(Pf * s) / s with s = peak / max(Pf)."""
    rng = rng_for(seed)
    Pf = np.asarray(Pf, np.float64)
    m = float(Pf.max())
    s = peak / m if m > 0 else 1.0
    return (rng.poisson(np.clip(Pf, 0, None) * s) / s).astype(np.float32)


def ellipsoid_phantom_device(shape, n_elem, n_ell=(6, 12), semi_axes=(4.0, 16.0), conc=(0.1, 1.0), seed=0,
                             device="cuda"):
    """Same draws as ellipsoid_phantom, evaluated with torch on `device` (for large volumes)."""
    import torch
    B, H, W = shape
    rng = rng_for(seed)
    out = torch.zeros((n_elem, B, H, W), dtype=torch.float32, device=device)
    zz = torch.arange(B, dtype=torch.float64, device=device)[:, None, None]
    yy = torch.arange(H, dtype=torch.float64, device=device)[None, :, None]
    xx = torch.arange(W, dtype=torch.float64, device=device)[None, None, :]
    cy, cx = (H - 1) / 2.0, (W - 1) / 2.0
    r_cyl = min(H, W) / 2.0
    for e in range(n_elem):
        n = int(rng.integers(n_ell[0], n_ell[1] + 1))
        for _ in range(n):
            ax = rng.uniform(semi_axes[0], semi_axes[1], size=3)
            rad = r_cyl * 0.75 * np.sqrt(rng.uniform(0.0, 1.0))
            phi = rng.uniform(0.0, 2 * np.pi)
            c = (rng.uniform(0, B - 1), cy + rad * np.sin(phi), cx + rad * np.cos(phi))
            val = float(np.float32(rng.uniform(conc[0], conc[1])))
            inside = (((zz - c[0]) / ax[0]) ** 2 + ((yy - c[1]) / ax[1]) ** 2 + ((xx - c[2]) / ax[2]) ** 2) <= 1.0
            out[e] += inside.to(torch.float32) * val
    return out
