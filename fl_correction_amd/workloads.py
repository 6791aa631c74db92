"""Seeded definitions of the benchmark workloads, shared by the tests and the benchmarks.

Each configuration fixes the volume, the elements (emission lines), the ellipsoid
phantom, the angles, the detector cone and the seeds; the inputs follow from them:

  plain_64     64^3,  2 lines,  60 angles over 180 deg, 3x3 cone over 10 deg, 20 MLEM iterations
  loop_128     128^3, 3 lines, 180 angles over 180 deg, 5x5 cone over 15 deg, 50 iterations,
               attenuation recomputed every 5
  model_256    256^3, 4 lines, 360 angles over 180 deg, 5x5 cone over 15 deg (the differentiable model)
  joint_512    128 x 512 x 512, 5 lines, 720 angles over 360 deg, 7x7 cone over 20 deg

Physics mapping onto the reference's parameters: rho = 1, frac = concentration,
pix = 1e-4 cm (1 um), em_cs = 1 (so em = pix per line), cross-sections uniform in
[20, 400] cm^2/g at the incident energy and at every emission line, beam along the
rows, detector 'r'.  Projections carry Poisson noise at 1e4 peak counts.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import phantom
from .atten_util import make_param
from .detector_mask import Cone

PIX = 1.0e-4


@dataclass(frozen=True)
class Workload:
    name: str
    shape: tuple          # (B, H, W)
    K: int                # elements = emission lines
    n_ell: tuple          # ellipsoids per element (inclusive range)
    semi_axes: tuple      # semi-axis range in voxels
    A: int                # angles
    span: float           # angle span in degrees
    cone_n: int           # n x n cone directions
    cone_half: float      # cone half-angle in degrees
    seed: int
    q: int = 8            # slope denominator of the cone directions

    def phantom(self, box=None):
        """(K, B, H, W) float32 concentrations (numpy), or the given sub-box of them."""
        return phantom.ellipsoid_phantom(self.shape, self.K, self.n_ell, self.semi_axes, seed=self.seed, box=box)

    def phantom_device(self, device="cuda"):
        return phantom.ellipsoid_phantom_device(self.shape, self.K, self.n_ell, self.semi_axes, seed=self.seed,
                                                device=device)

    def angles(self):
        return phantom.angle_list(self.A, self.span)

    def cross_sections(self):
        return phantom.cross_sections(self.K, seed=self.seed + 1)

    def param(self):
        cs_x, cs_line = self.cross_sections()
        elems = [f"E{k}" for k in range(self.K)]
        return make_param(elems, [f"C{k}" for k in range(self.K)], cs_x, cs_line, rho=1.0, pix=PIX,
                          em_cs={e: np.ones(1) for e in elems})

    def coefs(self):
        """(coef_x (K,), coef_xrf (K, K)) float32: cs * rho * pix as the reference rounds them."""
        cs_x, cs_line = self.cross_sections()
        return (cs_x * PIX).astype(np.float32), (cs_line * PIX).astype(np.float32)

    def cone(self):
        return Cone(n=self.cone_n, half_angle_deg=self.cone_half, q=self.q)

    def model(self, angles=None, device="cuda", **kw):
        from .model import XRFModel
        return XRFModel(self.param(), self.angles() if angles is None else angles, self.cone(), device=device, **kw)


PLAIN_64 = Workload("plain_64", (64, 64, 64), 2, (6, 12), (4.0, 16.0), 60, 180.0, 3, 10.0, 20260000)
LOOP_128 = Workload("loop_128", (128, 128, 128), 3, (10, 30), (6.0, 32.0), 180, 180.0, 5, 15.0, 20260001)
MODEL_256 = Workload("model_256", (256, 256, 256), 4, (10, 30), (6.0, 32.0), 360, 180.0, 5, 15.0, 20260002)
JOINT_512 = Workload("joint_512", (128, 512, 512), 5, (10, 30), (6.0, 32.0), 720, 360.0, 7, 20.0, 20260004)


def noisy(Pf, seed, peak=1.0e4):
    """Poisson noise at `peak` counts on the largest value (numpy in, numpy out)."""
    return phantom.poisson_noise(Pf, peak=peak, seed=seed)


def batch_samples(n=48, seed=20260003):
    """The uneven batch: n samples with lateral size in {64, 96, 128, 160, 192}, 16-64
    slices, 2-4 elements and 90-360 angles in multiples of 30, 3x3 cone over 10 deg,
    ellipsoid phantoms as in plain_64.  Returns a list of Workload (one per sample)."""
    rng = phantom.rng_for(seed)
    out = []
    for i in range(n):
        L = int(rng.choice([64, 96, 128, 160, 192]))
        B = int(rng.integers(16, 65))
        K = int(rng.integers(2, 5))
        A = int(rng.choice(np.arange(90, 361, 30)))
        out.append(Workload(f"sample_{i}", (B, L, L), K, (6, 12), (4.0, 16.0), A, 180.0, 3, 10.0, seed + 1000 * (i + 1)))
    return out
