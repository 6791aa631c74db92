"""Writes model_crop_fd.npz: FP64 central differences (step 1e-4) of the Poisson loss of
the attenuated model on the 32^3 crop case (tests/helpers.model_crop_case) at 64 seeded
voxels, from the oracle's FP64 model (oracle.model_forward_f64).

    python tests/golden/make_model_crop_fd.py      (about 3 minutes on 8 cores)

The inputs are regenerated from their seeds by the tests; the file holds the voxel
indices, the differences, the loss at the evaluation point and the step.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import helpers  # noqa: E402
from helpers import oracle  # noqa: E402

N_VOXELS = 64
SEED = 20260102


def loss_at(case, C):
    Pf = oracle.model_forward_f64(C, case["angles"], case["coef_x"], case["coef_xrf"], case["em"], case["cone"])
    return helpers.poisson_loss_f64(Pf, case["I"])


def voxel_indices(shape, n=N_VOXELS, seed=SEED):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, s, n) for s in shape], axis=1).astype(np.int32)


def central_difference(case, ix, step=helpers.MODEL_CROP_FD_STEP):
    C = case["C_eval"].astype(np.float64)
    Cp, Cm = C.copy(), C.copy()
    Cp[tuple(ix)] += step
    Cm[tuple(ix)] -= step
    return (loss_at(case, Cp) - loss_at(case, Cm)) / (2 * step)


def main():
    case = helpers.model_crop_case()
    idx = voxel_indices(case["C_eval"].shape)
    t0 = time.time()
    fd = np.array([central_difference(case, ix) for ix in idx])
    L0 = loss_at(case, case["C_eval"].astype(np.float64))
    out = os.path.join(HERE, "model_crop_fd.npz")
    np.savez(out, idx=idx, fd=fd, loss=np.float64(L0), step=np.float64(helpers.MODEL_CROP_FD_STEP))
    print(f"{out}: {len(fd)} differences in {time.time() - t0:.0f}s, loss {L0:.6e}, |fd| max {np.abs(fd).max():.3e}")


if __name__ == "__main__":
    main()
