"""The joint reconstruction's loss and misfit on the host (no GPU): the
reference's poisson_loss_image (FL/ml_lib/ml_loss.py:6-8) restated directly."""
import numpy as np
import torch


def test_poisson_loss_matches_definition():
    from fl_correction_amd.joint import data_residual, poisson_loss
    rng = np.random.default_rng(0)
    Pf = rng.uniform(-0.5, 2.0, (3, 5, 7))
    I = rng.uniform(0.0, 2.0, (3, 5, 7))
    P = np.maximum(Pf, 1e-6)
    ref = np.mean(P - I * np.log(P))
    got = float(poisson_loss(torch.from_numpy(Pf), torch.from_numpy(I)))
    assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
    assert abs(data_residual(torch.from_numpy(Pf), torch.from_numpy(I))
               - np.linalg.norm(Pf - I) / np.linalg.norm(I)) < 1e-12


def test_poisson_loss_minimum_at_data():
    from fl_correction_amd.joint import poisson_loss
    I = torch.tensor([0.5, 1.0, 2.0], dtype=torch.float64)
    base = float(poisson_loss(I, I))
    for f in (0.9, 1.1):
        assert float(poisson_loss(I * f, I)) > base
