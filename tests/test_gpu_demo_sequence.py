"""The reference demo's call sequence (FL/example/demo.py:262-311) through the package
aliases: per-path attenuation files, the total-attenuation files, the maps read back,
MLEM on them.  The files must hold exactly the device-resident maps, and MLEM on the
maps read back must equal MLEM on the device maps bit for bit."""
import os
import sys

import numpy as np
import pytest
import torch

from helpers import RTOL, oracle, rel_l2

pytestmark = pytest.mark.gpu


def _alias_reference_names():
    import fl_correction_amd
    import fl_correction_amd.cuda_lib as cl
    sys.modules["FL"] = fl_correction_amd
    sys.modules["FL.cuda_lib"] = cl
    sys.modules["FL.cuda_lib.atten_cuda"] = cl.atten_cuda
    sys.modules["FL.cuda_lib.autograd"] = cl.autograd
    import FL
    return FL


def test_demo_sequence_round_trips_the_files(tmp_path):
    FL = _alias_reference_names()
    from fl_correction_amd import core
    rng = np.random.default_rng(2026)
    elem_type = ["Co", "Mn"]
    K, B, N, A = 2, 4, 28, 6
    frac = rng.uniform(0, 1, (K, B, N, N)).astype(np.float32)
    angle_list = np.linspace(0, 180, A, endpoint=False)
    em = {el: rng.uniform(0.5, 2.0, A) for el in elem_type}
    param = FL.make_param(elem_type, ["CoO", "MnO"], rng.uniform(20, 400, K), rng.uniform(20, 400, (K, K)),
                          rho=1.0, pix=1e-4, em_cs=em)
    detector_mask = FL.generate_detector_mask(20, 20, N).astype(np.float32)
    frac_4D = torch.tensor(frac, dtype=torch.float, device="cuda")
    fsave = str(tmp_path / "Angle_prj_bin1_1")
    print(f"demo sequence K={K} B={B} N={N} A={A}", flush=True)

    FL.cal_xrf_atten_at_angles(frac_4D, param, angle_list, detector_mask, detector_offset_angle=0,
                               position_det="r", dict_use_ref={}, fsave=fsave)
    FL.cal_xray_atten_at_angles(frac_4D, param, angle_list, {}, fsave=fsave)
    FL.cal_total_atten_at_angles(fsave, param)
    assert len(os.listdir(fsave)) == 3 * K * A

    resident = core.cal_atten_at_angles(frac_4D, param, angle_list, detector_mask).cpu().numpy()
    for k, elem in enumerate(elem_type):
        atten = FL.read_attenuation_at_all_angle(angle_list, fsave, elem, "all")
        assert atten.dtype == np.float64 and atten.shape == (A, B, N, N)
        assert np.array_equal(atten.astype(np.float32), resident[k]), elem

        # the maps against the oracle's composition of rotation, mask correlation and incident product
        for a in (0, A - 1):
            rot = oracle.rotate_torch(frac, np.float32(angle_list[a] / 180 * np.pi))
            p = FL.extract_single_value_param(param, a)
            from fl_correction_amd.atten_util import coef_xray, coef_xrf
            a_x = oracle.incident_f32(oracle.mu_xray(rot, coef_xray(p)))
            a_f = oracle.xrf_atten_from_mask(oracle.mu_xrf(rot, coef_xrf(p)), detector_mask)
            assert rel_l2(atten[a].astype(np.float32), (a_f * a_x)[k]) <= RTOL

        em_cs = FL.atten_util.extract_em_cs_elem(param, elem, use_ref=False)
        C_true = rng.uniform(0, 1, (1, B, N, N)).astype(np.float32)
        theta = (angle_list / 180 * np.pi).astype(np.float32)
        em_col = (np.asarray(em_cs)[:, None] * 1.0 * 1e-4).astype(np.float32)
        prjs = oracle.forward_batch_f32(resident[k], C_true, em_col, theta)
        C_init = np.ones((1, B, N, N), np.float32)
        rec = FL.torch_mlem_recon_batch(C_init, prjs, angle_list, atten, em_cs, 1.0, 1e-4, 8, 0.0, 0.01, "cuda")
        rec_dev = FL.mlem_cuda_batch(torch.tensor(C_init, device="cuda"), torch.tensor(resident[k], device="cuda"),
                                     torch.tensor(em_col, device="cuda"), torch.tensor(theta, device="cuda"),
                                     torch.tensor(prjs, device="cuda"), 8, 0.0, 0.01).cpu().numpy()
        assert np.array_equal(rec, rec_dev), elem
        ref = oracle.mlem_batch_f32(C_init, resident[k], em_col, theta, prjs, 8, 0.0, 0.01)
        assert rel_l2(rec, ref) <= RTOL


def test_drivers_write_nothing_without_fsave(tmp_path, monkeypatch):
    FL = _alias_reference_names()
    monkeypatch.chdir(tmp_path)
    rng = np.random.default_rng(5)
    frac = torch.tensor(rng.uniform(0, 1, (1, 2, 12, 12)), dtype=torch.float, device="cuda")
    param = FL.make_param(["Fe"], ["Fe2O3"], [100.0], [[200.0]])
    out = FL.cal_xray_atten_at_angles(frac, param, np.array([0.0, 45.0]), fsave=None)
    assert out.shape == (1, 2, 2, 12, 12) and os.listdir(tmp_path) == []
    FL.cal_xray_atten_at_angles(frac, param, np.array([0.0, 45.0]))  # the reference's default fsave
    assert sorted(os.listdir(tmp_path / "Angle_prj")) == ["atten_incident_xray_Fe_prj_0000.tiff",
                                                          "atten_incident_xray_Fe_prj_0001.tiff"]
