"""Attenuation-map files: the reference's names and layout, float32 round trips (no GPU)."""
import os

import numpy as np
import pytest


def test_names_match_the_reference_patterns(tmp_path):
    from fl_correction_amd import file_io
    assert os.path.basename(file_io.attenuation_file(str(tmp_path), "Fe", 7)) == "atten_Fe_prj_0007.tiff"
    assert os.path.basename(file_io.attenuation_file(str(tmp_path), "Fe", 12, "fl")) == "atten_fl_Fe_prj_0012.tiff"
    assert os.path.basename(file_io.attenuation_file(str(tmp_path), "Cu", 0, "xray")) == \
        "atten_incident_xray_Cu_prj_0000.tiff"
    with pytest.raises(ValueError):
        file_io.attenuation_file(str(tmp_path), "Cu", 0, "other")


@pytest.mark.parametrize("shape", [(5, 7), (3, 6, 9), (1, 4, 4)])
def test_tiff_round_trip_is_exact(tmp_path, shape):
    from fl_correction_amd import file_io
    rng = np.random.default_rng(1)
    x = rng.uniform(-3, 3, shape).astype(np.float32)
    x.flat[0] = np.float32(1e-38)
    x.flat[-1] = np.float32(3.4e38)
    f = str(tmp_path / "m.tiff")
    file_io.write_tiff(f, x)
    y = file_io.read_tiff(f)
    assert y.dtype == np.float32
    assert y.shape == (shape if len(shape) == 3 and shape[0] > 1 else shape[-2:])
    assert np.array_equal(y.reshape(x.shape), x)


def test_stack_writer_and_readers(tmp_path):
    from fl_correction_amd import file_io
    rng = np.random.default_rng(2)
    K, A, B, H, W = 2, 3, 4, 5, 6
    atten = rng.uniform(0, 1, (K, A, B, H, W)).astype(np.float32)
    d = str(tmp_path / "Angle_prj")
    file_io.write_attenuation_stack(atten, ["Fe", "Cu"], d)
    assert sorted(os.listdir(d))[0] == "atten_Cu_prj_0000.tiff"
    back = file_io.read_attenuation_stack(A, ["Fe", "Cu"], d)
    assert np.array_equal(back, atten)
    one = file_io.read_attenuation_at_all_angle(np.zeros(A), d, "Cu")
    assert one.dtype == np.float64 and np.array_equal(one, atten[1].astype(np.float64))
    file_io.write_attenuation_fl("Fe", atten[0, 1], 1, d)
    file_io.write_attenuation_xray("Fe", atten[0, 2, 0], 2, d)
    assert np.array_equal(file_io.read_attenuation_fl(1, d, "Fe"), atten[0, 1])
    assert np.array_equal(file_io.read_attenuation_xray(2, d, "Fe"), atten[0, 2, 0])
