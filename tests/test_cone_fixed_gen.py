"""The compiled-in cone stencils (fl_correction_amd/csrc/cone_fixed.inc) are the
generator's output, and the generator's direction grids are the ones Cone
produces for n = 3 over 10 deg and n = 5 over 15 deg with slopes in 1/8."""
import importlib.util
import os

import numpy as np

from fl_correction_amd.detector_mask import Cone

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fl_correction_amd", "csrc")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_cone_fixed", os.path.join(CSRC, "gen_cone_fixed.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_include_matches_generator():
    with open(os.path.join(CSRC, "cone_fixed.inc")) as f:
        assert f.read() == _gen().render()


def test_grids_are_the_cones_directions():
    gen = _gen()
    for n, half, M in ((3, 10.0, 1), (5, 15.0, 2)):
        c = Cone(n=n, half_angle_deg=half, q=8)
        pz = [int(p) * (8 // int(q)) for p, q in zip(c.pz, c.qd)]
        py = [int(p) * (8 // int(q)) for p, q in zip(c.py, c.qd)]
        dirs, *_ = gen.geometry(M, 0)
        assert [d[0] for d in dirs] == pz and [d[1] for d in dirs] == py


def test_taps_are_the_bilinear_splats():
    """Each direction's taps over its Q steps carry weight Q (one per step) and
    sit on the two voxels either side of the ray in z and in y."""
    gen = _gen()
    for M in (1, 2):
        dirs, *_ = gen.geometry(M, 0)
        for pz, py in dirs:
            t = gen.taps(pz, py, 0)
            assert abs(sum(w for *_, w in t) - gen.Q) < 1e-12
            for dz, dy, pl, w in t:
                step = pl + 1
                assert abs(dz - step * pz / gen.Q) < 1 and abs(dy - step * py / gen.Q) < 1
            rt = gen.taps(pz, py, 1)
            assert sorted((-dz, -dy, gen.Q - 1 - pl, w) for dz, dy, pl, w in rt) == sorted(t)
