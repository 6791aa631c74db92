"""Host-side parts of the batched MLEM: workspace layout, argument checks that
run before any device call, and the autotuner's buckets and cache (no GPU)."""
import ctypes
import json
import os

import pytest

from fl_correction_amd import _lib


def _lib_or_skip():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    return _lib.load()


def _arr(shapes):
    arr = (_lib.MLEMProblem * len(shapes))()
    for i, (A, B, H, W) in enumerate(shapes):
        arr[i] = _lib.MLEMProblem(1, 1, 1, 1, 1 << 20 + i, A, B, H, W, 0, 0)
    return arr


def test_struct_layout_matches_header():
    # five pointers, four int64, two int32: 5*8 + 4*8 + 2*4
    assert ctypes.sizeof(_lib.MLEMProblem) == 80
    assert _lib.MLEMProblem.A.offset == 40 and _lib.MLEMProblem.fwd_angles.offset == 72


def test_workspace_bytes():
    lib = _lib_or_skip()
    shapes = [(30, 3, 64, 64), (12, 2, 40, 70), (9, 1, 97, 33)]
    arr = _arr(shapes)
    got = lib.fl_mlem_multi_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), len(shapes))
    al = lambda b: (b + 255) // 256 * 256
    per = sum(2 * al(4 * B * H * W) + al(4 * A * B * W) for A, B, H, W in shapes)
    # tables: 2 forward + 3 back-projector job tables, a copy table, 3 tile prefix arrays
    assert got > per and got - per < 5 * al(200 * len(shapes)) + al(32 * len(shapes)) + 3 * al(4 * (len(shapes) + 1)) + 1
    assert got % 256 == 0
    assert lib.fl_mlem_multi_workspace_bytes(None, 3) == 0
    assert lib.fl_mlem_multi_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), 0) == 0


def test_workspace_bytes_slice_major_packs():
    """Problems with the same theta pointer and (A, H, W) share one pack: its
    slice-major attenuation, ratio, sensitivity, two estimates and scratch are
    sized for the summed slice count; a pack past H*W*sum(B) = 2^29 makes the
    mode inapplicable (0)."""
    lib = _lib_or_skip()
    al = lambda b: (b + 255) // 256 * 256

    def sizes(shapes, thetas):
        arr = _arr(shapes)
        for i, t in enumerate(thetas):
            arr[i].theta = t
        p = ctypes.cast(arr, ctypes.c_void_p)
        return lib.fl_mlem_multi_workspace_bytes(p, len(shapes)), lib.fl_mlem_multi_workspace_bytes_sm(p, len(shapes))

    def pack_bytes(A, B, H, W):
        return (al(4 * A * H * W * B) + al(4 * A * W * B) + 3 * al(4 * B * H * W)
                + al(lib.fl_backproject_sm_scratch_bytes(A, B, H, W)))

    shapes = [(30, 3, 64, 64), (30, 5, 64, 64), (12, 2, 40, 70)]
    dense, packed = sizes(shapes, [8, 8, 16])  # lines 0 and 1 share theta
    _, apart = sizes(shapes, [8, 24, 16])
    per = pack_bytes(30, 8, 64, 64) + pack_bytes(12, 2, 40, 70)
    assert packed % 256 == 0 and packed > dense
    assert per < packed < per + 64 * 1024  # plus the tables
    per_apart = pack_bytes(30, 3, 64, 64) + pack_bytes(30, 5, 64, 64) + pack_bytes(12, 2, 40, 70)
    assert per_apart < apart < per_apart + 64 * 1024
    # two lines of 2^28 voxels each: fine apart, too large packed
    big = [(1, 4096, 256, 256), (1, 4096, 256, 256)]
    assert sizes(big, [8, 16])[1] > 0
    assert sizes(big, [8, 8])[1] == 0
    assert lib.fl_mlem_multi_workspace_bytes_sm(None, 3) == 0


def test_create_rejects_before_device_work():
    lib = _lib_or_skip()
    plan = ctypes.c_void_p()
    bad = _arr([(4, 1, 1, 8)])  # H < 2
    rc = lib.fl_mlem_multi_create(ctypes.cast(bad, ctypes.c_void_p), 1, 0.0, 0.01, None, 0, None, ctypes.byref(plan))
    assert rc == 1 and b"shape" in lib.fl_last_error()
    arr = _arr([(4, 1, 8, 8)])
    arr[0].fwd_angles = 6
    rc = lib.fl_mlem_multi_create(ctypes.cast(arr, ctypes.c_void_p), 1, 0.0, 0.01, None, 0, None, ctypes.byref(plan))
    assert rc == 1 and b"tile shape" in lib.fl_last_error()
    two = _arr([(4, 1, 8, 8), (4, 1, 8, 8)])
    two[1].C = two[0].C + 4 * 8  # overlaps the first estimate
    rc = lib.fl_mlem_multi_create(ctypes.cast(two, ctypes.c_void_p), 2, 0.0, 0.01, None, 0, None, ctypes.byref(plan))
    assert rc == 1 and b"overlap" in lib.fl_last_error()
    ok = _arr([(4, 1, 8, 8)])
    rc = lib.fl_mlem_multi_create(ctypes.cast(ok, ctypes.c_void_p), 1, 0.0, 0.01, None, 0, None, ctypes.byref(plan))
    assert rc == 4 and b"workspace" in lib.fl_last_error()
    assert not plan.value
    assert lib.fl_mlem_multi_run(None, 2, 1, None) == 1
    assert lib.fl_mlem_multi_destroy(None) == 0


def test_autotuner_buckets_and_cache(tmp_path):
    from fl_correction_amd.batch import Autotuner
    assert Autotuner.bucket(64, 64) == (64, 64)
    assert Autotuner.bucket(65, 97) == (96, 128)
    assert Autotuner.key("MI355X", 100, 160) == "MI355X|reference|128x160"
    assert Autotuner.key("MI355X", 100, 160, "slice_major") == "MI355X|slice_major|128x160"
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"X|64x64": [4, 15]}))
    tu = Autotuner(path=str(path))
    assert tu.table["X|64x64"] == [4, 15]
    # choices the library does not accept are tuned again, not used
    assert Autotuner.valid([4, 15]) and Autotuner.valid([8, 31])
    assert not Autotuner.valid([4, 16]) and not Autotuner.valid([5, 31]) and not Autotuner.valid(None)
    (tmp_path / "bad.json").write_text("{not json")
    assert Autotuner(path=str(tmp_path / "bad.json")).table == {}


def test_samples_with_equal_angles_share_a_pack():
    """merged_thetas: samples (problems sharing a theta tensor) with equal angles
    and (A, H, W) get one theta tensor, in order, up to PACK_MERGE_SLICES slices;
    different angles, image sizes or an exceeded cap keep their own tensor."""
    import torch
    from fl_correction_amd.batch import PACK_MERGE_SLICES, Problem, merged_thetas

    def sample(theta, lines, B, N):
        A = theta.numel()
        return [Problem(torch.zeros((A, B, N, N)), torch.ones(A), theta, torch.zeros((A, B, N)),
                        torch.ones((B, N, N))) for _ in range(lines)]

    ang = torch.linspace(0.0, 3.0, 30)
    t1, t2, t3, t5 = ang.clone(), ang.clone(), ang.clone(), ang.clone()
    t4 = ang.clone() + 0.01                 # other angles
    probs = (sample(t1, 2, 64, 16) + sample(t2, 2, 64, 16)  # 256 slices: one pack
             + sample(t3, 1, 40, 16)                        # would exceed the cap: its own
             + sample(t4, 2, 8, 16)                         # other angles: its own
             + sample(t5, 1, 8, 24))                        # other image size: its own
    assert PACK_MERGE_SLICES == 256
    got = [t.data_ptr() for t in merged_thetas(probs)]
    want = [t1.data_ptr()] * 4 + [t3.data_ptr()] + [t4.data_ptr()] * 2 + [t5.data_ptr()]
    assert got == want
    # a sample of one line with the same angles as t3 joins t3's pack
    t6 = ang.clone()
    got = [t.data_ptr() for t in merged_thetas(probs + sample(t6, 1, 8, 16))]
    assert got[-1] == t3.data_ptr()
