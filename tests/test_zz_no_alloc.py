"""fl_mlem_batch_noalloc (static device scratch, no allocation) against fl_mlem_batch
(caller workspace) through the C ABI: two calls in a row, the second on a smaller
problem, each equal to the workspace entry point bit for bit; a problem beyond the
static scratch is refused."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import RTOL, oracle, rel_l2

pytestmark = pytest.mark.gpu


def _case(A, B, R, H, W, seed):
    rng = np.random.default_rng(seed)
    atten = rng.uniform(0.3, 1.0, (A, B, H, W)).astype(np.float32)
    C_true = rng.uniform(0, 1, (R, B, H, W)).astype(np.float32)
    em = rng.uniform(0.5, 2.0, (A, R)).astype(np.float32)
    theta = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    I = oracle.forward_batch_f32(atten, C_true, em, theta)
    C0 = np.ones((R, B, H, W), np.float32)
    return [torch.from_numpy(x).cuda() for x in (C0, atten, em, theta, I)]


def _run(lib, name, C0, atten, em, theta, I, n_iter, beta, ws=None):
    from fl_correction_amd import _lib
    A, B, H, W = atten.shape
    R = C0.shape[0]
    out = torch.full_like(C0, -7.0)
    st = _lib.stream_of(C0.device)
    args = [_lib.ptr(t) for t in (C0, atten, em, theta, I, out)] + [A, B, R, H, W, n_iter, beta, 0.01]
    if ws is not None:
        args += [_lib.ptr(ws), ws.numel() * 4]
    rc = getattr(lib, name)(*args, st)
    torch.cuda.synchronize()
    return rc, out


def test_noalloc_matches_workspace_entry_point_twice():
    from fl_correction_amd import _lib
    lib = _lib.load()
    for (A, B, R, H, W, beta, seed) in ((24, 5, 2, 70, 90, 1e-3, 1), (9, 2, 1, 33, 40, 0.0, 2)):
        C0, atten, em, theta, I = _case(A, B, R, H, W, seed)
        print(f"noalloc A={A} B={B} R={R} H={H} W={W} beta={beta}", flush=True)
        nbytes = lib.fl_mlem_workspace_bytes(A, B, R, H, W)
        ws = torch.empty((nbytes + 3) // 4, dtype=torch.float32, device="cuda")
        rc1, ref = _run(lib, "fl_mlem_batch", C0, atten, em, theta, I, 6, beta, ws)
        rc2, got = _run(lib, "fl_mlem_batch_noalloc", C0, atten, em, theta, I, 6, beta)
        assert rc1 == 0 and rc2 == 0, lib.fl_last_error()
        assert torch.equal(got, ref)
        host = [t.cpu().numpy() for t in (C0, atten, em, theta, I)]
        assert rel_l2(got.cpu().numpy(), oracle.mlem_batch_f32(*host, 6, beta, 0.01)) <= RTOL


def test_noalloc_refuses_a_problem_beyond_the_scratch():
    from fl_correction_amd import _lib
    lib = _lib.load()
    A, B, R, H, W = 8, 64, 1, 512, 512  # 2 x 64 MiB of estimates alone
    assert lib.fl_mlem_workspace_bytes(A, B, R, H, W) > (64 << 20)
    x = torch.zeros(1, device="cuda")
    p = _lib.ptr(x)
    rc = lib.fl_mlem_batch_noalloc(p, p, p, p, p, p, A, B, R, H, W, 1, 0.0, 0.01, _lib.stream_of(x.device))
    assert rc == 4  # FL_ERR_WORKSPACE
    assert b"static scratch" in lib.fl_last_error()
