"""Batched MLEM (fl_mlem_multi_*): many uneven problems in one persistent launch
per pass, two iterations per replay of a captured hipGraph.

Bars: every problem's result equals the FP32 oracle to 1e-5 relative L2 and
the per-problem path (fl_mlem_batch) to 1e-6 with the reference layout (the
same kernels) or 1e-5 with the slice-major layout (problems packed along the
slice axis, the slice-major back-projector: another summation order); graph
replays equal eager runs bit for bit (the kernels are deterministic).
"""
import numpy as np
import pytest
import torch

from helpers import RTOL, oracle, rel_l2

pytestmark = pytest.mark.gpu

# (A, B, H, W): uneven, ragged against the tile sizes (64-column strips,
# 31 x 63 and 15 x 63 tiles), one single-slice problem, one tiny one
SHAPES = [(30, 3, 64, 64), (12, 2, 40, 70), (9, 1, 97, 33), (24, 4, 32, 32), (5, 2, 130, 66), (7, 1, 8, 8)]


def _problem(A, B, H, W, seed):
    rng = np.random.default_rng(seed)
    atten = rng.uniform(0.3, 1.0, (A, B, H, W)).astype(np.float32)
    em = rng.uniform(0.5, 2.0, A).astype(np.float32)
    theta = rng.uniform(0, np.pi, A).astype(np.float32)
    C_true = rng.uniform(0.0, 1.0, (1, B, H, W)).astype(np.float32)
    I = oracle.forward_batch_f32(atten, C_true, em[:, None], theta)
    C0 = np.ones((B, H, W), np.float32)
    return atten, em, theta, I, C0


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.fixture(scope="module")
def data():
    return [_problem(*s, seed=100 + i) for i, s in enumerate(SHAPES)]


LAYOUTS = ["reference", "slice_major"]


def _batch(data, shapes=None, beta=0.0, layout="auto", theta_of=None):
    """theta_of[i] = j: problem i takes problem j's theta tensor (the same array)."""
    from fl_correction_amd.batch import BatchMLEM, Problem
    probs = [Problem(_dev(a), _dev(e), _dev(t), _dev(I), _dev(C0)) for a, e, t, I, C0 in data]
    if theta_of is not None:
        for i, j in enumerate(theta_of):
            probs[i].theta = probs[j].theta
    run = BatchMLEM(probs, beta=beta, delta=0.01, shapes=shapes, layout=layout)
    if layout != "auto":
        assert run.layout == layout
    return run, probs


def _single(d, n_iter, beta):
    from fl_correction_amd.cuda_lib import atten_cuda
    a, e, t, I, C0 = d
    out = atten_cuda.mlem_cuda_batch(_dev(C0[None]), _dev(a), _dev(e[:, None]), _dev(t), _dev(I), n_iter, beta, 0.01)
    return out[0].cpu().numpy()


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("beta", [0.0, 1e-3])
@pytest.mark.parametrize("n_iter", [3, 4])
def test_batch_matches_single_and_oracle(data, beta, n_iter, layout):
    run, probs = _batch(data, beta=beta, layout=layout)
    print(f"batch {layout} beta={beta} n_iter={n_iter}: {len(probs)} problems", flush=True)
    run.run(n_iter, graph=True)
    torch.cuda.synchronize()
    for i, (d, p) in enumerate(zip(data, probs)):
        got = p.C.cpu().numpy()
        ref1 = _single(d, n_iter, beta)
        a, e, t, I, C0 = d
        ref = oracle.mlem_batch_f32(C0[None], a, e[:, None], t, I, n_iter, beta, 0.01)[0]
        e1, e32 = rel_l2(got, ref1), rel_l2(got, ref)
        print(f"  problem {i} {SHAPES[i]}: vs single {e1:.1e}, vs f32 oracle {e32:.1e}")
        assert e1 <= (1e-6 if layout == "reference" else RTOL)
        assert e32 <= RTOL
    run.close()


@pytest.mark.parametrize("beta", [0.0, 1e-3])
@pytest.mark.parametrize("n_iter", [1, 4])
def test_packed_lines_match_oracle(beta, n_iter):
    """Problems sharing a theta tensor and (A, H, W) -- emission lines of one
    sample, with different slice counts and emission factors -- are packed
    along the slice axis in the slice-major layout: each line against the
    oracle and against the reference layout; a lone problem beside them, and
    packs of 67 and 131 slices (two and four slices per lane, padded to 68 and
    132)."""
    groups = [(24, [3, 5, 1], 40, 52), (15, [2, 2], 33, 70), (10, [4], 20, 24), (6, [33, 34], 20, 24),
              (5, [70, 61], 18, 22)]
    data, theta_of = [], []
    for g, (A, Bs, H, W) in enumerate(groups):
        first = len(data)
        for k, B in enumerate(Bs):
            a, e, t, I, C0 = _problem(A, B, H, W, seed=500 + 10 * g + k)
            if k > 0:  # the line's data made with the sample's angles
                t = data[first][2]
                C_true = np.random.default_rng(600 + 10 * g + k).uniform(0.0, 1.0, (1, B, H, W)).astype(np.float32)
                I = oracle.forward_batch_f32(a, C_true, e[:, None], t)
            data.append((a, e, t, I, C0))
            theta_of.append(first)
    print(f"packed lines {groups} beta={beta} n_iter={n_iter}", flush=True)
    sm, sm_probs = _batch(data, beta=beta, layout="slice_major", theta_of=theta_of)
    ref, ref_probs = _batch(data, beta=beta, layout="reference", theta_of=theta_of)
    sm.run(n_iter, graph=True)
    ref.run(n_iter, graph=True)
    torch.cuda.synchronize()
    for i, (d, p, q) in enumerate(zip(data, sm_probs, ref_probs)):
        a, e, t, I, C0 = d
        want = oracle.mlem_batch_f32(C0[None], a, e[:, None], t, I, n_iter, beta, 0.01)[0]
        got = p.C.cpu().numpy()
        e32, er = rel_l2(got, want), rel_l2(got, q.C.cpu().numpy())
        print(f"  line {i} {a.shape}: vs f32 oracle {e32:.1e}, vs reference layout {er:.1e}")
        assert e32 <= RTOL and er <= RTOL
    sm.close()
    ref.close()


def test_samples_with_equal_angles_share_a_pack():
    """Samples whose angles are equal value for value, in separate theta tensors,
    are planned as one pack (merged_thetas): every line against the oracle and
    against the reference layout, which plans each problem on its own."""
    A, H, W = 12, 30, 44
    theta = np.random.default_rng(700).uniform(0, np.pi, A).astype(np.float32)
    data, theta_of, sample_of = [], [], []
    for g, Bs in enumerate([[5, 3], [7], [4, 6]]):
        first = len(data)
        for k, B in enumerate(Bs):
            a, e, _, _, C0 = _problem(A, B, H, W, seed=710 + 10 * g + k)
            C_true = np.random.default_rng(720 + 10 * g + k).uniform(0.0, 1.0, (1, B, H, W)).astype(np.float32)
            I = oracle.forward_batch_f32(a, C_true, e[:, None], theta)
            data.append((a, e, theta.copy(), I, C0))
            theta_of.append(first)
            sample_of.append(g)
    print(f"equal angles: samples of {[3, 1, 2]} lines, A={A} {H}x{W}", flush=True)
    sm, sm_probs = _batch(data, layout="slice_major", theta_of=theta_of)
    ref, ref_probs = _batch(data, layout="reference", theta_of=theta_of)
    # three samples, three theta tensors, one pack
    assert len({p.theta.data_ptr() for p in sm_probs}) == 3
    assert len({t.data_ptr() for t in sm._thetas}) == 1
    sm.run(4, graph=True)
    ref.run(4, graph=True)
    torch.cuda.synchronize()
    for i, (d, p, q) in enumerate(zip(data, sm_probs, ref_probs)):
        a, e, t, I, C0 = d
        want = oracle.mlem_batch_f32(C0[None], a, e[:, None], t, I, 4, 0.0, 0.01)[0]
        got = p.C.cpu().numpy()
        e32, er = rel_l2(got, want), rel_l2(got, q.C.cpu().numpy())
        print(f"  line {i} (sample {sample_of[i]}) {a.shape}: vs f32 oracle {e32:.1e}, vs reference layout {er:.1e}")
        assert e32 <= RTOL and er <= RTOL
    sm.close()
    ref.close()


def test_auto_layout_takes_slice_major(data):
    run, _ = _batch(data[:2])
    assert run.layout == "slice_major"
    run.close()


def test_tile_shapes_agree(data):
    runs = []
    for shp in [(8, 31), (4, 31), (8, 15), (4, 15)]:
        run, probs = _batch(data, shapes=[shp] * len(data), layout="reference")
        run.run(4, graph=True)
        torch.cuda.synchronize()
        runs.append([p.C.cpu().numpy() for p in probs])
        run.close()
    for k in range(1, 4):
        for i in range(len(data)):
            assert rel_l2(runs[k][i], runs[0][i]) <= 1e-6, (k, i)
    # forward angles per workgroup change no arithmetic
    for i in range(len(data)):
        assert np.array_equal(runs[1][i], runs[0][i])


@pytest.mark.parametrize("layout", LAYOUTS)
def test_graph_replays_equal_eager(data, layout):
    g_run, g_probs = _batch(data, beta=1e-3, layout=layout)
    e_run, e_probs = _batch(data, beta=1e-3, layout=layout)
    side = torch.ones(1 << 20, device="cuda")
    for rep in range(3):
        g_run.run(2, graph=True)  # one replay each time
        side.mul_(1.0001)          # eager work of the program between replays
        torch.cuda.synchronize()
        e_run.run(2, graph=False)
        torch.cuda.synchronize()
        for gp, ep in zip(g_probs, e_probs):
            assert torch.equal(gp.C, ep.C), f"replay {rep}"
    g_run.close()
    e_run.close()


def test_batch_errors(data):
    from fl_correction_amd import _lib
    from fl_correction_amd.batch import BatchMLEM, Problem
    a, e, t, I, C0 = data[0]
    p = Problem(_dev(a), _dev(e), _dev(t), _dev(I), _dev(C0))
    with pytest.raises(_lib.FLError, match="tile shape"):
        BatchMLEM([p], shapes=[(5, 31)])
    q = Problem(p.atten, p.em_cs, p.theta, p.I, p.C)  # same estimate twice
    with pytest.raises(_lib.FLError, match="overlap"):
        BatchMLEM([p, q])
    with pytest.raises(ValueError):
        BatchMLEM([Problem(p.atten, p.em_cs, p.theta, p.I, p.C[:1])])


def test_autotuner_picks_and_caches(tmp_path):
    from fl_correction_amd.batch import Autotuner, BWD_ROWS, FWD_ANGLES
    path = tmp_path / "tune.json"
    tu = Autotuner(path=str(path), reps=1, tune_angles=24)
    fa, br = tu.shapes(24, 2, 40, 40)
    assert fa in FWD_ANGLES and br in BWD_ROWS
    assert len(tu.timings) == len(FWD_ANGLES) * len(BWD_ROWS)
    # the slice-major layout tunes the forward angles only
    fs, bs = tu.shapes(24, 2, 40, 40, layout="slice_major")
    assert fs in FWD_ANGLES and bs == BWD_ROWS[0]
    assert len(tu.timings) == len(FWD_ANGLES) * (len(BWD_ROWS) + 1)
    print("autotune 64x64 bucket:", tu.timings)
    again = Autotuner(path=str(path))
    assert again.shapes(30, 1, 33, 64) == (fa, br)  # same bucket, from the cache
    assert again.shapes(30, 1, 33, 64, layout="slice_major") == (fs, bs)
    assert again.timings == {}


@pytest.mark.parametrize("layout", LAYOUTS)
def test_batch_edge_sizes_and_iteration_counts(layout):
    """Minimum shapes (H = W = 2, one angle, one slice) beside a normal one;
    n_iter = 0 leaves C unchanged, n_iter = 1 runs the odd-iteration path."""
    shapes = [(1, 1, 2, 2), (2, 1, 3, 5), (3, 2, 70, 2), (1, 3, 2, 65), (5, 2, 33, 33)]
    data = [_problem(*s, seed=300 + i) for i, s in enumerate(shapes)]
    print(f"batch edge sizes {shapes} {layout}", flush=True)
    run, probs = _batch(data, layout=layout)
    before = [p.C.clone() for p in probs]
    run.run(0, graph=True)
    torch.cuda.synchronize()
    for p, b in zip(probs, before):
        assert torch.equal(p.C, b)
    run.run(1, graph=True)
    torch.cuda.synchronize()
    for i, (d, p) in enumerate(zip(data, probs)):
        a, e, t, I, C0 = d
        ref = oracle.mlem_batch_f32(C0[None], a, e[:, None], t, I, 1, 0.0, 0.01)[0]
        assert rel_l2(p.C.cpu().numpy(), ref) <= RTOL, shapes[i]
    run.close()


@pytest.mark.parametrize("layout", LAYOUTS)
def test_batch_spread_angles(layout):
    """Grouped launches with angles in random order over [0, 2 pi) on images
    wider than two 64-column strips (footprint merges of far-apart angles)."""
    shapes = [(12, 1, 200, 200), (9, 2, 150, 220)]
    data = []
    for i, (A, B, H, W) in enumerate(shapes):
        atten, em, theta, I, C0 = _problem(A, B, H, W, seed=400 + i)
        theta = np.random.default_rng(410 + i).uniform(0, 2 * np.pi, A).astype(np.float32)
        C_true = np.random.default_rng(420 + i).uniform(0.0, 1.0, (1, B, H, W)).astype(np.float32)
        I = oracle.forward_batch_f32(atten, C_true, em[:, None], theta)
        data.append((atten, em, theta, I, C0))
    print(f"batch spread angles {shapes} {layout}", flush=True)
    run, probs = _batch(data, layout=layout)
    run.run(3, graph=True)
    torch.cuda.synchronize()
    for i, (d, p) in enumerate(zip(data, probs)):
        a, e, t, I, C0 = d
        ref = oracle.mlem_batch_f32(C0[None], a, e[:, None], t, I, 3, 0.0, 0.01)[0]
        assert rel_l2(p.C.cpu().numpy(), ref) <= RTOL, shapes[i]
    run.close()
