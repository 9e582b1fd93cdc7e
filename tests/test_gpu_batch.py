"""Batched local windows (vba_optimize_batch) on the GPU.

Each window of a batch must give exactly what vba_optimize gives it alone
(same kernels and arithmetic; LM control per window on the device), and both
must match the CPU FP64 oracle to the single-problem bar: same accept/reject
sequence, final cost within 1e-9 relative, camera centres within 1e-6 of the
scene extent."""
import numpy as np
import pytest

import oracle
import vba
from vba import abi

pytestmark = pytest.mark.gpu


def window_options(**kw):
    o = dict(linear_solver=abi.SOLVER_DENSE, max_iterations=10, robust_kernel=abi.KERNEL_CAUCHY, robust_scale=1.0)
    o.update(kw)
    return vba.default_options(**o)


def make_windows(n, seed0=100, kmax=14, lmax=900):
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(n):
        k = int(rng.integers(5, kmax + 1))
        l = int(rng.integers(300, lmax + 1))
        out.append(vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=k, num_landmarks=l, seed=seed0 + i))
    return out


def test_batch_equals_single_and_oracle(device):
    wins = make_windows(6)
    o = window_options()
    batch = device.optimize_batch([w.problem for w in wins], o)
    assert batch.status == abi.OK, batch.message
    for w, got in zip(wins, batch.windows):
        assert got.status == abi.OK
        alone = device.optimize(w.problem, o)
        assert alone.status == abi.OK, alone.message
        assert got.accepted == alone.accepted
        assert got.result.final_cost == alone.result.final_cost
        assert got.result.initial_cost == alone.result.initial_cost
        assert np.array_equal(got.poses, alone.poses)
        assert np.array_equal(got.landmarks, alone.landmarks)
        assert [it["lambda"] for it in got.iterations] == [it["lambda"] for it in alone.iterations]
        ref = oracle.optimize(w.problem, o)
        assert got.accepted == ref.accepted
        rel = abs(got.result.final_cost - ref.result.final_cost) / ref.result.final_cost
        assert rel <= 1e-9, rel
        d = np.linalg.norm(vba.camera_centres(got.poses) - vba.camera_centres(ref.poses), axis=1).max()
        assert d <= 1e-6 * w.scene_extent


def test_batch_mixed_kinds_and_errors(device):
    """Mono and rig windows in one batch; a window with an error is skipped
    with its own status and leaves the others unchanged."""
    circle = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=8, num_landmarks=300, seed=4)
    stereo = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=7, num_landmarks=350, seed=5)
    pose_only = vba.synthesize(abi.SYNTH_POSE_ONLY, num_landmarks=400, seed=6)
    broken = stereo.problem.copy()
    broken.rig_pose_ids = broken.rig_pose_ids.copy()
    broken.rig_pose_ids[0] = 999_999  # unknown pose id -> VBA_ERR_MISSING_POSE
    o = window_options(max_iterations=8)
    probs = [circle.problem, broken, stereo.problem, pose_only.problem]
    batch = device.optimize_batch(probs, o)
    assert batch.status == abi.OK, batch.message
    assert batch.windows[1].status == abi.ERR_MISSING_POSE
    assert batch.windows[1].result.error_id == 999_999
    for i in (0, 2, 3):
        got = batch.windows[i]
        alone = device.optimize(probs[i], o)
        assert got.status == alone.status == abi.OK
        assert got.accepted == alone.accepted
        assert got.result.final_cost == alone.result.final_cost
        assert np.array_equal(got.poses, alone.poses)


def test_batch_rejects_matrix_free(device):
    w = make_windows(1)[0]
    o = window_options(linear_solver=abi.SOLVER_MATRIX_FREE)
    batch = device.optimize_batch([w.problem], o)
    assert batch.status == abi.ERR_INVALID_CONFIGURATION


def test_batch_repeatable(device):
    wins = make_windows(4, seed0=300)
    o = window_options()
    a = device.optimize_batch([w.problem for w in wins], o)
    b = device.optimize_batch([w.problem for w in reversed(wins)], o)
    for x, y in zip(a.windows, reversed(b.windows)):
        assert x.result.final_cost == y.result.final_cost
        assert np.array_equal(x.poses, y.poses)


def test_prepared_batch_resolves_from_initial_state(device):
    wins = make_windows(3, seed0=500)
    o = window_options()
    one_shot = device.optimize_batch([w.problem for w in wins], o)
    pb = device.prepare_batch([w.problem for w in wins], o)
    try:
        for _ in range(2):
            out = pb.solve(o)
            assert out.status == abi.OK, out.message
            for i, (a, b) in enumerate(zip(out.windows, one_shot.windows)):
                assert a.result.final_cost == b.result.final_cost
                assert a.accepted == b.accepted
                poses, lms = pb.state(i)
                assert np.array_equal(poses, b.poses) and np.array_equal(lms, b.landmarks)
    finally:
        pb.free()


def test_graph_replays_match_eager(device):
    """A prepared batch records its LM iterations as a HIP graph on the first
    solve and replays it; every replay, with other device work between them,
    equals the kernel-by-kernel solve bitwise, and new options re-record."""
    wins = make_windows(4, seed0=700)
    o = window_options()
    eager_o = window_options()
    eager_o.flags = abi.FLAG_NO_GRAPH
    pb = device.prepare_batch([w.problem for w in wins], o)
    try:
        ref = pb.solve(eager_o)
        assert ref.status == abi.OK, ref.message
        ref_state = [pb.state(i) for i in range(len(wins))]
        other = make_windows(1, seed0=900)[0]
        for _ in range(3):
            out = pb.solve(o)
            assert out.status == abi.OK, out.message
            for i, (a, b) in enumerate(zip(out.windows, ref.windows)):
                assert a.result.final_cost == b.result.final_cost
                assert a.accepted == b.accepted
                poses, lms = pb.state(i)
                assert np.array_equal(poses, ref_state[i][0]) and np.array_equal(lms, ref_state[i][1])
            device.optimize(other.problem, o)  # eager work between replays
        o.max_iterations = 4
        eager_o.max_iterations = 4
        short_ref = pb.solve(eager_o)
        short = pb.solve(o)
        for a, b in zip(short.windows, short_ref.windows):
            assert a.result.final_cost == b.result.final_cost and a.accepted == b.accepted
    finally:
        pb.free()
