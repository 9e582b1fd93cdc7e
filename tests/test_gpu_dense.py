"""The dense normal-equation solve on the GPU: vba_dense_solve
(solve_normal_equations, pose_graph.rs:2259-2270) against the oracle's
restatement on a positive definite system (the Cholesky path), a symmetric
indefinite one (Cholesky fails, LU with partial pivoting solves it) and a
singular one (both fail: SingularSystem); and the LM's singular-system path
(lm.rs:608-662: the step is rejected and lambda raised) on a problem whose
reduced camera matrix has an exactly zero block, alone and in a batch."""
import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi

pytestmark = pytest.mark.gpu

SOLVE_RTOL = 1e-10  # blocked MFMA Cholesky / unblocked LU vs nalgebra's order


def spd(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    return a @ a.T + n * np.eye(n), rng.standard_normal(n)


def indefinite(n, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = rng.uniform(0.5, 2.0, n) * np.where(np.arange(n) % 3 == 1, -1.0, 1.0)
    return (q * ev) @ q.T, rng.standard_normal(n)


@pytest.mark.parametrize("n", [6, 37, 96, 300])
def test_dense_solve_cholesky_path(device, n):
    h, b = spd(n, n)
    st, x = device.dense_solve(h, b)
    ref = oracle.solve_normal_equations(h, b)
    assert st == abi.OK and ref is not None
    assert np.abs(x - ref).max() <= SOLVE_RTOL * np.abs(ref).max()
    assert np.abs(h @ x - b).max() <= 1e-9 * np.abs(b).max()


def test_dense_solve_cholesky_reads_the_lower_triangle(device):
    h, b = spd(50, 9)
    hu = h.copy()
    hu[np.triu_indices(50, 1)] = 1e3
    st, x = device.dense_solve(h, b)
    stu, xu = device.dense_solve(hu, b)
    assert st == stu == abi.OK
    assert np.array_equal(x, xu)


@pytest.mark.parametrize("n", [5, 48, 130])
def test_dense_solve_lu_fallback(device, n):
    h, b = indefinite(n, 100 + n)
    assert np.linalg.eigvalsh(h).min() < 0  # Cholesky must fail
    st, x = device.dense_solve(h, b)
    ref = oracle.solve_normal_equations(h, b)
    assert st == abi.OK and ref is not None
    assert np.abs(x - ref).max() <= SOLVE_RTOL * np.abs(ref).max()
    assert np.abs(h @ x - b).max() <= 1e-9 * np.abs(b).max()


def test_dense_solve_lu_reads_the_whole_matrix(device):
    """LU takes H as given, not its lower triangle: a non-symmetric H with a
    negative pivot is solved as H, column-major as the caller passes it."""
    rng = np.random.default_rng(7)
    h = rng.standard_normal((40, 40)) + 40 * np.diag(np.where(np.arange(40) % 2, -1.0, 1.0))
    b = rng.standard_normal(40)
    st, x = device.dense_solve(h, b)
    assert st == abi.OK
    assert np.abs(x - np.linalg.solve(h, b)).max() <= 1e-10 * np.abs(x).max()


@pytest.mark.parametrize("n", [257, 700, 1500])
def test_dense_solve_device_wide_cholesky(device, n):
    """Beyond 256 unknowns the factor runs across the device (64-column
    panels, MFMA trailing updates)."""
    h, b = spd(n, n)
    st, x = device.dense_solve(h, b)
    ref = oracle.solve_normal_equations(h, b)
    assert st == abi.OK and ref is not None
    assert np.abs(x - ref).max() <= SOLVE_RTOL * np.abs(ref).max()
    assert np.abs(h @ x - b).max() <= 1e-9 * np.abs(b).max()


@pytest.mark.parametrize("n", [300, 900])
def test_dense_solve_device_wide_lu_fallback(device, n):
    """The device-wide factor meets a negative pivot; the LU fallback runs on
    the copy saved before it."""
    h, b = indefinite(n, 300 + n)
    st, x = device.dense_solve(h, b)
    ref = oracle.solve_normal_equations(h, b)
    assert st == abi.OK and ref is not None
    assert np.abs(x - ref).max() <= SOLVE_RTOL * np.abs(ref).max()


def test_dense_solve_device_wide_singular(device):
    h, b = spd(400, 4)
    h[123, :] = 0.0
    h[:, 123] = 0.0
    assert oracle.solve_normal_equations(h, b) is None
    st, x = device.dense_solve(h, b)
    assert st == abi.ERR_SINGULAR_SYSTEM and x is None


def test_dense_solve_singular(device):
    h, b = spd(24, 3)
    h[7, :] = 0.0
    h[:, 7] = 0.0
    assert oracle.solve_normal_equations(h, b) is None
    st, x = device.dense_solve(h, b)
    assert st == abi.ERR_SINGULAR_SYSTEM and x is None
    assert "singular" in device.last_error()


def singular_bundle():
    """truth_bundle with KF30 drifted and a free pose 40 that no observation
    sees: with pure Gauss-Newton (no initial lambda) the reduced camera
    matrix has an exactly zero 6x6 block, Cholesky meets a zero pivot and LU
    a zero column, so every LM iteration takes the singular-system path."""
    ba = scenes.truth_bundle(drift30=[1.05, 0.02, 0.03])
    poses = [(int(i), p) for i, p in zip(ba.pose_ids, ba.poses)] + [(40, scenes.pose_at([1.5, 0.0, 0.0]))]
    lms = [(int(i), x) for i, x in zip(ba.landmark_ids, ba.landmarks)]
    obs = [(int(p), int(l), uv) for p, l, uv in zip(ba.obs_pose_ids, ba.obs_landmark_ids, ba.obs_xy)]
    return scenes.make_problem(scenes.pinhole_camera(), poses, lms, obs, fixed_poses=(10, 20))


def test_singular_system_rejects_every_step(device):
    ba = singular_bundle()
    o = vba.default_options(has_initial_lambda=0, max_iterations=6)
    ref = oracle.optimize(ba, o)
    gpu = device.optimize(ba, o)
    assert ref.status == gpu.status == abi.OK
    assert ref.accepted == gpu.accepted == [False] * 6
    for a, b in zip(gpu.iterations, ref.iterations):
        assert a["lambda"] == b["lambda"] == 0.0
        assert a["cost_after"] == a["cost_before"]
    assert gpu.result.final_cost == gpu.result.initial_cost
    assert abs(gpu.result.final_cost - ref.result.final_cost) <= 1e-12 * ref.result.final_cost
    assert np.array_equal(gpu.poses, ba.poses)


def test_singular_window_in_a_batch(device):
    """A singular window is rejected step after step on the device's
    per-window LM control while the window beside it converges as alone."""
    sing = singular_bundle()
    good = scenes.truth_bundle(drift30=[1.05, 0.02, 0.03])
    o = vba.default_options(has_initial_lambda=0, max_iterations=6)
    out = device.optimize_batch([good, sing], o)
    assert out.status == abi.OK
    alone = device.optimize(good, o)
    w0, w1 = out.windows
    assert w0.status == w1.status == abi.OK
    assert w0.accepted == alone.accepted
    assert w0.result.final_cost == alone.result.final_cost
    assert w1.accepted == [False] * 6
    assert w1.result.final_cost == w1.result.initial_cost


def dense_loop(cams, lms, seed):
    return vba.synthesize(abi.SYNTH_LOOP, num_cameras=cams, num_landmarks=lms, seed=seed,
                          scene_scale=200.0 * cams / 1200.0)


@pytest.mark.parametrize("solver", ["dense", "sparse"])
def test_lm_with_the_device_wide_factor(device, solver):
    """A 200-camera loop (1,194 unknowns in the reduced system) through the
    dense LM, whose factor now runs across the device; the Sparse solver
    (Cholesky only) takes the same steps while the factor succeeds."""
    s = dense_loop(200, 12000, 31)
    ls = abi.SOLVER_DENSE if solver == "dense" else abi.SOLVER_SPARSE
    o = vba.default_options(linear_solver=ls, max_iterations=6, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0)
    ref = oracle.optimize(s.problem, o, threads=8)
    gpu = device.optimize(s.problem, o)
    assert ref.status == gpu.status == abi.OK, gpu.message
    assert gpu.accepted == ref.accepted
    rel = abs(gpu.result.final_cost - ref.result.final_cost) / ref.result.final_cost
    assert rel <= 1e-9, rel
    d = np.linalg.norm(vba.camera_centres(gpu.poses) - vba.camera_centres(ref.poses), axis=1).max()
    assert d <= 1e-6 * s.scene_extent
    if solver == "sparse":
        dense = device.optimize(s.problem, vba.default_options(
            linear_solver=abi.SOLVER_DENSE, max_iterations=6, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0))
        assert dense.result.final_cost == gpu.result.final_cost


def test_sparse_solver_has_no_lu_fallback(device):
    """LinearSolver::Sparse (normal_equations.rs:1483-1535): a failed block
    Cholesky is SingularSystem, which the LM takes as a rejected step; the
    singular problem is rejected step after step as with Dense, where LU
    fails too."""
    ba = singular_bundle()
    o = vba.default_options(has_initial_lambda=0, max_iterations=4, linear_solver=abi.SOLVER_SPARSE)
    ref = oracle.optimize(ba, o)
    gpu = device.optimize(ba, o)
    assert ref.status == gpu.status == abi.OK
    assert ref.accepted == gpu.accepted == [False] * 4
