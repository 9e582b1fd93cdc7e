"""GPU parity: libvba (HIP, through the C ABI) against the CPU FP64 oracle on
the same seeded problems.

Bar (FP64 solver): same LM accept/reject sequence, final cost within 1e-9
relative, camera centres within 1e-6 of the scene extent. Mixed precision
(FP32 Jacobian storage and PCG matvecs, FP64 refinement): final cost within
1e-6 relative of the FP64 solver."""
import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi

pytestmark = pytest.mark.gpu

FP64_COST_RTOL = 1e-9
MIXED_COST_RTOL = 1e-6
CENTRE_TOL = 1e-6  # x scene extent


def compare(gpu, ref, extent, cost_rtol=FP64_COST_RTOL, same_sequence=True):
    assert gpu.status == ref.status == abi.OK, gpu.message
    if same_sequence:
        assert gpu.accepted == ref.accepted
    rel = abs(gpu.result.final_cost - ref.result.final_cost) / max(abs(ref.result.final_cost), 1e-300)
    assert rel <= cost_rtol, (gpu.result.final_cost, ref.result.final_cost, rel)
    assert abs(gpu.result.initial_cost - ref.result.initial_cost) <= 1e-12 * abs(ref.result.initial_cost) + 1e-12
    d = np.linalg.norm(vba.camera_centres(gpu.poses) - vba.camera_centres(ref.poses), axis=1).max()
    assert d <= CENTRE_TOL * extent, (d, extent)


def test_drifted_pose_recovered_on_gpu(device):
    ba = scenes.truth_bundle(drift30=[1.05, 0.02, 0.03])
    out = device.optimize(ba, vba.default_options())
    assert out.status == abi.OK, out.message
    assert out.result.final_cost < 1e-12 and out.result.converged
    assert np.linalg.norm(vba.camera_centres(out.poses)[2] - [1.0, 0.0, 0.0]) < 1e-6


@pytest.mark.parametrize("solver", [abi.SOLVER_DENSE, abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT])
def test_circle_parity(device, solver):
    s = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=20, num_landmarks=600, seed=1)
    o = vba.default_options(linear_solver=solver, max_iterations=20)
    ref = oracle.optimize(s.problem, o)
    gpu = device.optimize(s.problem, o)
    compare(gpu, ref, s.scene_extent)


@pytest.mark.parametrize("solver", [abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT])
@pytest.mark.parametrize("policy", [abi.PCG_MAXIT_REJECT, abi.PCG_MAXIT_ACCEPT_ITERATE])
def test_loop_pcg_parity(device, policy, solver):
    """The oracle runs the implicit operator for both PCG solvers: the
    explicit matrix is the same S, so only rounding separates the two."""
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=80, num_landmarks=6000, seed=2, scene_scale=16.0)
    o = vba.default_options(linear_solver=solver, max_iterations=12, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                            pcg_maxit_policy=policy)
    ref = oracle.optimize(s.problem, o)
    gpu = device.optimize(s.problem, o)
    compare(gpu, ref, s.scene_extent, cost_rtol=1e-6)


@pytest.mark.parametrize("solver", [abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT])
def test_mixed_precision_matches_fp64(device, solver):
    """Mixed precision stores G in FP32 and pairs it with the inverse of the
    V summed from those same FP32 blocks; the operator's arithmetic stays
    FP64 and PCG is checked against the FP64 true residual. On the reference
    GPU backend's PCG settings (rtol 1e-4, at most 100 iterations, the last
    iterate used at the cap) both precisions take the same LM path."""
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=80, num_landmarks=6000, seed=3, scene_scale=16.0)
    o = vba.default_options(linear_solver=solver, max_iterations=15, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                            pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE)
    fp64 = device.optimize(s.problem, o)
    o.precision = abi.PRECISION_MIXED
    mixed = device.optimize(s.problem, o)
    assert fp64.status == mixed.status == abi.OK, mixed.message
    assert mixed.accepted == fp64.accepted
    rel = abs(mixed.result.final_cost - fp64.result.final_cost) / fp64.result.final_cost
    assert rel <= MIXED_COST_RTOL, rel
    d = np.linalg.norm(vba.camera_centres(mixed.poses) - vba.camera_centres(fp64.poses), axis=1).max()
    assert d <= CENTRE_TOL * s.scene_extent, (d, s.scene_extent)


def test_stereo_rig_window_parity(device):
    s = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=10, num_landmarks=400, seed=6)
    for solver in (abi.SOLVER_DENSE, abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT):
        o = vba.default_options(linear_solver=solver, max_iterations=10, robust_kernel=abi.KERNEL_CAUCHY,
                                robust_scale=1.0)
        compare(device.optimize(s.problem, o), oracle.optimize(s.problem, o), s.scene_extent)


@pytest.mark.parametrize("solver", [abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT])
def test_city_opencv_rig_parity(device, solver):
    s = vba.synthesize(abi.SYNTH_CITY, num_cameras=60, num_landmarks=3000, seed=8, scene_scale=30.0)
    o = vba.default_options(linear_solver=solver, max_iterations=8, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=2.0)
    compare(device.optimize(s.problem, o), oracle.optimize(s.problem, o), s.scene_extent, cost_rtol=1e-6)


def test_feasibility_gate_on_gpu(device):
    ba = scenes.make_problem(scenes.pinhole_camera(), [(10, scenes.pose_at([0, 0, 0]))],
                             [(1, np.array([1.0, 0.0, 0.1]))], [(10, 1, np.array([105320.0, 240.0]))],
                             fixed_poses=(10,))
    out = device.optimize(ba, vba.default_options(max_iterations=1))
    assert out.status == abi.OK, out.message
    assert np.array_equal(out.landmarks[0], [1.0, 0.0, 0.1])
    assert len(out.iterations) == 1 and not out.iterations[0]["step_accepted"]


def test_error_codes_match_oracle(device):
    ba = scenes.truth_bundle()
    bad = ba.copy()
    bad.obs_landmark_ids[5] = 1234
    out = device.optimize(bad, vba.default_options())
    assert out.status == abi.ERR_MISSING_LANDMARK and out.result.error_id == 1234
    mf = vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE)
    nofix = scenes.truth_bundle(fixed=())
    assert device.optimize(nofix, mf).status == oracle.optimize(nofix, mf).status == abi.ERR_INELIGIBLE


def test_repeat_runs_are_bitwise_identical(device):
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=60, num_landmarks=4000, seed=12, scene_scale=14.0)
    o = vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_iterations=6, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-4, precision=abi.PRECISION_MIXED)
    a = device.optimize(s.problem, o)
    b = device.optimize(s.problem, o)
    assert a.status == b.status == abi.OK
    assert np.array_equal(a.poses, b.poses) and np.array_equal(a.landmarks, b.landmarks)
    assert a.result.final_cost == b.result.final_cost


def test_explicit_matches_implicit_per_iteration(device):
    """implicit_schur.rs:952-975 and :1190-1197 compare the implicit operator
    with the explicitly formed Schur complement; here the two PCG solvers run
    the same LM solve, so every LM iteration takes the same step decision,
    the same PCG iteration count (one either way at a tie) and a cost within
    FP64 rounding."""
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=60, num_landmarks=4000, seed=21, scene_scale=14.0)
    o = vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_iterations=8, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0)
    mf = device.optimize(s.problem, o)
    o.linear_solver = abi.SOLVER_PCG_EXPLICIT
    ex = device.optimize(s.problem, o)
    assert mf.status == ex.status == abi.OK, ex.message
    assert mf.accepted == ex.accepted
    for a, b in zip(mf.iterations, ex.iterations):
        assert abs(a["pcg_iterations"] - b["pcg_iterations"]) <= 1, (a, b)
        assert abs(a["cost_after"] - b["cost_after"]) <= 1e-9 * abs(a["cost_after"]), (a, b)
    rel = abs(mf.result.final_cost - ex.result.final_cost) / mf.result.final_cost
    assert rel <= 1e-9, rel


def test_explicit_repeat_is_bitwise_identical(device):
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=60, num_landmarks=4000, seed=12, scene_scale=14.0)
    o = vba.default_options(linear_solver=abi.SOLVER_PCG_EXPLICIT, max_iterations=6, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-4, precision=abi.PRECISION_MIXED)
    a = device.optimize(s.problem, o)
    b = device.optimize(s.problem, o)
    assert a.status == b.status == abi.OK, a.message
    assert np.array_equal(a.poses, b.poses) and np.array_equal(a.landmarks, b.landmarks)
    assert a.result.final_cost == b.result.final_cost
