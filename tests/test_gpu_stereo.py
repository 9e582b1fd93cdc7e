"""Rectified and general stereo observations on the device (vba_problem.stereo)
against the CPU FP64 oracle.

The reference's stereo scenes (pipelines/slam/tests/bundle_adjustment.rs:
1519-1853) meet the reference's own assertions on the device; larger scenes
with all observation kinds (monocular, rectified stereo, general stereo, rig)
meet the parity bar of the other device tests: the same LM accept/reject
sequence, final cost within 1e-9 relative (1e-6 behind a loose PCG
tolerance), camera centres within 1e-6 of the scene extent. Weights and GNC
follow the flattened observation order; batched windows with stereo terms
give exactly what each window gives alone."""
import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi

from test_gpu_gnc import gnc_compare
from test_gpu_parity import MIXED_COST_RTOL, compare

pytestmark = pytest.mark.gpu


def reference_scenes():
    return {
        "metric_pose_single_anchor": (scenes.truth_stereo_bundle(
            0.5, fixed=(10,), poses={30: scenes.pose_at([1.07, 0.04, 0.05])}), 1e-12),
        "drifted_landmark": (scenes.truth_stereo_bundle(
            0.5, fixed=(10, 20, 30), landmarks={5: np.array([0.20, 0.10, 6.30])}), 1e-14),
        "general_landmark": (scenes.truth_general_stereo_bundle(start=[0.45, 0.08, 4.8])[0], 1e-14),
        "general_pose": (scenes.general_pose_bundle()[0], 1e-12),
        "mixed_mono_stereo": (scenes.mixed_mono_stereo_bundle(), 1e-12),
    }


@pytest.mark.parametrize("name", list(reference_scenes()))
def test_reference_stereo_scenes_on_device(device, name):
    p, bar = reference_scenes()[name]
    o = vba.default_options()
    gpu = device.optimize(p, o)
    ref = oracle.optimize(p, o)
    assert gpu.status == ref.status == abi.OK, gpu.message
    assert abs(gpu.result.initial_cost - ref.result.initial_cost) <= 1e-12 * ref.result.initial_cost
    assert gpu.result.final_cost < bar
    d = np.linalg.norm(vba.camera_centres(gpu.poses) - vba.camera_centres(ref.poses), axis=1).max()
    assert d < 1e-7
    assert np.abs(gpu.landmarks - ref.landmarks).max() < 1e-7


def test_stereo_validation_on_device(device):
    """MissingStereoBaseline, the pinhole requirement and the general pair's
    checks, with the oracle's codes (lm.rs:168-217)."""
    for b in (float("nan"), 0.0, -1.0):
        p = scenes.truth_stereo_bundle(0.5, fixed=(10,))
        p.stereo_baseline = b
        assert device.optimize(p, vba.default_options()).status == abi.ERR_MISSING_STEREO_BASELINE
    p = scenes.truth_stereo_bundle(0.12, fixed=(10,))
    p.camera = vba.camera(abi.CAMERA_OPENCV, [500.0, 500.0, 320.0, 240.0, 0, 0, 0, 0], 640, 480)
    assert device.optimize(p, vba.default_options()).status == abi.ERR_UNSUPPORTED_CAMERA_MODEL
    g, _ = scenes.truth_general_stereo_bundle()
    g.pairs = [vba.sensor(vba.camera(abi.CAMERA_PINHOLE, [500.0, 500.0], 640, 480), scenes.left_to_right())]
    g = g.copy()
    assert device.optimize(g, vba.default_options()).status == abi.ERR_UNSUPPORTED_CAMERA_MODEL
    g.general_landmark_ids[:] = 99
    out = device.optimize(g, vba.default_options())
    assert out.status == abi.ERR_MISSING_LANDMARK and out.result.error_id == 99
    s = scenes.truth_stereo_bundle(0.5, fixed=(10,))
    s.stereo_obs[3, 2] = np.nan
    mf = vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE)
    assert device.optimize(s, mf).status == oracle.optimize(s, mf).status == abi.ERR_INELIGIBLE


def loop_with_stereo(seed=4):
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=40, num_landmarks=3000, seed=seed, scene_scale=16.0)
    p = s.problem.copy()
    p.camera = vba.camera(abi.CAMERA_PINHOLE, [800.0, 800.0, 512.0, 384.0], 1024, 768)
    return scenes.split_into_stereo(p, np.random.default_rng(seed)), s.scene_extent


@pytest.mark.parametrize("solver", [abi.SOLVER_DENSE, abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT])
def test_loop_with_stereo_parity(device, solver):
    p, extent = loop_with_stereo()
    if solver == abi.SOLVER_DENSE:
        o = vba.default_options(linear_solver=solver, max_iterations=12, robust_kernel=abi.KERNEL_HUBER,
                                robust_scale=1.0)
        compare(device.optimize(p, o), oracle.optimize(p, o), extent)
    else:
        o = vba.default_options(linear_solver=solver, max_iterations=10, max_pcg_iterations=100,
                                pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                                pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE)
        compare(device.optimize(p, o), oracle.optimize(p, o), extent, cost_rtol=1e-6)


def rig_window_with_stereo(seed=6, k=10, l=400):
    s = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=k, num_landmarks=l, seed=seed)
    return scenes.add_stereo_from_state(s.problem, np.random.default_rng(seed), 600, 400), s.scene_extent


def test_rig_window_with_both_stereo_kinds_parity(device):
    """Rig, rectified and general observations of the same edges."""
    p, extent = rig_window_with_stereo()
    for solver in (abi.SOLVER_DENSE, abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT):
        o = vba.default_options(linear_solver=solver, max_iterations=10, robust_kernel=abi.KERNEL_CAUCHY,
                                robust_scale=1.0)
        compare(device.optimize(p, o), oracle.optimize(p, o), extent)


def test_stereo_weighted_parity(device):
    p, extent = loop_with_stereo(seed=5)
    w = np.random.default_rng(1).uniform(0.0, 1.0, p.num_observations)
    o = vba.default_options(linear_solver=abi.SOLVER_DENSE, max_iterations=10)
    compare(device.optimize_weighted(p, o, w), oracle.optimize_weighted(p, o, w), extent)
    nm, ns = len(p.obs_pose_ids), len(p.stereo_pose_ids)
    w_st = np.concatenate([np.ones(nm), np.zeros(ns), np.ones(p.num_observations - nm - ns)])
    gpu = device.optimize_weighted(p, o, w_st)
    ref = oracle.optimize_weighted(p, o, w_st)
    compare(gpu, ref, extent)


def test_stereo_gnc_parity(device):
    """GNC on the flattened order: wrong matches among all three kinds."""
    p, extent = loop_with_stereo(seed=6)
    rng = np.random.default_rng(2)
    for arr in (p.obs_xy, p.stereo_obs, p.general_obs):
        bad = rng.choice(len(arr), size=len(arr) // 40, replace=False)
        arr[bad] += rng.uniform(-60.0, 60.0, (len(bad), arr.shape[1]))
    o = vba.default_options(linear_solver=abi.SOLVER_DENSE)
    g = vba.default_gnc_options(kernel=abi.GNC_TRUNCATED_LEAST_SQUARES, c=3.0, max_outer=8, inner_iterations=3)
    gnc_compare(device.optimize_gnc(p, o, g), oracle.optimize_gnc(p, o, g), extent)


def test_stereo_windows_batch_equals_single_and_oracle(device):
    wins = [rig_window_with_stereo(seed=200 + i, k=6 + i, l=300 + 50 * i) for i in range(4)]
    mono = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=8, num_landmarks=300, seed=4)
    probs = [w for w, _ in wins] + [mono.problem]
    o = vba.default_options(linear_solver=abi.SOLVER_DENSE, max_iterations=10, robust_kernel=abi.KERNEL_CAUCHY,
                            robust_scale=1.0)
    batch = device.optimize_batch(probs, o)
    assert batch.status == abi.OK, batch.message
    for i, (p, got) in enumerate(zip(probs, batch.windows)):
        alone = device.optimize(p, o)
        assert got.status == alone.status == abi.OK, alone.message
        assert got.accepted == alone.accepted
        assert got.result.final_cost == alone.result.final_cost
        assert np.array_equal(got.poses, alone.poses) and np.array_equal(got.landmarks, alone.landmarks)
        extent = wins[i][1] if i < len(wins) else mono.scene_extent
        compare(got, oracle.optimize(p, o), extent)


def test_stereo_mixed_precision_matches_fp64(device):
    p, extent = loop_with_stereo(seed=7)
    o = vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_iterations=12, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                            pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE)
    fp64 = device.optimize(p, o)
    o.precision = abi.PRECISION_MIXED
    mixed = device.optimize(p, o)
    assert fp64.status == mixed.status == abi.OK, mixed.message
    assert mixed.accepted == fp64.accepted
    rel = abs(mixed.result.final_cost - fp64.result.final_cost) / fp64.result.final_cost
    assert rel <= MIXED_COST_RTOL, rel


def test_stereo_requests_served(device):
    """The request loop copies (or borrows) the stereo terms with the rest of
    the problem: served results equal the solo solves bit for bit."""
    cauchy = vba.default_options(max_iterations=10, robust_kernel=abi.KERNEL_CAUCHY, robust_scale=1.0)
    reqs = [(rig_window_with_stereo(seed=300 + i, k=6 + i, l=300)[0], cauchy) for i in range(3)]
    reqs.append((loop_with_stereo(seed=8)[0],
                 vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_iterations=8, max_pcg_iterations=100,
                                     pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                                     pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE)))
    with vba.Server(window_streams=2, submap_streams=1, batch_wait_seconds=0.005) as srv:
        tickets = [srv.submit(p, o, borrow=(i % 2 == 1)) for i, (p, o) in enumerate(reqs)]
        served = [srv.wait(t) for t in tickets]
    for (p, o), got in zip(reqs, served):
        alone = device.optimize(p, o)
        assert got.status == alone.status == abi.OK, (got.status, alone.message)
        assert got.result.final_cost == alone.result.final_cost
        assert np.array_equal(got.poses, alone.poses) and np.array_equal(got.landmarks, alone.landmarks)
