"""Observation weights and graduated non-convexity on the device, against
the oracle (tests/test_gnc_oracle.py pins the oracle with the reference's own
tests). FP64 bar: the same LM accept/reject sequence, final cost within 1e-9
relative, camera centres within 1e-6 of the scene extent; GNC: the same
outer-level count and verdict, weights within 1e-6."""
import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

GM, TLS = abi.GNC_GEMAN_MCCLURE, abi.GNC_TRUNCATED_LEAST_SQUARES


def test_zero_weight_mutes_the_outlier_on_device(device):
    ba = scenes.over_determined_bundle(inject_outlier=True)
    w = np.ones(len(ba.obs_pose_ids))
    w[[i for i in range(len(w)) if ba.obs_pose_ids[i] == 50 and ba.obs_landmark_ids[i] == 1]] = 0.0
    gpu = device.optimize_weighted(ba, vba.default_options(), w)
    ref = oracle.optimize_weighted(ba, vba.default_options(), w)
    assert gpu.status == abi.OK, gpu.message
    assert gpu.result.final_cost < 1e-18
    assert np.linalg.norm(vba.camera_centres(gpu.poses)[list(ba.pose_ids).index(50)] - [2, 0, 0]) < 1e-10
    assert gpu.accepted == ref.accepted


@pytest.mark.parametrize("solver", [abi.SOLVER_DENSE, abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT])
def test_weighted_parity(device, solver):
    s = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=16, num_landmarks=500, seed=3)
    rng = np.random.default_rng(1)
    w = rng.uniform(0.2, 2.0, s.problem.num_observations)
    o = vba.default_options(linear_solver=solver, max_iterations=12, robust_kernel=abi.KERNEL_HUBER,
                            robust_scale=2.0)
    ref = oracle.optimize_weighted(s.problem, o, w)
    gpu = device.optimize_weighted(s.problem, o, w)
    compare(gpu, ref, s.scene_extent)
    # unit weights are the unweighted device solve bit for bit
    one = device.optimize_weighted(s.problem, o, np.ones(len(w)))
    plain = device.optimize(s.problem, o)
    assert one.iterations == plain.iterations and one.result.final_cost == plain.result.final_cost
    np.testing.assert_array_equal(one.poses, plain.poses)


def test_invalid_weights_on_device(device):
    ba = scenes.over_determined_bundle(inject_outlier=False)
    n = len(ba.obs_pose_ids)
    assert device.optimize_weighted(ba, vba.default_options(), [1.0]).status == abi.ERR_OBSERVATION_WEIGHT_COUNT
    w = np.ones(n)
    w[3] = np.nan
    out = device.optimize_weighted(ba, vba.default_options(), w)
    assert out.status == abi.ERR_INVALID_OBSERVATION_WEIGHT and out.result.error_id == 3
    o = vba.default_options(refine_intrinsics=1)
    assert device.optimize_weighted(ba, o, np.ones(n)).status == abi.ERR_WEIGHTS_WITH_INTRINSICS


def gnc_compare(dev_out, ref_out, extent):
    st, res, poses, _, w = dev_out
    rst, rres, rposes, _, rw = ref_out
    assert st == rst == abi.OK
    assert res.outer_iterations == rres.outer_iterations and res.converged == rres.converged
    assert res.observation_count == rres.observation_count
    np.testing.assert_allclose(w, rw, atol=1e-6, rtol=0)
    for a, b in [(res.initial_cost, rres.initial_cost), (res.final_cost, rres.final_cost),
                 (res.inlier_cost, rres.inlier_cost)]:
        assert abs(a - b) <= 1e-6 * max(abs(b), 1e-9) + 1e-12, (a, b)
    d = np.linalg.norm(vba.camera_centres(poses) - vba.camera_centres(rposes), axis=1).max()
    assert d <= 1e-6 * extent, d


@pytest.mark.parametrize("kernel,outliers,shift", [(GM, [7, 18, 33, 41, 52], 70.0), (TLS, [25, 38, 50], 90.0),
                                                  (TLS, [], 0.0)])
def test_gnc_parity_on_the_reference_scene(device, kernel, outliers, shift):
    ba, truth = scenes.gnc_bundle(outliers, shift)
    o = vba.default_options(linear_solver=abi.SOLVER_DENSE)
    g = vba.default_gnc_options(kernel=kernel, c=5.0)
    dev = device.optimize_gnc(ba, o, g)
    ref = oracle.optimize_gnc(ba, o, g)
    gnc_compare(dev, ref, 7.0)
    w = dev[4]
    if outliers:
        assert max(w[i] for i in outliers) < 1e-3
    if kernel == TLS:
        err = np.linalg.norm(vba.camera_centres(dev[2]) - vba.camera_centres(np.array([p for _, p in truth])),
                             axis=1).max()
        assert err < 1e-4


def test_gnc_auto_scale_parity(device):
    ba, _ = scenes.gnc_bundle([7, 18, 33, 41, 52], 80.0)
    o = vba.default_options(linear_solver=abi.SOLVER_DENSE)
    g = vba.default_gnc_options(kernel=TLS, c=2.0, has_auto_scale=1, auto_scale=3.5, auto_scale_readapt=1)
    gnc_compare(device.optimize_gnc(ba, o, g), oracle.optimize_gnc(ba, o, g), 7.0)


def test_gnc_with_outliers_on_a_loop(device):
    """GNC around the matrix-free solver on a loop scene with 2 % of the
    observations replaced by wrong matches. The inner solves stop at PCG
    rtol 1e-6, so the device and the oracle (with different rounding) end
    within that tolerance of each other, and a truncated-least-squares weight
    in the transition band moves with the residual at the band's slope
    (sqrt(mu (mu + 1)) c / r^2, several hundred per pixel at the last level):
    the bar is the GNC cost to 1e-6 relative and the weights to 1e-3."""
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=40, num_landmarks=2000, seed=6, scene_scale=16.0)
    p = s.problem.copy()
    rng = np.random.default_rng(2)
    bad = rng.choice(len(p.obs_xy), size=len(p.obs_xy) // 50, replace=False)
    p.obs_xy[bad] += rng.uniform(-60.0, 60.0, (len(bad), 2))
    o = vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-6)
    g = vba.default_gnc_options(kernel=TLS, c=3.0, max_outer=12, inner_iterations=3)
    dev = device.optimize_gnc(p, o, g)
    ref = oracle.optimize_gnc(p, o, g)
    st, res, _, _, w = dev
    assert st == abi.OK
    assert res.outer_iterations == ref[1].outer_iterations
    assert np.mean(w[bad] < 0.5) > 0.9
    assert abs(res.final_cost - ref[1].final_cost) <= 1e-6 * ref[1].final_cost
    np.testing.assert_allclose(w, ref[4], atol=1e-3, rtol=0)
