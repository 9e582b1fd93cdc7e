"""The oracle's rectified and general stereo terms, pinned.

Restates the stereo tests of pipelines/slam/tests/bundle_adjustment.rs
(:825-850, :1519-1853) on the CPU oracle, and pins what those leave open with
independent numpy geometry: the cost against the textbook right camera (a
rectified right camera sits at +b along the left x axis; a general one sees
T_right<-left X_c through its own pinhole), and the first LM step against the
full damped normal system built from finite-difference Jacobians of those
residuals, so a dropped Jacobian term or a wrong sign moves the step.
"""
import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi


def textbook_residuals(p):
    """Every observation's residual rows from the scene's own geometry."""
    P = {int(i): p.poses[k] for k, i in enumerate(p.pose_ids)}
    X = {int(i): p.landmarks[k] for k, i in enumerate(p.landmark_ids)}
    out = []
    for k in range(len(p.obs_pose_ids)):
        xc = scenes.transform(P[int(p.obs_pose_ids[k])], X[int(p.obs_landmark_ids[k])])
        out.append(scenes.project_pinhole((500.0, 500.0), (320.0, 240.0), xc) - p.obs_xy[k])
    for k in range(len(p.stereo_pose_ids)):
        pose, x = P[int(p.stereo_pose_ids[k])], X[int(p.stereo_landmark_ids[k])]
        uv = scenes.project_pinhole((500.0, 500.0), (320.0, 240.0), scenes.transform(pose, x))
        ur = scenes.rectified_right_u(pose, x, p.stereo_baseline)
        out.append(np.array([uv[0], uv[1], ur]) - p.stereo_obs[k])
    for k in range(len(p.general_pose_ids)):
        pair = p.pairs[int(p.general_pair[k])]
        lr = np.array(list(pair.sensor_from_rig))
        pred = scenes.general_obs(P[int(p.general_pose_ids[k])], X[int(p.general_landmark_ids[k])], lr)
        out.append(np.array(pred) - p.general_obs[k])
    return out


def textbook_cost(p):
    return sum(float(r @ r) for r in textbook_residuals(p))


def test_stereo_cost_is_zero_at_truth():
    """bundle_adjustment.rs:1553-1561 and :1669-1677."""
    c, bad = oracle.cost(scenes.truth_stereo_bundle(0.5))
    assert c < 1e-20 and bad == 0
    g, truth = scenes.truth_general_stereo_bundle()
    c, bad = oracle.cost(g)
    assert c < 1e-20 and bad == 0


def test_cost_matches_the_textbook_right_cameras():
    rng = np.random.default_rng(3)
    lms = {l: X + rng.normal(0, 0.05, 3) for l, X in scenes.world_grid()}
    p = scenes.truth_stereo_bundle(0.5, landmarks=lms)
    assert np.isclose(oracle.cost(p)[0], textbook_cost(p), rtol=1e-12)
    g, _ = scenes.truth_general_stereo_bundle(start=[0.45, 0.08, 4.8])
    assert np.isclose(oracle.cost(g)[0], textbook_cost(g), rtol=1e-12)
    m = scenes.mixed_mono_stereo_bundle()
    assert oracle.cost(m)[0] > 1.0
    assert np.isclose(oracle.cost(m)[0], textbook_cost(m), rtol=1e-12)


def test_point_behind_a_stereo_camera_is_nonprojectable():
    """nonprojectable_observation_count (problem.rs:264-290): a landmark
    behind the left camera drops its rectified term; one in front of the
    left camera but behind the right one drops the general term."""
    p = scenes.truth_stereo_bundle(0.5, landmarks={5: np.array([0.0, 0.0, -2.0])})
    c, bad = oracle.cost(p)
    assert bad == 3 and c == pytest.approx(textbook_cost(scenes.truth_stereo_bundle(0.5)), abs=1e-9)
    lr = scenes.pose_from_rt(np.eye(3), np.array([0.0, 0.0, -4.5]))  # right camera 4.5 m ahead
    g, _ = scenes.truth_general_stereo_bundle()
    g.pairs = [vba.sensor(scenes.right_camera(), lr)]
    g = g.copy()
    assert oracle.cost(g)[1] == 1


@pytest.mark.parametrize("baseline", [float("nan"), 0.0, -0.5, float("inf")])
def test_missing_baseline_is_rejected(baseline):
    """bundle_stereo_optimize_rejects_missing_baseline (:1563-1582)."""
    p = scenes.truth_stereo_bundle(0.5)
    p.stereo_baseline = baseline
    out = oracle.optimize(p, vba.default_options())
    assert out.status == abi.ERR_MISSING_STEREO_BASELINE


def test_stereo_needs_a_pinhole_camera():
    """non_pinhole_lens_with_stereo_or_unknown_model_is_rejected (:825-850)."""
    for model, params in ((abi.CAMERA_OPENCV, [500.0, 500.0, 320.0, 240.0, 0, 0, 0, 0]),
                          (abi.CAMERA_SIMPLE_RADIAL, [500.0, 320.0, 240.0, 0.0])):
        p = scenes.truth_stereo_bundle(0.12, fixed=(10,))
        p.camera = vba.camera(model, params, 640, 480)
        assert oracle.optimize(p, vba.default_options()).status == abi.ERR_UNSUPPORTED_CAMERA_MODEL
    p = scenes.truth_stereo_bundle(0.12, fixed=(10,))
    p.camera = vba.camera(abi.CAMERA_SIMPLE_PINHOLE, [500.0, 320.0, 240.0], 640, 480)
    assert oracle.optimize(p, vba.default_options()).status == abi.OK


def test_general_pair_needs_right_intrinsics_and_ids():
    """lm.rs:207-217: ids first, then the right camera's intrinsics."""
    g, _ = scenes.truth_general_stereo_bundle()
    bad = g.copy()
    bad.pairs = [vba.sensor(vba.camera(abi.CAMERA_PINHOLE, [500.0, 500.0], 640, 480), scenes.left_to_right())]
    bad = bad.copy()
    assert oracle.optimize(bad, vba.default_options()).status == abi.ERR_UNSUPPORTED_CAMERA_MODEL
    bad.general_landmark_ids[:] = 99
    out = oracle.optimize(bad, vba.default_options())
    assert out.status == abi.ERR_MISSING_LANDMARK and out.result.error_id == 99


def test_stereo_recovers_metric_pose_with_single_anchor():
    """bundle_stereo_recovers_metric_pose_with_single_anchor (:1583-1617)."""
    p = scenes.truth_stereo_bundle(0.5, fixed=(10,), poses={30: scenes.pose_at([1.07, 0.04, 0.05])})
    assert oracle.cost(p)[0] > 1e-2
    out = oracle.optimize(p, vba.default_options())
    assert out.status == abi.OK and out.result.final_cost < 1e-12
    c = vba.camera_centres(out.poses)
    assert np.linalg.norm(c[2] - [1.0, 0, 0]) < 1e-7 and np.linalg.norm(c[1] - [0.5, 0, 0]) < 1e-7


def test_stereo_recovers_drifted_landmark():
    """bundle_stereo_recovers_drifted_landmark (:1618-1634)."""
    p = scenes.truth_stereo_bundle(0.5, fixed=(10, 20, 30), landmarks={5: np.array([0.20, 0.10, 6.30])})
    out = oracle.optimize(p, vba.default_options())
    assert out.status == abi.OK and out.result.final_cost < 1e-14
    assert np.linalg.norm(out.landmarks[4] - [0.0, 0.0, 6.0]) < 1e-8


def test_general_stereo_recovers_landmark():
    """bundle_general_stereo_recovers_landmark_with_rotated_right_camera (:1679-1700)."""
    g, truth = scenes.truth_general_stereo_bundle(start=[0.45, 0.08, 4.8])
    assert oracle.cost(g)[0] > 1.0
    out = oracle.optimize(g, vba.default_options())
    assert out.status == abi.OK and out.result.final_cost < 1e-14
    assert np.linalg.norm(out.landmarks[0] - truth) < 1e-8


def test_general_stereo_recovers_pose():
    """bundle_general_stereo_recovers_pose_with_rotated_right_camera (:1701-1755)."""
    g, truth = scenes.general_pose_bundle()
    out = oracle.optimize(g, vba.default_options())
    assert out.status == abi.OK and out.result.final_cost < 1e-12
    assert np.linalg.norm(vba.camera_centres(out.poses)[0] - vba.camera_centres(truth[None])[0]) < 1e-7
    R, Rt = scenes.quat_to_matrix(out.poses[0, :4]), scenes.quat_to_matrix(truth[:4])
    angle = np.arccos(np.clip((np.trace(R @ Rt.T) - 1) / 2, -1, 1))
    assert angle < 1e-7


def test_stereo_dense_matches_matrix_free():
    """bundle_stereo_sparse_solver_matches_dense (:1756-1793), with the
    matrix-free PCG in place of the sparse Cholesky."""
    p = scenes.truth_stereo_bundle(0.5, fixed=(10,), poses={30: scenes.pose_at([1.03, -0.02, 0.04])})
    d = oracle.optimize(p, vba.default_options())
    m = oracle.optimize(p, vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE))
    assert d.status == m.status == abi.OK and d.result.converged == m.result.converged
    assert abs(d.result.final_cost - m.result.final_cost) < 1e-12
    assert np.linalg.norm(vba.camera_centres(d.poses)[2] - vba.camera_centres(m.poses)[2]) < 1e-9


def test_mixed_mono_and_stereo_optimize_together():
    """bundle_mixed_mono_and_stereo_observations_optimize_together (:1794-1853)."""
    out = oracle.optimize(scenes.mixed_mono_stereo_bundle(), vba.default_options())
    assert out.status == abi.OK and out.result.final_cost < 1e-12
    assert np.linalg.norm(vba.camera_centres(out.poses)[2] - [1.5, 0, 0]) < 1e-7


def step_scene():
    """Mono, rectified and general observations together, off the truth."""
    rng = np.random.default_rng(5)
    p = scenes.mixed_mono_stereo_bundle()
    truth20 = scenes.pose_at([0.5, 0.0, 0.0])
    gen = [(20, l, scenes.general_obs(truth20, X), 0) for l, X in scenes.world_grid()]
    p = scenes.with_stereo(p, stereo=[(int(a), int(b), tuple(c)) for a, b, c in
                                      zip(p.stereo_pose_ids, p.stereo_landmark_ids, p.stereo_obs)],
                           baseline=p.stereo_baseline, general=gen,
                           pairs=[vba.sensor(scenes.right_camera(), scenes.left_to_right())])
    p.landmarks = p.landmarks + rng.normal(0, 0.03, p.landmarks.shape)
    return p.copy()


def test_first_lm_step_equals_finite_difference_normal_system():
    p = step_scene()
    lam = 1e-4
    var = [i for i in range(len(p.pose_ids)) if not p.pose_fixed[i]]
    n = 6 * len(var) + 3 * len(p.landmark_ids)
    r0 = textbook_residuals(p)
    rows = sum(len(r) for r in r0)
    J = np.zeros((rows, n))
    h = 1e-6
    for k, i in enumerate(var):
        for a in range(6):
            cols = []
            for s in (h, -h):
                xi = np.zeros(6)
                xi[a] = s
                q = p.copy()
                q.poses[i] = oracle.se3_compose(p.poses[i], oracle.se3_exp(xi))
                cols.append(np.concatenate(textbook_residuals(q)))
            J[:, 6 * k + a] = (cols[0] - cols[1]) / (2 * h)
    for l in range(len(p.landmark_ids)):
        for a in range(3):
            cols = []
            for s in (h, -h):
                q = p.copy()
                q.landmarks[l, a] += s
                cols.append(np.concatenate(textbook_residuals(q)))
            J[:, 6 * len(var) + 3 * l + a] = (cols[0] - cols[1]) / (2 * h)
    r = np.concatenate(r0)
    delta = np.linalg.solve(J.T @ J + lam * np.eye(n), -J.T @ r)
    out = oracle.optimize(p, vba.default_options(max_iterations=1))
    assert out.iterations[0]["step_accepted"]
    assert np.abs(delta).max() > 1e-2  # a step a wrong Jacobian term would visibly move
    for k, i in enumerate(var):
        expect = oracle.se3_compose(p.poses[i], oracle.se3_exp(delta[6 * k:6 * k + 6]))
        assert np.allclose(out.poses[i], expect, atol=1e-8)
    assert np.allclose(out.landmarks, p.landmarks + delta[6 * len(var):].reshape(-1, 3), atol=1e-8)


def test_flattened_order_of_weights_and_residuals():
    """Weights and GNC residuals are indexed mono, rectified, general, rig
    (problem.rs:486-495, 748-760): zero weight on every stereo term leaves
    the mono-only cost."""
    p = step_scene()
    nm, ns, ng = len(p.obs_pose_ids), len(p.stereo_pose_ids), len(p.general_pose_ids)
    w = np.concatenate([np.ones(nm), np.zeros(ns + ng)])
    out = oracle.optimize_weighted(p, vba.default_options(max_iterations=1), w)
    mono = sum(float(r @ r) for r in textbook_residuals(p)[:nm])
    assert out.status == abi.OK and np.isclose(out.result.initial_cost, mono, rtol=1e-12)
    bad = oracle.optimize_weighted(p, vba.default_options(), np.ones(nm + ns))
    assert bad.status == abi.ERR_OBSERVATION_WEIGHT_COUNT
