"""The oracle's joint pose / landmark / intrinsics solve, pinned.

Restates the joint-intrinsics tests of pipelines/slam/tests/bundle_adjustment.rs
(:108-320, :875-980) on the CPU oracle, and pins what they leave open with
numpy: the scenes are rendered through a textbook Brown-Conrady lens, and the
first LM step equals a numpy solve of the full damped normal system built from
finite-difference Jacobians of that lens in the poses, the landmarks and the
camera parameters (fx, fy, cx, cy[, k1, k2[, p1, p2]]), with and without
rectified stereo terms and with the shared focal fold.
"""
import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi

PIN = abi.CAMERA_PINHOLE
OCV = abi.CAMERA_OPENCV
R, D, T, SF = abi.REFINE_INTRINSICS, abi.REFINE_DISTORTION, abi.REFINE_TANGENTIAL, abi.REFINE_SHARED_FOCAL


def params(cam):
    return list(cam.params[:cam.num_params])


def solve(p, flags, **kw):
    out, cam = oracle.optimize_intrinsics(p, vba.default_options(refine_intrinsics=flags, **kw))
    assert out.status == abi.OK
    return out, cam


def test_recovers_perturbed_focal_length():
    """refine_intrinsics_recovers_perturbed_focal_length (:151-191)."""
    p = scenes.truth_bundle_through(PIN, [500.0, 500.0, 320.0, 240.0],
                                    vba.camera(PIN, [520.0, 515.0, 326.0, 233.0], 640, 480))
    before = oracle.cost(p)[0]
    out, cam = solve(p, R)
    fx, fy, cx, cy = params(cam)
    assert abs(fx - 500) < 0.5 and abs(fy - 500) < 0.5 and abs(cx - 320) < 0.5 and abs(cy - 240) < 0.5
    assert out.result.final_cost < 0.1 and out.result.final_cost < before
    assert cam.model == PIN and cam.num_params == 4


def test_self_calibrates_radial_distortion():
    """joint_ba_self_calibrates_radial_distortion (:108-149)."""
    p = scenes.truth_bundle_through(PIN, [500.0, 500.0, 320.0, 240.0, -0.05, 0.01], scenes.pinhole_camera())
    out, cam = solve(p, R | D)
    assert cam.num_params == 6
    k1, k2 = params(cam)[4:6]
    assert abs(k1 + 0.05) < 5e-3 and abs(k2 - 0.01) < 5e-3
    assert out.result.final_cost < 0.1


def test_shared_focal_with_distortion():
    """shared_focal_keeps_fx_equal_to_fy_while_recovering_focal_and_distortion (:192-238)."""
    p = scenes.truth_bundle_through(PIN, [500.0, 500.0, 320.0, 240.0, -0.05, 0.01],
                                    vba.camera(PIN, [520.0, 515.0, 326.0, 233.0], 640, 480))
    out, cam = solve(p, R | D | SF)
    fx, fy, cx, cy, k1, k2 = params(cam)
    assert fx == fy and abs(fx - 500) < 0.5 and abs(cx - 320) < 0.5 and abs(cy - 240) < 0.5
    assert abs(k1 + 0.05) < 5e-3 and abs(k2 - 0.01) < 5e-3 and out.result.final_cost < 0.1


def test_shared_focal_fits_one_focal_to_anisotropic_data():
    """shared_focal_fits_the_best_single_focal_to_anisotropic_data (:239-306)."""
    start = vba.camera(PIN, [500.0, 500.0, 320.0, 240.0], 640, 480)
    p = scenes.truth_bundle_through(PIN, [510.0, 490.0, 320.0, 240.0], start)
    out, cam = solve(p, R | SF)
    fx, fy = params(cam)[:2]
    assert fx == fy and 490 < fx < 510 and out.result.final_cost > 1e-3
    out, cam = solve(p, R)
    fx, fy = params(cam)[:2]
    assert abs(fx - 510) < 0.5 and abs(fy - 490) < 0.5


def test_disabled_refinement_leaves_the_camera():
    """refine_intrinsics_is_noop_when_disabled (:307-320), and the plain solve
    for the cases optimize() sends there (optimize.rs:8-26): another lens
    model, rig observations."""
    p = scenes.truth_bundle()
    p.camera = vba.camera(PIN, [520.0, 520.0, 320.0, 240.0], 640, 480)
    out, cam = solve(p, 0)
    assert params(cam) == [520.0, 520.0, 320.0, 240.0]
    plain = oracle.optimize(p, vba.default_options())
    assert out.result.final_cost == plain.result.final_cost and np.array_equal(out.poses, plain.poses)
    q = p.copy()
    q.camera = vba.camera(abi.CAMERA_SIMPLE_RADIAL, [520.0, 320.0, 240.0, 0.0], 640, 480)
    out, cam = solve(q, R)
    plain = oracle.optimize(q, vba.default_options())
    assert params(cam) == [520.0, 320.0, 240.0, 0.0] and out.result.final_cost == plain.result.final_cost


def test_opencv_tangential_terms_held_fixed():
    """joint_intrinsics_keeps_fixed_tangential_terms_of_an_opencv_camera (:875-898)."""
    truth = [500.0, 500.0, 320.0, 240.0, -0.15, 0.05, 0.004, -0.003]
    p = scenes.fixed_structure_bundle_through(OCV, truth)
    p.camera = vba.camera(OCV, [520.0, 515.0, 326.0, 233.0] + truth[4:], 640, 480)
    out, cam = solve(p, R)
    assert out.result.final_cost < 1e-8 and cam.model == OCV
    assert np.allclose(params(cam)[:4], truth[:4], atol=1e-3) and params(cam)[4:] == truth[4:]


def test_tangential_self_calibration():
    """joint_intrinsics_self_calibrates_tangential_distortion_when_asked (:900-952)."""
    truth = [500.0, 500.0, 320.0, 240.0, -0.05, 0.01, 0.003, -0.002]
    start = vba.camera(PIN, [510.0, 505.0, 323.0, 236.0], 640, 480)
    p = scenes.fixed_structure_bundle_through(OCV, truth)
    p.camera = start
    out, cam = solve(p, R | D | T)
    got = params(cam)
    assert cam.model == OCV and cam.num_params == 8 and out.result.final_cost < 1e-8
    assert abs(got[6] - 0.003) < 1e-5 and abs(got[7] + 0.002) < 1e-5
    assert abs(got[4] + 0.05) < 1e-4 and abs(got[5] - 0.01) < 1e-4
    assert np.allclose(got[:4], truth[:4], atol=1e-3)
    radial, rcam = solve(p, R | D)
    assert rcam.model == PIN and radial.result.final_cost > 1e3 * max(out.result.final_cost, 1e-12)


def test_tangential_flag_inert_without_distortion():
    """tangential_refinement_flag_is_inert_without_distortion_refinement (:953-978)."""
    p = scenes.fixed_structure_bundle_through(PIN, [500.0, 500.0, 320.0, 240.0, -0.05, 0.01])
    p.camera = vba.camera(PIN, [510.0, 505.0, 323.0, 236.0], 640, 480)
    a, ca = solve(p, R | T)
    b, cb = solve(p, R)
    assert params(ca) == params(cb) and ca.model == PIN and ca.num_params == 4
    assert a.iterations == b.iterations and np.array_equal(a.poses, b.poses)


def test_many_view_joint_solve_converges():
    """The scene of sparse_joint_intrinsics_matches_dense (:3717-3785): the
    dense joint solve brings the cost down by two orders of magnitude."""
    p = scenes.perturbed_many_view_bundle([500.0, 500.0, 320.0, 240.0],
                                          vba.camera(PIN, [520.0, 512.0, 324.0, 236.0], 640, 480))
    before = oracle.cost(p)[0]
    out, cam = solve(p, R)
    assert out.result.final_cost < 1e-2 * before


# ---------------------------------------------------------------------------
# First LM step against finite differences of the textbook lens
# ---------------------------------------------------------------------------
def residuals(p, cam_params, poses, lms):
    P = {int(i): poses[k] for k, i in enumerate(p.pose_ids)}
    X = {int(i): lms[k] for k, i in enumerate(p.landmark_ids)}
    out = []
    for k in range(len(p.obs_pose_ids)):
        xc = scenes.transform(P[int(p.obs_pose_ids[k])], X[int(p.obs_landmark_ids[k])])
        out.append(scenes.project_brown(cam_params, xc) - p.obs_xy[k])
    for k in range(len(p.stereo_pose_ids)):
        pose, x = P[int(p.stereo_pose_ids[k])], X[int(p.stereo_landmark_ids[k])]
        xc = scenes.transform(pose, x)
        fx, fy, cx, cy = cam_params[:4]
        u, v = fx * xc[0] / xc[2] + cx, fy * xc[1] / xc[2] + cy
        ur = fx * (xc[0] - p.stereo_baseline) / xc[2] + cx
        out.append(np.array([u, v, ur]) - p.stereo_obs[k])
    return np.concatenate(out)


def fd_step(p, cam0, kd, shared, lam):
    var = [i for i in range(len(p.pose_ids)) if not p.pose_fixed[i]]
    lv = [i for i in range(len(p.landmark_ids)) if not p.landmark_fixed[i]]
    n = 6 * len(var) + 3 * len(lv) + kd
    r0 = residuals(p, cam0, p.poses, p.landmarks)
    J = np.zeros((len(r0), n))
    h = 1e-6
    for k, i in enumerate(var):
        for a in range(6):
            cols = []
            for s in (h, -h):
                xi = np.zeros(6)
                xi[a] = s
                q = p.poses.copy()
                q[i] = oracle.se3_compose(p.poses[i], oracle.se3_exp(xi))
                cols.append(residuals(p, cam0, q, p.landmarks))
            J[:, 6 * k + a] = (cols[0] - cols[1]) / (2 * h)
    for k, l in enumerate(lv):
        for a in range(3):
            cols = []
            for s in (h, -h):
                q = p.landmarks.copy()
                q[l, a] += s
                cols.append(residuals(p, cam0, p.poses, q))
            J[:, 6 * len(var) + 3 * k + a] = (cols[0] - cols[1]) / (2 * h)
    for a in range(kd):
        hs = h * max(1.0, abs(cam0[a]))
        cols = []
        for s in (hs, -hs):
            c = list(cam0)
            c[a] += s
            cols.append(residuals(p, c, p.poses, p.landmarks))
        J[:, n - kd + a] = (cols[0] - cols[1]) / (2 * hs)
    A = J.T @ J + lam * np.eye(n)
    g = -J.T @ r0
    if shared:  # fx and fy move together
        f = n - kd
        Tm = np.delete(np.eye(n), f + 1, axis=1)
        Tm[f + 1, f] = 1.0
        delta = Tm @ np.linalg.solve(Tm.T @ A @ Tm, Tm.T @ g)
    else:
        delta = np.linalg.solve(A, g)
    return var, lv, delta


@pytest.mark.parametrize("case", ["k4", "k6_shared", "k8", "k4_stereo"])
def test_first_joint_step_equals_finite_difference_normal_system(case):
    lam = 1e-4
    if case == "k4_stereo":
        p = scenes.mixed_mono_stereo_bundle()
        p.camera = vba.camera(PIN, [506.0, 497.0, 322.0, 238.0], 640, 480)
        p = p.copy()
        cam0, kd, flags = [506.0, 497.0, 322.0, 238.0], 4, R
    else:
        truth = [500.0, 500.0, 320.0, 240.0, -0.05, 0.01, 0.003, -0.002]
        p = scenes.drifted_bundle_through(OCV, truth)
        p.landmark_fixed[:] = 0
        p.landmark_fixed[0] = 1
        if case == "k4":
            p.camera, cam0, kd, flags = vba.camera(OCV, [505.0, 498.0, 323.0, 236.0] + truth[4:], 640, 480), \
                [505.0, 498.0, 323.0, 236.0] + truth[4:], 4, R
        elif case == "k6_shared":
            p.camera, cam0, kd, flags = vba.camera(PIN, [503.0, 503.0, 323.0, 236.0], 640, 480), \
                [503.0, 503.0, 323.0, 236.0, 0.0, 0.0], 6, R | D | SF
        else:
            p.camera, cam0, kd, flags = vba.camera(PIN, [505.0, 498.0, 323.0, 236.0], 640, 480), \
                [505.0, 498.0, 323.0, 236.0, 0.0, 0.0, 0.0, 0.0], 8, R | D | T
        p = p.copy()
    var, lv, delta = fd_step(p, cam0, kd, flags & SF, lam)
    out, cam = oracle.optimize_intrinsics(p, vba.default_options(refine_intrinsics=flags, max_iterations=1))
    assert out.status == abi.OK and out.iterations[0]["step_accepted"]
    assert np.abs(delta[-kd:]).max() > 1e-3  # the camera moves: a wrong camera column would show
    for k, i in enumerate(var):
        expect = oracle.se3_compose(p.poses[i], oracle.se3_exp(delta[6 * k:6 * k + 6]))
        assert np.allclose(out.poses[i], expect, atol=1e-7)
    for k, l in enumerate(lv):
        assert np.allclose(out.landmarks[l], p.landmarks[l] + delta[6 * len(var) + 3 * k:][:3], atol=1e-7)
    got = np.array(params(cam)[:kd])
    assert np.allclose(got, np.array(cam0[:kd]) + delta[-kd:], rtol=0, atol=1e-6 * np.maximum(1, np.abs(got)))
