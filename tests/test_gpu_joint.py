"""The joint pose / landmark / camera solve (vba_optimize_intrinsics) on the
device against the CPU FP64 oracle.

Bar: the same LM accept/reject sequence, final cost within 1e-9 relative
(or both below 1e-12 where the scene is exact), camera parameters within
1e-7 relative, camera centres and landmarks within 1e-7. Scenes: the
reference's joint-intrinsics tests (pipelines/slam/tests/bundle_adjustment.rs
:108-320, :875-980, :3717-3822) and the finite-difference step scenes of the
oracle's own tests, with every camera block size (4, 6, 8), the shared focal,
rectified stereo, Huber and damping."""
import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi

from test_joint_oracle import OCV, PIN, D, R, SF, T, params

pytestmark = pytest.mark.gpu


def compare_joint(gpu, gcam, ref, rcam):
    assert gpu.status == ref.status == abi.OK, gpu.message
    assert gpu.accepted == ref.accepted
    a, b = gpu.result.final_cost, ref.result.final_cost
    assert (a < 1e-12 and b < 1e-12) or abs(a - b) <= 1e-9 * abs(b), (a, b)
    assert abs(gpu.result.initial_cost - ref.result.initial_cost) <= 1e-12 * ref.result.initial_cost
    assert gcam.model == rcam.model and gcam.num_params == rcam.num_params
    np.testing.assert_allclose(params(gcam), params(rcam), rtol=1e-7, atol=1e-9)
    d = np.linalg.norm(vba.camera_centres(gpu.poses) - vba.camera_centres(ref.poses), axis=1).max()
    assert d < 1e-7
    assert np.abs(gpu.landmarks - ref.landmarks).max() < 1e-7


def cases():
    truth = [500.0, 500.0, 320.0, 240.0]
    dist = truth + [-0.05, 0.01]
    ocv = truth + [-0.15, 0.05, 0.004, -0.003]
    tan = truth + [-0.05, 0.01, 0.003, -0.002]
    c = {}
    c["focal"] = (scenes.truth_bundle_through(PIN, truth, vba.camera(PIN, [520.0, 515.0, 326.0, 233.0], 640, 480)),
                  R)
    c["radial"] = (scenes.truth_bundle_through(PIN, dist, scenes.pinhole_camera()), R | D)
    c["shared_radial"] = (scenes.truth_bundle_through(PIN, dist, vba.camera(PIN, [520.0, 515.0, 326.0, 233.0],
                                                                            640, 480)), R | D | SF)
    c["shared_anisotropic"] = (scenes.truth_bundle_through(PIN, [510.0, 490.0, 320.0, 240.0],
                                                           vba.camera(PIN, truth, 640, 480)), R | SF)
    p = scenes.fixed_structure_bundle_through(OCV, ocv)
    p.camera = vba.camera(OCV, [520.0, 515.0, 326.0, 233.0] + ocv[4:], 640, 480)
    c["opencv_fixed_tangential"] = (p.copy(), R)
    p = scenes.fixed_structure_bundle_through(OCV, tan)
    p.camera = vba.camera(PIN, [510.0, 505.0, 323.0, 236.0], 640, 480)
    c["tangential"] = (p.copy(), R | D | T)
    p = scenes.drifted_bundle_through(OCV, tan)
    p.landmark_fixed[:] = 0
    p.landmark_fixed[0] = 1
    p.camera = vba.camera(PIN, [505.0, 498.0, 323.0, 236.0], 640, 480)
    c["free_structure_k8"] = (p.copy(), R | D | T)
    m = scenes.mixed_mono_stereo_bundle()
    m.camera = vba.camera(PIN, [506.0, 497.0, 322.0, 238.0], 640, 480)
    c["mono_and_stereo"] = (m.copy(), R)
    return c


@pytest.mark.parametrize("name", list(cases()))
def test_joint_parity(device, name):
    p, flags = cases()[name]
    o = vba.default_options(refine_intrinsics=flags)
    gpu, gcam = device.optimize_intrinsics(p, o)
    ref, rcam = oracle.optimize_intrinsics(p, o)
    compare_joint(gpu, gcam, ref, rcam)


@pytest.mark.parametrize("config", ["plain", "shared_distortion", "huber_damping"])
def test_many_view_joint_parity(device, config):
    """The scenes and settings of the reference's sparse-vs-dense joint tests
    (bundle_adjustment.rs:3773-3821)."""
    if config == "plain":
        truth, start, o = [500.0, 500.0, 320.0, 240.0], [520.0, 512.0, 324.0, 236.0], \
            vba.default_options(refine_intrinsics=R)
    elif config == "shared_distortion":
        truth, start, o = [500.0, 500.0, 320.0, 240.0, -0.05, 0.01], [515.0, 508.0, 324.0, 236.0], \
            vba.default_options(refine_intrinsics=R | D | SF)
    else:
        truth, start, o = [500.0, 500.0, 320.0, 240.0], [490.0, 507.0, 317.0, 243.0], \
            vba.default_options(refine_intrinsics=R, robust_kernel=abi.KERNEL_HUBER, robust_scale=3.0,
                                initial_lambda=1e-3)
    p = scenes.perturbed_many_view_bundle(truth, vba.camera(PIN, start, 640, 480))
    gpu, gcam = device.optimize_intrinsics(p, o)
    ref, rcam = oracle.optimize_intrinsics(p, o)
    compare_joint(gpu, gcam, ref, rcam)
    assert gpu.result.final_cost < 1e-2 * gpu.result.initial_cost


def test_fallbacks_and_declines(device):
    """optimize() runs the pose / structure solve for another lens model or
    rig observations (optimize.rs:8-26); vba_optimize declines the flags;
    a fully fixed structure still refines the camera."""
    p = scenes.truth_bundle()
    p.camera = vba.camera(abi.CAMERA_SIMPLE_RADIAL, [520.0, 320.0, 240.0, 0.0], 640, 480)
    gpu, gcam = device.optimize_intrinsics(p, vba.default_options(refine_intrinsics=R))
    plain = device.optimize(p, vba.default_options())
    assert gpu.status == abi.OK and params(gcam) == [520.0, 320.0, 240.0, 0.0]
    assert gpu.result.final_cost == plain.result.final_cost and np.array_equal(gpu.poses, plain.poses)
    s = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=6, num_landmarks=200, seed=3)
    gpu, gcam = device.optimize_intrinsics(s.problem, vba.default_options(refine_intrinsics=R))
    plain = device.optimize(s.problem, vba.default_options())
    assert gpu.status == abi.OK and gpu.result.final_cost == plain.result.final_cost
    assert device.optimize(p, vba.default_options(refine_intrinsics=R)).status == abi.DECLINED
    q = scenes.truth_bundle_through(PIN, [500.0, 500.0, 320.0, 240.0],
                                    vba.camera(PIN, [505.0, 503.0, 321.0, 238.0], 640, 480))
    q.pose_fixed[:] = 1
    q = q.copy()
    gpu, gcam = device.optimize_intrinsics(q, vba.default_options(refine_intrinsics=R))
    ref, rcam = oracle.optimize_intrinsics(q, vba.default_options(refine_intrinsics=R))
    compare_joint(gpu, gcam, ref, rcam)
    assert abs(params(gcam)[0] - 500.0) < 1e-2
