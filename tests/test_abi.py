"""C-ABI library checks that need no GPU: the library loads, exports every
symbol its headers declare, declines what it does not implement, and the
seeded generators are deterministic."""
import ctypes as C
import os
import re

import numpy as np

import scenes
import vba
from vba import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(vba_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = abi.library()
    names = declared_functions()
    assert "vba_optimize" in names and "vba_synth_generate" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(abi.SIGNATURES), set(names) - set(abi.SIGNATURES)
    assert lib.vba_abi_version() == abi.ABI_VERSION


def test_struct_sizes_match_c_layout():
    # sizes implied by the C declarations (natural alignment, LP64)
    assert C.sizeof(abi.Camera) == 112
    assert C.sizeof(abi.Sensor) == 168
    assert C.sizeof(abi.IterationStats) == 72
    assert C.sizeof(abi.Result) == 56
    assert C.sizeof(abi.Tuning) == 16
    assert C.sizeof(abi.StereoTerms) == 96
    assert C.sizeof(abi.Problem) == 280 and abi.Problem.stereo.offset == 272


def test_default_options_are_baconfig_default():
    o = vba.default_options()
    assert (o.max_iterations, o.has_initial_lambda, o.initial_lambda) == (20, 1, 1e-4)
    assert (o.lambda_increase_factor, o.lambda_decrease_factor, o.max_lambda, o.min_lambda) == (10.0, 0.1, 1e12, 1e-9)
    assert (o.step_tolerance, o.cost_tolerance, o.has_relative_cost_tolerance) == (1e-7, 1e-9, 0)
    assert o.linear_solver == abi.SOLVER_DENSE and o.robust_kernel == abi.KERNEL_NONE
    assert (o.max_pcg_iterations, o.pcg_relative_tolerance, o.pcg_absolute_tolerance) == (128, 1e-12, 1e-12)


def test_check_declines_what_the_library_does_not_implement():
    ba = scenes.truth_bundle()
    assert vba.check(ba, vba.default_options()) == abi.OK
    assert vba.check(ba, vba.default_options(refine_intrinsics=1)) == abi.DECLINED
    assert vba.check(ba, vba.default_options(robust_kernel=7)) == abi.DECLINED
    for feature in (abi.FEATURE_STEREO_OBSERVATIONS, abi.FEATURE_PRIORS, abi.FEATURE_INERTIAL,
                    abi.FEATURE_FIXED_ROTATIONS, abi.FEATURE_OBSERVATION_WEIGHTS):
        b = ba.copy()
        b.declined_features = feature
        assert vba.check(b, vba.default_options()) == abi.DECLINED
    b = ba.copy()
    b.camera = vba.camera(9, [1.0, 2.0, 3.0, 4.0])
    assert vba.check(b, vba.default_options()) == abi.DECLINED
    assert "declined" in abi.library().vba_last_error(None).decode()


def test_check_declines_out_of_range_option_enums():
    """Every enum of vba_options out of its range declines (the options half
    of the support check, which the prepared solves also apply)."""
    ba = scenes.truth_bundle()
    for field, bad in (("robust_kernel", 3), ("linear_solver", 4), ("linear_solver", 7), ("precision", 2),
                       ("pcg_maxit_policy", 2)):
        assert vba.check(ba, vba.default_options(**{field: bad})) == abi.DECLINED, field
    for field, good in (("linear_solver", abi.SOLVER_SPARSE), ("linear_solver", abi.SOLVER_MATRIX_FREE),
                        ("linear_solver", abi.SOLVER_PCG_EXPLICIT),
                        ("precision", abi.PRECISION_MIXED), ("pcg_maxit_policy", abi.PCG_MAXIT_ACCEPT_ITERATE)):
        assert vba.check(ba, vba.default_options(**{field: good})) == abi.OK, field


def test_generators_are_deterministic_and_sized():
    a = vba.synthesize(abi.SYNTH_LOOP, num_cameras=50, num_landmarks=3000, seed=9, scene_scale=15.0)
    b = vba.synthesize(abi.SYNTH_LOOP, num_cameras=50, num_landmarks=3000, seed=9, scene_scale=15.0)
    c = vba.synthesize(abi.SYNTH_LOOP, num_cameras=50, num_landmarks=3000, seed=10, scene_scale=15.0)
    assert np.array_equal(a.problem.obs_xy, b.problem.obs_xy)
    assert np.array_equal(a.problem.poses, b.problem.poses)
    assert not np.array_equal(a.problem.obs_xy, c.problem.obs_xy)
    assert a.problem.pose_fixed[0] == 1 and a.problem.pose_fixed[1:].sum() == 0
    # every observation of the truth projects in frame, tracks are >= 2 long
    counts = np.bincount(a.problem.obs_landmark_ids.astype(np.int64))
    assert counts[1:].min() >= 2
    mean = a.problem.num_observations / len(a.problem.landmark_ids)
    assert 3.0 < mean < 11.0  # a 15 m loop bends tracks out of view early


def test_full_size_loop_workload_has_3_5m_observations():
    s = vba.synthesize(abi.SYNTH_LOOP)
    p = s.problem
    assert len(p.pose_ids) == 1200 and len(p.landmark_ids) == 400_000
    assert 3.3e6 < p.num_observations < 3.7e6
    assert np.bincount(p.obs_landmark_ids.astype(np.int64))[1:].min() >= 2


def test_circle_workload_observes_every_landmark_in_every_camera():
    """The CIRCLE geometry (radius 10 m, 4 m cube, f=500 at 640x480) puts
    every landmark in frame in every camera: N = cameras x landmarks."""
    s = vba.synthesize(abi.SYNTH_CIRCLE, num_landmarks=500)
    assert len(s.problem.pose_ids) == 50
    assert s.problem.num_observations == 50 * 500


def test_rig_and_city_generators_build_rig_observations():
    s = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=8, num_landmarks=200, seed=4)
    assert len(s.problem.sensors) == 2 and len(s.problem.rig_pose_ids) > 0
    assert s.problem.sensors[1].sensor_from_rig[4] == -0.11
    c = vba.synthesize(abi.SYNTH_CITY, num_cameras=60, num_landmarks=2000, seed=4)
    assert len(c.problem.sensors) == 60 and c.problem.sensors[0].camera.model == abi.CAMERA_OPENCV


def test_create_without_device_reports_an_error_or_succeeds():
    lib = abi.library()
    ctx = lib.vba_create(0)
    if ctx:
        lib.vba_destroy(ctx)
    else:
        assert lib.vba_last_error(None).decode().startswith("vba_create")
