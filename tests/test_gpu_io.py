"""Problems on disk through the device solver: a COLMAP model and a problem
dump are read (include/vba_io.h), solved by libvba on the GPU and by the CPU
FP64 oracle on the same view, and the solved state is written back and read
again. The bar is that of the other parity tests: the same LM accept/reject
sequence, final cost within 1e-9 relative (FP64), camera centres within
1e-6 of the scene extent."""
import numpy as np
import pytest

import oracle
import vba
from vba import abi

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def unit_rows(q):
    n = np.sqrt(((q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2]) + q[:, 3] * q[:, 3]) + q[:, 0] * q[:, 0])
    return q / n[:, None]


@pytest.mark.parametrize("binary", [False, True])
def test_colmap_model_solved_on_device_and_written_back(device, tmp_path, binary):
    s = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=12, num_landmarks=400, seed=7)
    vba.Model.from_problem(s.problem).write_colmap(tmp_path / "in", binary=binary)
    model = vba.Model.read_colmap(tmp_path / "in", binary=binary)
    prob, _ = model.problem()
    assert len(prob.pose_ids) == 12 and prob.pose_fixed.sum() == 2  # the SfM's two-pose gauge
    o = vba.default_options(linear_solver=abi.SOLVER_DENSE, max_iterations=15)
    ref = oracle.optimize(prob, o)
    gpu = device.optimize(prob, o)
    compare(gpu, ref, s.scene_extent)
    model.set_state(gpu.poses, gpu.landmarks)
    model.write_colmap(tmp_path / "out", binary=True)
    back, _ = vba.Model.read_colmap(tmp_path / "out", binary=True).problem()
    np.testing.assert_array_equal(back.poses[:, 4:], gpu.poses[:, 4:])
    np.testing.assert_array_equal(back.poses[:, :4], unit_rows(gpu.poses[:, :4]))
    np.testing.assert_array_equal(back.landmarks, gpu.landmarks)


def test_problem_dump_replayed_as_a_local_window(device, tmp_path):
    """ba_gpu_replay with --local 5 (crates/ba-gpu/examples/ba_gpu_replay.rs:
    43-66) on a dumped loop problem, matrix-free PCG as the reference GPU
    backend runs it."""
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=40, num_landmarks=3000, seed=4, scene_scale=16.0)
    p = s.problem.copy()
    p.camera = vba.camera(abi.CAMERA_PINHOLE, [800.0, 800.0, 512.0, 384.0], 1024, 768)
    vba.Model.from_problem(p).write_ba_problem(tmp_path / "dump.txt")
    model = vba.Model.read_ba_problem(tmp_path / "dump.txt")
    model.keep_local_window(5)
    prob, w = model.problem()
    assert w is None and (prob.pose_fixed == 0).sum() == 5
    o = vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_iterations=10, max_pcg_iterations=100,
                            pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                            pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE)
    ref = oracle.optimize(prob, o)
    gpu = device.optimize(prob, o)
    compare(gpu, ref, s.scene_extent, cost_rtol=1e-6)


def test_weighted_dump_is_declined(device, tmp_path):
    """Observation weights are declined, as the SfM sends weighted solves to
    its CPU solver (refinement.rs:469-473)."""
    s = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=6, num_landmarks=60, seed=2)
    p = s.problem.copy()
    p.camera = vba.camera(abi.CAMERA_PINHOLE, [500.0, 500.0, 320.0, 240.0], 640, 480)
    vba.Model.from_problem(p, np.ones(len(p.obs_pose_ids))).write_ba_problem(tmp_path / "w.txt")
    prob, w = vba.Model.read_ba_problem(tmp_path / "w.txt").problem()
    assert w is not None
    out = device.optimize(prob, vba.default_options())
    assert out.status == abi.DECLINED


def test_opencv_known_answer_on_device(device):
    """The cv2.projectPoints pixel of crates/io/tests/colmap.rs:651-668 is a
    zero residual for the device's OPENCV projection too."""
    cam = vba.camera(abi.CAMERA_OPENCV, [1353.09, 1338.03, 962.7, 539.93, 0.0346, -0.0235, 0.0012, -0.0008],
                     1920, 1080)
    uv = np.array([1643.9695843529998, 136.0453501334])
    p = vba.BundleProblem(
        camera=cam, pose_ids=[1, 2], poses=[[1, 0, 0, 0, 0, 0, 0], [1, 0, 0, 0, 0, 0, 0]], pose_fixed=[1, 0],
        landmark_ids=[9], landmarks=[[0.5, -0.3, 1.0]], landmark_fixed=[0], obs_pose_ids=[1, 2],
        obs_landmark_ids=[9, 9], obs_xy=[uv, uv])
    out = device.optimize(p, vba.default_options(max_iterations=1))
    assert out.status == abi.OK, out.message
    assert out.result.initial_cost < 2 * (1e-8) ** 2 * 2
