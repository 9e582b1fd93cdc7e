"""Small scenes of the reference's tests and independent numpy geometry.

The scenes restate the fixtures of pipelines/slam/tests/bundle_adjustment.rs
(pinhole 640x480, f=500, c=(320,240), :17-77 and :1312-1367). The numpy
geometry below is written from the textbook definitions (unit quaternion
rotation, pinhole / radial projection), not from the oracle's code, so tests
that compare against it pin the oracle.
"""
from __future__ import annotations

import numpy as np

import vba
from vba import abi


def quat_to_matrix(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def matrix_to_quat(R):
    w = np.sqrt(max(0.0, 1 + R[0, 0] + R[1, 1] + R[2, 2])) / 2
    x = np.copysign(np.sqrt(max(0.0, 1 + R[0, 0] - R[1, 1] - R[2, 2])) / 2, R[2, 1] - R[1, 2])
    y = np.copysign(np.sqrt(max(0.0, 1 - R[0, 0] + R[1, 1] - R[2, 2])) / 2, R[0, 2] - R[2, 0])
    z = np.copysign(np.sqrt(max(0.0, 1 - R[0, 0] - R[1, 1] + R[2, 2])) / 2, R[1, 0] - R[0, 1])
    q = np.array([w, x, y, z])
    return q / np.linalg.norm(q)


def rodrigues(w):
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K


def pose_from_rt(R, t):
    return np.concatenate([matrix_to_quat(R), t])


def pose_at(centre, yaw=0.0):
    """pose_at / pose_with_yaw (bundle_adjustment.rs:21-31): rotation about y."""
    R = rodrigues(np.array([0.0, yaw, 0.0]))
    return pose_from_rt(R, -R @ np.asarray(centre, float))


def transform(pose, X):
    return quat_to_matrix(pose[:4]) @ X + pose[4:7]


def project_pinhole(f, c, Xc):
    return np.array([f[0] * Xc[0] / Xc[2] + c[0], f[1] * Xc[1] / Xc[2] + c[1]])


PINHOLE = dict(model=abi.CAMERA_PINHOLE, params=[500.0, 500.0, 320.0, 240.0])


def pinhole_camera():
    return vba.camera(abi.CAMERA_PINHOLE, [500.0, 500.0, 320.0, 240.0], 640, 480)


def world_grid():
    return [
        (1, np.array([-1.0, -0.5, 5.0])), (2, np.array([1.0, -0.5, 5.0])), (3, np.array([-1.0, 0.5, 5.0])),
        (4, np.array([1.0, 0.5, 5.0])), (5, np.array([0.0, 0.0, 6.0])), (6, np.array([0.5, 0.25, 7.0])),
    ]


def make_problem(cam, poses, landmarks, observations, fixed_poses=(), fixed_landmarks=()):
    """poses: [(id, pose7)], landmarks: [(id, xyz)], observations: [(pose id, lm id, uv)]"""
    pose_ids = np.array([p[0] for p in poses], np.uint64)
    lm_ids = np.array([l[0] for l in landmarks], np.uint64)
    return vba.BundleProblem(
        camera=cam,
        pose_ids=pose_ids,
        poses=np.array([p[1] for p in poses]),
        pose_fixed=np.array([1 if p[0] in fixed_poses else 0 for p in poses], np.uint8),
        landmark_ids=lm_ids,
        landmarks=np.array([l[1] for l in landmarks]),
        landmark_fixed=np.array([1 if l[0] in fixed_landmarks else 0 for l in landmarks], np.uint8),
        obs_pose_ids=np.array([o[0] for o in observations], np.uint64),
        obs_landmark_ids=np.array([o[1] for o in observations], np.uint64),
        obs_xy=np.array([o[2] for o in observations]).reshape(-1, 2),
    )


def truth_bundle(fixed=(10, 20), drift30=None):
    """truth_bundle (bundle_adjustment.rs:47-77), optional drifted KF30."""
    truth = [(10, pose_at([0.0, 0.0, 0.0])), (20, pose_at([0.5, 0.0, 0.0])), (30, pose_at([1.0, 0.0, 0.0]))]
    obs = []
    for kid, pose in truth:
        for lid, X in world_grid():
            obs.append((kid, lid, project_pinhole((500.0, 500.0), (320.0, 240.0), transform(pose, X))))
    poses = list(truth)
    if drift30 is not None:
        poses[2] = (30, pose_at(drift30))
    return make_problem(pinhole_camera(), poses, world_grid(), obs, fixed_poses=fixed)


def over_determined_bundle(inject_outlier: bool):
    """over_determined_bundle_with_optional_outlier (bundle_adjustment.rs:1312-1367)."""
    truth = [(10 * (i + 1), pose_at([0.5 * i, 0.0, 0.0])) for i in range(5)]
    lms = []
    for r in range(3):
        for c in range(4):
            lid = r * 4 + c + 1
            lms.append((lid, np.array([-1.5 + c, -1.0 + r, 5.0 + 0.2 * np.sin(r * 4 + c)])))
    obs = []
    for kid, pose in truth:
        for lid, X in lms:
            uv = project_pinhole((500.0, 500.0), (320.0, 240.0), transform(pose, X))
            if inject_outlier and kid == 50 and lid == 1:
                uv = uv + 50.0
            obs.append((kid, lid, uv))
    return make_problem(pinhole_camera(), truth, lms, obs, fixed_poses=(10, 20, 30, 40))


def gnc_bundle(outliers=(), shift=0.0):
    """build_with_outliers (pipelines/slam/tests/gnc_robust_ba.rs:18-106): five
    yawing keyframes 10..14 at truth, a 3x4 landmark grid 4.5-7 m ahead seen
    by every keyframe, the end poses fixed; observations whose row-major
    insertion index is in `outliers` are shifted by (+shift, -shift) px."""
    truth = [(10 + i, pose_at([0.3 * i, 0.0, 0.04 * i], 0.015 * i)) for i in range(5)]
    lms = []
    for r in range(3):
        for c in range(4):
            lms.append((r * 4 + c + 1, np.array([-1.2 + 0.8 * c, -0.8 + 0.8 * r, 4.5 + 0.4 * (r + c)])))
    obs = []
    idx = 0
    for kid, pose in truth:
        for lid, X in lms:
            uv = project_pinhole((500.0, 500.0), (320.0, 240.0), transform(pose, X))
            if idx in outliers:
                uv = uv + np.array([shift, -shift])
            obs.append((kid, lid, uv))
            idx += 1
    return make_problem(pinhole_camera(), truth, lms, obs, fixed_poses=(10, 14)), truth
