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


# ---------------------------------------------------------------------------
# Stereo scenes (pipelines/slam/tests/bundle_adjustment.rs:1519-1853)
# ---------------------------------------------------------------------------
def euler_matrix(roll, pitch, yaw):
    """UnitQuaternion::from_euler_angles: R = Rz(yaw) Ry(pitch) Rx(roll)."""
    return rodrigues([0.0, 0.0, yaw]) @ rodrigues([0.0, pitch, 0.0]) @ rodrigues([roll, 0.0, 0.0])


def rectified_right_u(pose, X, baseline, f=500.0, c=320.0):
    """The right camera of a rectified pair sits at +baseline along the left
    camera's x axis: u_r = f (X_c - b) / Z_c + c_x."""
    xc = transform(pose, X)
    return f * (xc[0] - baseline) / xc[2] + c


def with_stereo(problem, stereo=(), baseline=float("nan"), general=(), pairs=()):
    """The problem plus rectified observations [(pose id, lm id, (u_l, v_l,
    u_r))] and general ones [(pose id, lm id, (xl, yl, xr, yr), pair)]."""
    import dataclasses
    return dataclasses.replace(
        problem,
        stereo_pose_ids=np.array([s[0] for s in stereo], np.uint64),
        stereo_landmark_ids=np.array([s[1] for s in stereo], np.uint64),
        stereo_obs=np.array([s[2] for s in stereo], float).reshape(-1, 3), stereo_baseline=baseline,
        general_pose_ids=np.array([g[0] for g in general], np.uint64),
        general_landmark_ids=np.array([g[1] for g in general], np.uint64),
        general_obs=np.array([g[2] for g in general], float).reshape(-1, 4),
        general_pair=np.array([g[3] for g in general], np.uint32), pairs=list(pairs))


STEREO_TRUTH = [(10, [0.0, 0.0, 0.0]), (20, [0.5, 0.0, 0.0]), (30, [1.0, 0.0, 0.0])]


def truth_stereo_bundle(baseline, fixed=(), poses=None, landmarks=None):
    """truth_stereo_bundle (bundle_adjustment.rs:1519-1551): three keyframes
    seeing the six-point grid through a rectified pair; `poses` / `landmarks`
    replace initial states by id."""
    truth = [(k, pose_at(c)) for k, c in STEREO_TRUTH]
    st = []
    for kid, pose in truth:
        for lid, X in world_grid():
            uv = project_pinhole((500.0, 500.0), (320.0, 240.0), transform(pose, X))
            st.append((kid, lid, (uv[0], uv[1], rectified_right_u(pose, X, baseline))))
    init = [(k, (poses or {}).get(k, p)) for k, p in truth]
    lms = [(l, (landmarks or {}).get(l, X)) for l, X in world_grid()]
    p = make_problem(pinhole_camera(), init, lms, [], fixed_poses=fixed)
    return with_stereo(p, stereo=st, baseline=baseline)


def right_camera():
    """The general pair's right camera (bundle_adjustment.rs:1638)."""
    return vba.camera(abi.CAMERA_PINHOLE, [510.0, 495.0, 318.0, 242.0], 640, 480)


def left_to_right():
    """T_right<-left with rotation, vertical translation and unequal
    intrinsics (bundle_adjustment.rs:1641-1644)."""
    return pose_from_rt(euler_matrix(0.01, -0.02, 0.015), np.array([-0.11, 0.002, 0.001]))


def general_obs(pose, X, lr=None):
    lr = left_to_right() if lr is None else lr
    pl = transform(pose, X)
    pr = transform(lr, pl)
    return (*project_pinhole((500.0, 500.0), (320.0, 240.0), pl), *project_pinhole((510.0, 495.0), (318.0, 242.0), pr))


def truth_general_stereo_bundle(start=None):
    """truth_general_stereo_bundle (bundle_adjustment.rs:1636-1668): one fixed
    identity keyframe, landmark 5 at (0.2, -0.1, 4) seen by the general pair;
    `start` replaces the landmark's initial position."""
    pose = np.array([1.0, 0, 0, 0, 0, 0, 0])
    truth = np.array([0.2, -0.1, 4.0])
    p = make_problem(pinhole_camera(), [(10, pose)], [(5, truth if start is None else np.asarray(start, float))], [],
                     fixed_poses=(10,))
    return with_stereo(p, general=[(10, 5, general_obs(pose, truth), 0)],
                       pairs=[vba.sensor(right_camera(), left_to_right())]), truth


def general_pose_bundle():
    """bundle_general_stereo_recovers_pose_with_rotated_right_camera
    (bundle_adjustment.rs:1702-1755): one variable keyframe starting off its
    truth, the six fixed grid landmarks seen through the general pair."""
    truth = pose_at([0.25, -0.03, 0.04], 0.035)
    start = pose_from_rt(euler_matrix(0.025, 0.07, -0.018), np.array([-0.39, 0.12, -0.08]))
    grid = world_grid()
    p = make_problem(pinhole_camera(), [(10, start)], grid, [], fixed_landmarks=[l for l, _ in grid])
    gen = [(10, l, general_obs(truth, X), 0) for l, X in grid]
    return with_stereo(p, general=gen, pairs=[vba.sensor(right_camera(), left_to_right())]), truth


def mixed_mono_stereo_bundle(baseline=0.5):
    """bundle_mixed_mono_and_stereo_observations_optimize_together
    (bundle_adjustment.rs:1795-1853): keyframes 10 and 20 stereo, 40
    monocular and drifted, 10 fixed."""
    truth = [(10, pose_at([0.0, 0.0, 0.0])), (20, pose_at([0.5, 0.0, 0.0])), (40, pose_at([1.5, 0.0, 0.0]))]
    mono, st = [], []
    for kid, pose in truth:
        for lid, X in world_grid():
            uv = project_pinhole((500.0, 500.0), (320.0, 240.0), transform(pose, X))
            if kid == 40:
                mono.append((kid, lid, uv))
            else:
                st.append((kid, lid, (uv[0], uv[1], rectified_right_u(pose, X, baseline))))
    init = truth[:2] + [(40, pose_at([1.55, 0.04, 0.06]))]
    p = make_problem(pinhole_camera(), init, world_grid(), mono, fixed_poses=(10,))
    return with_stereo(p, stereo=st, baseline=baseline)


def split_into_stereo(p, rng, rectified=0.3, general=0.2, baseline=0.12, noise=0.5):
    """A mixed problem from one with monocular observations and a PINHOLE
    camera: a share of them become rectified stereo observations (u_r from
    the current state's depth) and another share general stereo
    observations (the right pixel of the current state through the
    right_camera() pair), both with pixel noise; the rest stay monocular."""
    fx, fy, cx, cy = (p.camera.params[i] for i in range(4))
    pidx = {int(i): k for k, i in enumerate(p.pose_ids)}
    lidx = {int(i): k for k, i in enumerate(p.landmark_ids)}
    n = len(p.obs_pose_ids)
    pk = np.array([pidx[int(i)] for i in p.obs_pose_ids])
    lk = np.array([lidx[int(i)] for i in p.obs_landmark_ids])
    Rs = np.array([quat_to_matrix(q) for q in p.poses[:, :4]])
    xc = np.einsum("nij,nj->ni", Rs[pk], p.landmarks[lk]) + p.poses[pk, 4:7]
    kind = rng.choice(3, size=n, p=[1 - rectified - general, rectified, general])
    kind[xc[:, 2] <= 0.1] = 0
    lr = left_to_right()
    Rlr = quat_to_matrix(lr[:4])
    xr = xc @ Rlr.T + lr[4:7]
    kind[(kind == 2) & (xr[:, 2] <= 0.1)] = 0
    st = [(int(p.obs_pose_ids[k]), int(p.obs_landmark_ids[k]),
           (p.obs_xy[k, 0], p.obs_xy[k, 1], p.obs_xy[k, 0] - fx * baseline / xc[k, 2] + rng.normal(0, noise)))
          for k in np.flatnonzero(kind == 1)]
    ge = [(int(p.obs_pose_ids[k]), int(p.obs_landmark_ids[k]),
           (p.obs_xy[k, 0], p.obs_xy[k, 1], 510.0 * xr[k, 0] / xr[k, 2] + 318.0 + rng.normal(0, noise),
            495.0 * xr[k, 1] / xr[k, 2] + 242.0 + rng.normal(0, noise)), 0)
          for k in np.flatnonzero(kind == 2)]
    import dataclasses
    keep = kind == 0
    mono = dataclasses.replace(p, obs_pose_ids=p.obs_pose_ids[keep], obs_landmark_ids=p.obs_landmark_ids[keep],
                               obs_xy=p.obs_xy[keep])
    return with_stereo(mono, stereo=st, baseline=baseline, general=ge,
                       pairs=[sensor_pair()])


def sensor_pair():
    return vba.sensor(right_camera(), left_to_right())


def add_stereo_from_state(p, rng, n_rect, n_gen, baseline=0.12, noise=0.5):
    """p plus rectified and general stereo observations of (pose, landmark)
    pairs the problem already observes, measured at the current state with
    pixel noise through p's pinhole intrinsics (and right_camera() for the
    general pair): edges then hold observations of several kinds."""
    fx, fy, cx, cy = (p.camera.params[i] for i in range(4))
    pairs = sorted(set(zip(np.concatenate([p.obs_pose_ids, p.rig_pose_ids]).tolist(),
                           np.concatenate([p.obs_landmark_ids, p.rig_landmark_ids]).tolist())))
    pick = rng.choice(len(pairs), size=min(len(pairs), n_rect + n_gen), replace=False)
    P = {int(i): p.poses[k] for k, i in enumerate(p.pose_ids)}
    X = {int(i): p.landmarks[k] for k, i in enumerate(p.landmark_ids)}
    lr = left_to_right()
    st, ge = [], []
    for j, k in enumerate(pick):
        pid, lid = pairs[k]
        xc = transform(P[pid], X[lid])
        xr = transform(lr, xc)
        if xc[2] <= 0.1 or xr[2] <= 0.1:
            continue
        u, v = fx * xc[0] / xc[2] + cx + rng.normal(0, noise), fy * xc[1] / xc[2] + cy + rng.normal(0, noise)
        if j < n_rect:
            st.append((pid, lid, (u, v, fx * (xc[0] - baseline) / xc[2] + cx + rng.normal(0, noise))))
        else:
            ge.append((pid, lid, (u, v, 510.0 * xr[0] / xr[2] + 318.0 + rng.normal(0, noise),
                                  495.0 * xr[1] / xr[2] + 242.0 + rng.normal(0, noise)), 0))
    return with_stereo(p, stereo=st, baseline=baseline, general=ge, pairs=[sensor_pair()])


# ---------------------------------------------------------------------------
# Joint intrinsics scenes (pipelines/slam/tests/bundle_adjustment.rs:78-320,
# 404-433, 864-873, 3645-3715)
# ---------------------------------------------------------------------------
def project_brown(params, Xc):
    """The textbook Brown-Conrady lens: (fx, fy, cx, cy[, k1, k2[, p1, p2]])."""
    p = list(params) + [0.0] * (8 - len(params))
    fx, fy, cx, cy, k1, k2, p1, p2 = p
    x, y = Xc[0] / Xc[2], Xc[1] / Xc[2]
    r2 = x * x + y * y
    rad = 1 + k1 * r2 + k2 * r2 * r2
    xd = x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y * rad + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    return np.array([fx * xd + cx, fy * yd + cy])


def drifted_bundle_through(model, params):
    """drifted_bundle_through (bundle_adjustment.rs:404-433): three views
    rendered through the camera, 10 and 20 fixed, 30 and landmark 5 drifted."""
    truth = [(10, pose_at([0.0, 0.0, 0.0])), (20, pose_at([0.5, 0.0, 0.0])), (30, pose_at([1.0, 0.0, 0.0]))]
    obs = [(k, l, project_brown(params, transform(p, X))) for k, p in truth for l, X in world_grid()]
    poses = truth[:2] + [(30, pose_at([1.05, 0.02, 0.03]))]
    lms = [(l, np.array([0.03, -0.02, 6.1]) if l == 5 else X) for l, X in world_grid()]
    return make_problem(vba.camera(model, params, 640, 480), poses, lms, obs, fixed_poses=(10, 20))


def fixed_structure_bundle_through(model, params):
    """fixed_structure_bundle_through (:864-873): the drifted bundle reset to
    truth, every landmark fixed."""
    p = drifted_bundle_through(model, params)
    p.poses[2] = pose_at([1.0, 0.0, 0.0])
    p.landmarks[4] = [0.0, 0.0, 6.0]
    p.landmark_fixed[:] = 1
    return p


def truth_bundle_through(model, params, start, fixed_structure=True):
    """truth_bundle_distorted / the perturbed-focal scenes (:78-107,
    :151-191): exact three-view data through (model, params), solved from the
    `start` camera with poses 10, 20 and every landmark fixed."""
    truth = [(10, pose_at([0.0, 0.0, 0.0])), (20, pose_at([0.5, 0.0, 0.0])), (30, pose_at([1.0, 0.0, 0.0]))]
    obs = [(k, l, project_brown(params, transform(p, X))) for k, p in truth for l, X in world_grid()]
    fixed_l = [l for l, _ in world_grid()] if fixed_structure else []
    return make_problem(start, truth, world_grid(), obs, fixed_poses=(10, 20), fixed_landmarks=fixed_l)


def perturbed_many_view_bundle(truth_params, start):
    """perturbed_many_view_bundle (:3645-3715): ten yawing views of 80 points
    rendered through the truth camera, poses 100 and 109 fixed, the others
    and every point perturbed; the solve starts from `start`."""
    import oracle
    poses, obs = [], []
    for i in range(10):
        x = i * 0.4 - 1.8
        truth = pose_at([x, 0.1 * (i % 3), 0.0], 0.04 * x)
        if i < 2:
            start_pose = truth
        else:
            s = float(i)
            jit = [0.01 * np.sin(s), -0.01 * np.cos(s), 0.008 * np.sin(2 * s), 0.003 * np.cos(s),
                   -0.002 * np.sin(s), 0.002 * np.cos(3 * s)]
            start_pose = oracle.se3_compose(truth, oracle.se3_exp(jit))
        poses.append((100 + i, truth, start_pose))
    pts = []
    for j in range(80):
        a = float(j)
        pts.append((1000 + j, np.array([np.sin(a * 0.37) * 2.5, np.cos(a * 0.61) * 1.2, 6.0 + np.sin(a * 0.23) * 1.5])))
    init_pts = [(i, X + [0.02 * np.sin(float(i)), -0.02 * np.cos(float(i)), 0.03 * np.sin(0.5 * i)]) for i, X in pts]
    for pid, truth, _ in poses:
        for lid, X in pts:
            xc = transform(truth, X)
            if xc[2] > 0:
                obs.append((pid, lid, project_brown(truth_params, xc)))
    return make_problem(start, [(p, s) for p, _, s in poses], init_pts, obs, fixed_poses=(100, 109))


def elimination_problem(n=120, n_points=2500, seed=3, extras=2, cluster=8, coarse_group=16,
                        no_inverse=0.1):
    """A point elimination of the rig solver's shape (vba.PointEliminationInput):
    n nodes in capture order, the last `extras` of them extra nodes (rig
    sensors / shared cameras) joined to earlier nodes only by direct
    couplings; each point seen by a run of 2-12 consecutive nodes with
    random gaps; random 6x3 camera-point blocks and SPD point inverses, a
    fraction of the points without one (None). Edges are every node pair
    sharing a point plus the extra couplings, sorted."""
    rng = np.random.default_rng(seed)
    n_cam = n - extras
    point_nodes, point_off = [], [0]
    for _ in range(n_points):
        length = int(rng.integers(2, 13))
        start = int(rng.integers(0, n_cam - 1))
        nodes = sorted(set(min(n_cam - 1, start + int(v)) for v in np.cumsum(rng.integers(1, 3, length)) - 1))
        if len(nodes) < 2:
            nodes = [start, start + 1]
        point_nodes.extend(nodes)
        point_off.append(len(point_nodes))
    point_nodes = np.array(point_nodes, np.uint32)
    point_off = np.array(point_off, np.uint64)
    pairs = set()
    for p in range(n_points):
        nd = point_nodes[point_off[p]:point_off[p + 1]]
        for i in range(len(nd)):
            for j in range(i + 1, len(nd)):
                pairs.add((int(nd[i]), int(nd[j])))
    extra_pairs = []
    for x in range(extras):
        for a in sorted(rng.choice(n_cam, size=min(n_cam, 6), replace=False)):
            extra_pairs.append((int(a), n_cam + x))
    pairs |= set(extra_pairs)
    edges = np.array(sorted(pairs), np.uint32)
    row_off = np.searchsorted(edges[:, 0], np.arange(n + 1)).astype(np.uint64)
    slots_by_node = [[] for _ in range(n)]
    for p in range(n_points):
        for k, s in enumerate(point_nodes[point_off[p]:point_off[p + 1]]):
            slots_by_node[int(s)].append((p, k))
    slot_off = np.cumsum([0] + [len(v) for v in slots_by_node]).astype(np.uint64)
    slots = np.array([sl for v in slots_by_node for sl in v], np.uint32).reshape(-1, 2)
    hcp = rng.standard_normal((len(point_nodes), 6, 3))
    a = rng.standard_normal((n_points, 3, 3))
    inv = np.linalg.inv(a @ np.swapaxes(a, 1, 2) + 0.5 * np.eye(3))
    inv = 0.5 * (inv + np.swapaxes(inv, 1, 2))
    has = (rng.uniform(size=n_points) >= no_inverse).astype(np.uint8)
    eid = {tuple(e): i for i, e in enumerate(edges.tolist())}
    extra_edge = np.array([eid[e] for e in extra_pairs], np.uint32)
    extra_blocks = rng.standard_normal((len(extra_pairs), 6, 6))
    run_edges = np.array([i for i, (a_, b_) in enumerate(edges.tolist()) if a_ // cluster == b_ // cluster],
                         np.uint32)
    return vba.PointEliminationInput(
        num_nodes=n, edges=edges, row_off=row_off, slot_off=slot_off, slots=slots, point_off=point_off,
        point_nodes=point_nodes, hcp=hcp, point_inverse=inv, has_inverse=has,
        scale=np.abs(rng.standard_normal((n, 6))) + 1.0, extra_edge=extra_edge, extra_blocks=extra_blocks,
        run_edges=run_edges, coarse=(coarse_group, n_cam) if coarse_group else None,
        structure_id=1000 + seed, linearization_id=1)


def elimination_brute_force(el):
    """The eliminated blocks from the dense Schur complement: with Hcp the
    (6n x 3P) camera-point matrix and Vinv the block-diagonal point
    inverses (zero for points without one), block (t, s) of
    -Hcp Vinv Hcp^T, plus the direct couplings."""
    n = el.num_nodes
    P = len(el.point_off) - 1
    H = np.zeros((6 * n, 3 * P))
    for p in range(P):
        for k in range(int(el.point_off[p]), int(el.point_off[p + 1])):
            s = int(el.point_nodes[k])
            H[6 * s:6 * s + 6, 3 * p:3 * p + 3] = el.hcp[k]
    V = np.zeros((3 * P, 3 * P))
    for p in range(P):
        if el.has_inverse[p]:
            V[3 * p:3 * p + 3, 3 * p:3 * p + 3] = el.point_inverse[p]
    S = -(H @ V @ H.T)
    out = np.array([S[6 * t:6 * t + 6, 6 * s:6 * s + 6] for s, t in el.edges.tolist()])
    for e, blk in zip(el.extra_edge.tolist(), el.extra_blocks):
        out[e] = out[e] + blk
    return out
