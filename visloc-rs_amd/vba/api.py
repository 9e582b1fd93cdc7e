"""Python host side of the bundle-adjustment library.

`BundleProblem` owns numpy arrays laid out as the C view `vba_problem`
(the caller-side image of visloc-rs `BundleAdjustment`,
pipelines/slam/src/bundle/mod.rs:529-622); `Device.optimize` is
`BaAccelerator::optimize` (pipelines/slam/src/ba_accel.rs:22-32) through the
C ABI and returns the reference's `BaResult` fields.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import abi


def camera(model: int, params, width: int = 0, height: int = 0) -> abi.Camera:
    cam = abi.Camera()
    cam.model = model
    cam.width = width
    cam.height = height
    params = list(params)
    cam.num_params = len(params)
    for i, v in enumerate(params):
        cam.params[i] = float(v)
    return cam


def sensor(cam: abi.Camera, transform) -> abi.Sensor:
    """A rig sensor (camera, T_sensor<-rig) or a general stereo pair (right
    camera, T_right<-left); the transform is (qw qx qy qz tx ty tz)."""
    s = abi.Sensor()
    s.camera = cam
    for i, v in enumerate(np.asarray(transform, np.float64).reshape(7)):
        s.sensor_from_rig[i] = float(v)
    return s


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None or a.size == 0:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class BundleProblem:
    camera: abi.Camera
    pose_ids: np.ndarray           # uint64 [P]
    poses: np.ndarray              # float64 [P, 7] (qw qx qy qz tx ty tz), world->camera
    pose_fixed: np.ndarray         # uint8 [P]
    landmark_ids: np.ndarray       # uint64 [L]
    landmarks: np.ndarray          # float64 [L, 3]
    landmark_fixed: np.ndarray     # uint8 [L]
    obs_pose_ids: np.ndarray       # uint64 [N]
    obs_landmark_ids: np.ndarray   # uint64 [N]
    obs_xy: np.ndarray             # float64 [N, 2]
    rig_pose_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    rig_landmark_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    rig_xy: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    rig_sensor: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    sensors: List[abi.Sensor] = field(default_factory=list)
    declined_features: int = 0
    # rectified stereo (BaStereoObservation) and its baseline
    stereo_pose_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    stereo_landmark_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    stereo_obs: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))      # u_l, v_l, u_r
    stereo_baseline: float = float("nan")                                         # None in the reference
    # general stereo (BaGeneralStereoObservation): pair q = (right camera,
    # T_right<-left) as an abi.Sensor
    general_pose_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    general_landmark_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    general_obs: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))     # left x, y, right x, y
    general_pair: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    pairs: List[abi.Sensor] = field(default_factory=list)

    def __post_init__(self):
        self.pose_ids = np.ascontiguousarray(self.pose_ids, np.uint64)
        self.poses = np.ascontiguousarray(self.poses, np.float64).reshape(-1, 7)
        self.pose_fixed = np.ascontiguousarray(self.pose_fixed, np.uint8)
        self.landmark_ids = np.ascontiguousarray(self.landmark_ids, np.uint64)
        self.landmarks = np.ascontiguousarray(self.landmarks, np.float64).reshape(-1, 3)
        self.landmark_fixed = np.ascontiguousarray(self.landmark_fixed, np.uint8)
        self.obs_pose_ids = np.ascontiguousarray(self.obs_pose_ids, np.uint64)
        self.obs_landmark_ids = np.ascontiguousarray(self.obs_landmark_ids, np.uint64)
        self.obs_xy = np.ascontiguousarray(self.obs_xy, np.float64).reshape(-1, 2)
        self.rig_pose_ids = np.ascontiguousarray(self.rig_pose_ids, np.uint64)
        self.rig_landmark_ids = np.ascontiguousarray(self.rig_landmark_ids, np.uint64)
        self.rig_xy = np.ascontiguousarray(self.rig_xy, np.float64).reshape(-1, 2)
        self.rig_sensor = np.ascontiguousarray(self.rig_sensor, np.uint32)
        self._sensor_array = (abi.Sensor * max(1, len(self.sensors)))(*self.sensors)
        self.stereo_pose_ids = np.ascontiguousarray(self.stereo_pose_ids, np.uint64)
        self.stereo_landmark_ids = np.ascontiguousarray(self.stereo_landmark_ids, np.uint64)
        self.stereo_obs = np.ascontiguousarray(self.stereo_obs, np.float64).reshape(-1, 3)
        self.general_pose_ids = np.ascontiguousarray(self.general_pose_ids, np.uint64)
        self.general_landmark_ids = np.ascontiguousarray(self.general_landmark_ids, np.uint64)
        self.general_obs = np.ascontiguousarray(self.general_obs, np.float64).reshape(-1, 4)
        self.general_pair = np.ascontiguousarray(self.general_pair, np.uint32)
        self._pair_array = (abi.Sensor * max(1, len(self.pairs)))(*self.pairs)

    @property
    def has_stereo(self) -> bool:
        return len(self.stereo_pose_ids) > 0 or len(self.general_pose_ids) > 0

    @property
    def num_observations(self) -> int:
        """All observations, in the flattened order mono, rectified stereo,
        general stereo, rig."""
        return (len(self.obs_pose_ids) + len(self.stereo_pose_ids) + len(self.general_pose_ids) +
                len(self.rig_pose_ids))

    def view(self) -> abi.Problem:
        """The C view; valid while this object is alive and unmodified."""
        v = abi.Problem()
        v.camera = self.camera
        v.num_poses = len(self.pose_ids)
        v.pose_ids = _ptr(self.pose_ids, C.c_uint64)
        v.poses = _ptr(self.poses, C.c_double)
        v.pose_fixed = _ptr(self.pose_fixed, C.c_uint8)
        v.num_landmarks = len(self.landmark_ids)
        v.landmark_ids = _ptr(self.landmark_ids, C.c_uint64)
        v.landmarks = _ptr(self.landmarks, C.c_double)
        v.landmark_fixed = _ptr(self.landmark_fixed, C.c_uint8)
        v.num_observations = len(self.obs_pose_ids)
        v.obs_pose_ids = _ptr(self.obs_pose_ids, C.c_uint64)
        v.obs_landmark_ids = _ptr(self.obs_landmark_ids, C.c_uint64)
        v.obs_xy = _ptr(self.obs_xy, C.c_double)
        v.num_rig_observations = len(self.rig_pose_ids)
        v.rig_pose_ids = _ptr(self.rig_pose_ids, C.c_uint64)
        v.rig_landmark_ids = _ptr(self.rig_landmark_ids, C.c_uint64)
        v.rig_xy = _ptr(self.rig_xy, C.c_double)
        v.rig_sensor = _ptr(self.rig_sensor, C.c_uint32)
        v.num_sensors = len(self.sensors)
        v.sensors = C.cast(self._sensor_array, C.POINTER(abi.Sensor)) if self.sensors else None
        v.declined_features = self.declined_features
        if self.has_stereo or np.isfinite(self.stereo_baseline):
            # one StereoTerms per problem, refreshed in place, so that every
            # view handed out keeps pointing at live memory
            t = self.__dict__.setdefault("_stereo_terms", abi.StereoTerms())
            t.baseline = self.stereo_baseline
            t.num_stereo = len(self.stereo_pose_ids)
            t.stereo_pose_ids = _ptr(self.stereo_pose_ids, C.c_uint64)
            t.stereo_landmark_ids = _ptr(self.stereo_landmark_ids, C.c_uint64)
            t.stereo_obs = _ptr(self.stereo_obs, C.c_double)
            t.num_general = len(self.general_pose_ids)
            t.general_pose_ids = _ptr(self.general_pose_ids, C.c_uint64)
            t.general_landmark_ids = _ptr(self.general_landmark_ids, C.c_uint64)
            t.general_obs = _ptr(self.general_obs, C.c_double)
            t.general_pair = _ptr(self.general_pair, C.c_uint32)
            t.num_pairs = len(self.pairs)
            t.pairs = C.cast(self._pair_array, C.POINTER(abi.Sensor)) if self.pairs else None
            v.stereo = C.pointer(t)
        return v

    def copy(self) -> "BundleProblem":
        return BundleProblem(
            camera=self.camera, pose_ids=self.pose_ids.copy(), poses=self.poses.copy(),
            pose_fixed=self.pose_fixed.copy(), landmark_ids=self.landmark_ids.copy(),
            landmarks=self.landmarks.copy(), landmark_fixed=self.landmark_fixed.copy(),
            obs_pose_ids=self.obs_pose_ids.copy(), obs_landmark_ids=self.obs_landmark_ids.copy(),
            obs_xy=self.obs_xy.copy(), rig_pose_ids=self.rig_pose_ids.copy(),
            rig_landmark_ids=self.rig_landmark_ids.copy(), rig_xy=self.rig_xy.copy(),
            rig_sensor=self.rig_sensor.copy(), sensors=list(self.sensors),
            declined_features=self.declined_features, stereo_pose_ids=self.stereo_pose_ids.copy(),
            stereo_landmark_ids=self.stereo_landmark_ids.copy(), stereo_obs=self.stereo_obs.copy(),
            stereo_baseline=self.stereo_baseline, general_pose_ids=self.general_pose_ids.copy(),
            general_landmark_ids=self.general_landmark_ids.copy(), general_obs=self.general_obs.copy(),
            general_pair=self.general_pair.copy(), pairs=list(self.pairs))

    def with_state(self, poses: np.ndarray, landmarks: np.ndarray) -> "BundleProblem":
        p = self.copy()
        p.poses = np.ascontiguousarray(poses, np.float64).reshape(-1, 7)
        p.landmarks = np.ascontiguousarray(landmarks, np.float64).reshape(-1, 3)
        return p


def _np_from(ptr, n, dtype, shape):
    if n == 0 or not ptr:
        return np.zeros(shape, dtype)
    buf = np.ctypeslib.as_array(ptr, shape=(int(np.prod(shape)),))
    return np.array(buf, dtype=dtype).reshape(shape)


@dataclass
class Synthetic:
    problem: BundleProblem
    truth_poses: np.ndarray
    truth_landmarks: np.ndarray
    scene_extent: float
    params: abi.SynthParams


def synth_params(kind: int, **overrides) -> abi.SynthParams:
    lib = abi.library()
    p = abi.SynthParams()
    lib.vba_synth_default(kind, C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def synthesize(kind: int, **overrides) -> Synthetic:
    """Generate a seeded synthetic problem (include/vba_synth.h) and copy it
    into numpy arrays."""
    lib = abi.library()
    params = synth_params(kind, **overrides)
    h = lib.vba_synth_generate(C.byref(params))
    if not h:
        raise ValueError(lib.vba_last_error(None).decode())
    try:
        v = abi.Problem()
        lib.vba_synth_view(h, C.byref(v))
        P, L = v.num_poses, v.num_landmarks
        N, R = v.num_observations, v.num_rig_observations
        sensors = [abi.Sensor.from_buffer_copy(v.sensors[i]) for i in range(v.num_sensors)]
        prob = BundleProblem(
            camera=abi.Camera.from_buffer_copy(v.camera),
            pose_ids=_np_from(v.pose_ids, P, np.uint64, (P,)),
            poses=_np_from(v.poses, P, np.float64, (P, 7)),
            pose_fixed=_np_from(v.pose_fixed, P, np.uint8, (P,)),
            landmark_ids=_np_from(v.landmark_ids, L, np.uint64, (L,)),
            landmarks=_np_from(v.landmarks, L, np.float64, (L, 3)),
            landmark_fixed=_np_from(v.landmark_fixed, L, np.uint8, (L,)),
            obs_pose_ids=_np_from(v.obs_pose_ids, N, np.uint64, (N,)),
            obs_landmark_ids=_np_from(v.obs_landmark_ids, N, np.uint64, (N,)),
            obs_xy=_np_from(v.obs_xy, N, np.float64, (N, 2)),
            rig_pose_ids=_np_from(v.rig_pose_ids, R, np.uint64, (R,)),
            rig_landmark_ids=_np_from(v.rig_landmark_ids, R, np.uint64, (R,)),
            rig_xy=_np_from(v.rig_xy, R, np.float64, (R, 2)),
            rig_sensor=_np_from(v.rig_sensor, R, np.uint32, (R,)),
            sensors=sensors)
        tp, tl = abi.P64(), abi.P64()
        lib.vba_synth_truth(h, C.byref(tp), C.byref(tl))
        truth_poses = _np_from(tp, P, np.float64, (P, 7))
        truth_lms = _np_from(tl, L, np.float64, (L, 3))
        extent = lib.vba_synth_scene_extent(h)
    finally:
        lib.vba_synth_free(h)
    return Synthetic(prob, truth_poses, truth_lms, extent, params)


def default_gnc_options(**overrides) -> abi.GncOptions:
    """GncConfig::default (vba_default_gnc_options) with field overrides."""
    g = abi.GncOptions()
    abi.library().vba_default_gnc_options(C.byref(g))
    for k, v in overrides.items():
        setattr(g, k, v)
    return g


def default_options(**overrides) -> abi.Options:
    """BaConfig::default() (+ MatrixFreeBaOptions::default()), with overrides."""
    o = abi.Options()
    abi.library().vba_default_options(C.byref(o))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


@dataclass
class SolveOutput:
    status: int
    result: abi.Result
    iterations: List[dict]
    poses: Optional[np.ndarray]
    landmarks: Optional[np.ndarray]
    message: str = ""

    @property
    def accepted(self) -> List[bool]:
        return [bool(it["step_accepted"]) for it in self.iterations]


def stats_list(stats, n) -> List[dict]:
    out = []
    for i in range(n):
        s = stats[i]
        out.append({f[0].rstrip("_"): getattr(s, f[0]) for f in abi.IterationStats._fields_})
    return out


def run_solver(fn, problem: BundleProblem, options: abi.Options, cap: int = 1024, **extra) -> SolveOutput:
    poses = np.empty_like(problem.poses)
    lms = np.empty_like(problem.landmarks)
    stats = (abi.IterationStats * cap)()
    res = abi.Result()
    view = problem.view()
    st = fn(view, options, poses, lms, stats, cap, res, **extra)
    n = min(res.num_iterations, cap)
    ok = st == abi.OK
    return SolveOutput(st, res, stats_list(stats, n) if ok else [], poses if ok else None, lms if ok else None)


@dataclass
class PointEliminationInput:
    """`PointElimination` (pipelines/slam/src/ba_accel.rs:71-121) as numpy
    arrays: edges (E, 2) sorted (s, t), s < t; row_off (n + 1); slot_off
    (n + 1) and slots (S, 2) = (point, k), ascending point per node;
    point_off (P + 1) and point_nodes (ascending per point); hcp (H, 6, 3)
    per (point, k), row-major here; point_inverse (P, 3, 3) with
    has_inverse (P) for Some / None; scale (n, 6); extra as (edge ids,
    (X, 6, 6) blocks); run_edges; coarse as (group, nodes) or None."""
    num_nodes: int
    edges: np.ndarray
    row_off: np.ndarray
    slot_off: np.ndarray
    slots: np.ndarray
    point_off: np.ndarray
    point_nodes: np.ndarray
    hcp: np.ndarray
    point_inverse: np.ndarray
    has_inverse: np.ndarray
    scale: np.ndarray
    extra_edge: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    extra_blocks: np.ndarray = field(default_factory=lambda: np.zeros((0, 6, 6)))
    run_edges: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    coarse: Optional[tuple] = None
    structure_id: int = 1
    linearization_id: int = 1

    def coarse_aggregates(self) -> int:
        if self.coarse is None:
            return 0
        group, nodes = self.coarse
        return -(-nodes // group)

    def view(self):
        """(abi.PointElimination, arrays to keep alive while it is used)."""
        col = lambda m: np.ascontiguousarray(np.swapaxes(np.asarray(m, np.float64), -1, -2))
        arrs = dict(
            edges=np.ascontiguousarray(np.asarray(self.edges, np.uint32).reshape(-1, 2)),
            row_off=np.ascontiguousarray(self.row_off, np.uint64),
            slot_off=np.ascontiguousarray(self.slot_off, np.uint64),
            slots=np.ascontiguousarray(np.asarray(self.slots, np.uint32).reshape(-1, 2)),
            point_off=np.ascontiguousarray(self.point_off, np.uint64),
            point_nodes=np.ascontiguousarray(self.point_nodes, np.uint32),
            hcp=col(np.asarray(self.hcp).reshape(-1, 6, 3)),
            point_inverse=col(np.asarray(self.point_inverse).reshape(-1, 3, 3)),
            has_inverse=np.ascontiguousarray(self.has_inverse, np.uint8),
            scale=np.ascontiguousarray(np.asarray(self.scale, np.float64).reshape(-1, 6)),
            extra_edge=np.ascontiguousarray(self.extra_edge, np.uint32),
            extra_blocks=col(np.asarray(self.extra_blocks).reshape(-1, 6, 6)),
            run_edges=np.ascontiguousarray(self.run_edges, np.uint32),
        )
        v = abi.PointElimination()
        v.structure_id = self.structure_id
        v.linearization_id = self.linearization_id
        v.num_nodes = self.num_nodes
        v.num_edges = len(arrs["edges"])
        v.num_points = len(arrs["point_off"]) - 1
        v.num_extra = len(arrs["extra_edge"])
        v.num_run_edges = len(arrs["run_edges"])
        if self.coarse is not None:
            v.coarse_group, v.coarse_nodes = int(self.coarse[0]), int(self.coarse[1])
        ptr = {np.dtype(np.uint32): abi.PU32, np.dtype(np.uint64): abi.PU64, np.dtype(np.uint8): abi.PU8,
               np.dtype(np.float64): abi.P64}
        for k, a in arrs.items():
            setattr(v, k, a.ctypes.data_as(ptr[a.dtype]))
        return v, arrs


class Device:
    """A libvba context on one HIP device (vba_create / vba_destroy)."""

    def __init__(self, device: int = 0):
        self.lib = abi.library()
        self.ctx = self.lib.vba_create(device)
        if not self.ctx:
            raise RuntimeError("vba_create failed: " + self.lib.vba_last_error(None).decode())

    def close(self):
        if self.ctx:
            self.lib.vba_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        return self.lib.vba_last_error(self.ctx).decode()

    def prepare(self, problem: BundleProblem, options: abi.Options) -> "PreparedProblem":
        return PreparedProblem(self, problem, options)

    def profile(self, enable: bool = True, reset: bool = True, stride: int = 1):
        """Record the device time of every `stride`-th launch of each kernel
        family (vba_profile_enable)."""
        if reset:
            self.lib.vba_profile_reset(self.ctx)
        self.lib.vba_profile_enable(self.ctx, max(1, int(stride)) if enable else 0)

    def set_tuning(self, pcg_chunk: int = 10, pcg_lookahead: int = 0, pcg_predict: bool = False,
                   apply_waves_per_cu: int = 0) -> None:
        """The PCG launch schedule of this context (vba_set_tuning)."""
        t = abi.Tuning(int(pcg_chunk), int(pcg_lookahead), int(bool(pcg_predict)), int(apply_waves_per_cu))
        st = self.lib.vba_set_tuning(self.ctx, C.byref(t))
        if st != abi.OK:
            raise ValueError(self.last_error())

    def tuning(self) -> dict:
        t = abi.Tuning()
        self.lib.vba_get_tuning(self.ctx, C.byref(t))
        return t.as_dict()

    def profile_query(self, kernel: str) -> dict:
        launches = C.c_uint64()
        secs = C.c_double()
        nbytes = C.c_double()
        self.lib.vba_profile_query(self.ctx, kernel.encode(), C.byref(launches), C.byref(secs), C.byref(nbytes))
        return {"launches": int(launches.value), "seconds": float(secs.value), "bytes": float(nbytes.value)}

    def optimize(self, problem: BundleProblem, options: abi.Options, cap: int = 1024) -> SolveOutput:
        def call(view, opts, poses, lms, stats, cap_, res):
            return self.lib.vba_optimize(
                self.ctx, C.byref(view), C.byref(opts), poses.ctypes.data_as(abi.P64),
                lms.ctypes.data_as(abi.P64), stats, cap_, C.byref(res))

        out = run_solver(call, problem, options, cap)
        if out.status != abi.OK:
            out.message = self.last_error()
        return out

    def optimize_intrinsics(self, problem: BundleProblem, options: abi.Options, cap: int = 1024):
        """vba_optimize_intrinsics (BundleAdjustment::optimize with
        options.refine_intrinsics = abi.REFINE_* bits): returns the solve and
        the refined abi.Camera."""
        cam = abi.Camera()

        def call(view, opts, poses, lms, stats, cap_, res):
            return self.lib.vba_optimize_intrinsics(
                self.ctx, C.byref(view), C.byref(opts), poses.ctypes.data_as(abi.P64),
                lms.ctypes.data_as(abi.P64), C.byref(cam), stats, cap_, C.byref(res))

        out = run_solver(call, problem, options, cap)
        if out.status != abi.OK:
            out.message = self.last_error()
        return out, cam

    def optimize_weighted(self, problem: BundleProblem, options: abi.Options, weights, cap: int = 1024) -> SolveOutput:
        """vba_optimize_weighted: `weights`, one per observation in the
        flattened order (monocular, rectified stereo, general stereo, rig), multiply the robust-kernel weights
        (BundleAdjustment::optimize_with_observation_weights)."""
        w = np.ascontiguousarray(weights, np.float64)

        def call(view, opts, poses, lms, stats, cap_, res):
            return self.lib.vba_optimize_weighted(
                self.ctx, C.byref(view), C.byref(opts), w.ctypes.data_as(abi.P64), len(w),
                poses.ctypes.data_as(abi.P64), lms.ctypes.data_as(abi.P64), stats, cap_, C.byref(res))

        out = run_solver(call, problem, options, cap)
        if out.status != abi.OK:
            out.message = self.last_error()
        return out

    def optimize_gnc(self, problem: BundleProblem, options: abi.Options, gnc: abi.GncOptions):
        """vba_optimize_gnc (BundleAdjustment::optimize_gnc): (status,
        GncResult, poses, landmarks, weights)."""
        view = problem.view()
        poses = np.empty_like(problem.poses)
        lms = np.empty_like(problem.landmarks)
        w = np.empty(problem.num_observations)
        res = abi.GncResult()
        st = self.lib.vba_optimize_gnc(self.ctx, C.byref(view), C.byref(options), C.byref(gnc),
                                       poses.ctypes.data_as(abi.P64), lms.ctypes.data_as(abi.P64),
                                       w.ctypes.data_as(abi.P64), C.byref(res))
        return st, res, poses, lms, w

    def dense_solve(self, h, b, noalloc: bool = False):
        """solve_normal_equations on the device (vba_dense_solve, or
        vba_dense_solve_noalloc with `noalloc`): returns (status, x); x is
        None unless the status is OK."""
        h = np.asfortranarray(h, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        n = len(b)
        assert h.shape == (n, n)
        x = np.zeros(n)
        fn = self.lib.vba_dense_solve_noalloc if noalloc else self.lib.vba_dense_solve
        st = fn(self.ctx, n, h.ravel(order="F").ctypes.data_as(abi.P64), b.ctypes.data_as(abi.P64),
                x.ctypes.data_as(abi.P64))
        return st, (x if st == abi.OK else None)

    def reduced_solve(self, diag, edges, offdiag, rhs, tol: float, max_iterations: int = 500, clusters=None,
                      coarse=None, structure_id: int = 0):
        """vba_reduced_solve, `ReducedSolveAccelerator::solve`
        (pipelines/slam/src/ba_accel.rs:138-145). Blocks are numpy 6x6
        (row-major here, passed column-major); `offdiag[e]` is edge e's
        (row b, column a) block, stored as FP32; `clusters` a list of
        (first node, inverse) over runs of equal length, `coarse` a tuple
        (group, nodes covered, inverse). Returns (step (n, 6), iterations),
        or None where the trait returns None (vba_last_error says why)."""
        diag = np.asarray(diag, dtype=np.float64)
        n = len(diag)
        edges = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1, 2))
        col = lambda m: np.ascontiguousarray(np.swapaxes(np.asarray(m), -1, -2))
        d = col(diag).astype(np.float64)
        off = None if offdiag is None else col(np.asarray(offdiag, dtype=np.float32).reshape(-1, 6, 6)).astype(
            np.float32, copy=False)
        b = np.ascontiguousarray(np.asarray(rhs, dtype=np.float64).reshape(n, 6))
        sysv = abi.ReducedSystem()
        sysv.structure_id = structure_id
        sysv.num_nodes = n
        sysv.num_edges = len(edges)
        sysv.diag = d.ctypes.data_as(abi.P64)
        sysv.edges = edges.ctypes.data_as(C.POINTER(C.c_uint32))
        sysv.offdiag = None if off is None else off.ctypes.data_as(C.POINTER(C.c_float))
        sysv.rhs = b.ctypes.data_as(abi.P64)
        sysv.tol = tol
        sysv.max_iterations = max_iterations
        keep = [d, edges, off, b]
        if clusters is not None:
            # runs of k nodes from node 0 (the last may be shorter), as the device solve takes them
            k = clusters[0][1].shape[0] // 6
            if len(clusters) != -(-n // k):
                return None
            for c, (start, inv) in enumerate(clusters):
                m = min(k, n - c * k)
                if start != c * k or np.shape(inv) != (6 * m, 6 * m):
                    return None
            cl = np.ascontiguousarray(np.concatenate([np.asarray(inv, dtype=np.float64).T.reshape(-1)
                                                      for _, inv in clusters]))
            sysv.cluster_size = k
            sysv.cluster_inverses = cl.ctypes.data_as(abi.P64)
            keep.append(cl)
        if coarse is not None:
            group, nodes, inv = coarse
            m = -(-nodes // group)
            if np.shape(inv) != (6 * m, 6 * m):
                return None
            ci = np.ascontiguousarray(np.asarray(inv, dtype=np.float64).T)
            sysv.coarse_group = group
            sysv.coarse_nodes = nodes
            sysv.coarse_inverse = ci.ctypes.data_as(abi.P64)
            keep.append(ci)
        step = np.zeros((n, 6))
        its = C.c_uint32(0)
        st = self.lib.vba_reduced_solve(self.ctx, C.byref(sysv), step.ctypes.data_as(abi.P64), C.byref(its))
        if st == abi.DECLINED:
            return None
        if st != abi.OK:
            raise RuntimeError(f"vba_reduced_solve: {st} {self.last_error()}")
        return step, int(its.value)

    def reduced_eliminate(self, elim: "PointEliminationInput"):
        """vba_reduced_eliminate, `ReducedSolveAccelerator::eliminate`
        (pipelines/slam/src/ba_accel.rs:150-156): returns (run_blocks
        (len(run_edges), 6, 6), coarse_sums (m, m, 6, 6) or None), row-major
        6x6 here, or None where the trait returns None. The blocks stay on
        the device for the next reduced_solve(offdiag=None) of the
        structure."""
        view, keep = elim.view()
        runs = np.zeros((max(1, len(elim.run_edges)), 36))
        m = elim.coarse_aggregates()
        coarse = np.zeros((max(1, m * m), 36))
        st = self.lib.vba_reduced_eliminate(self.ctx, C.byref(view), runs.ctypes.data_as(abi.P64),
                                            coarse.ctypes.data_as(abi.P64))
        del keep
        if st == abi.DECLINED:
            return None
        if st != abi.OK:
            raise RuntimeError(f"vba_reduced_eliminate: {st} {self.last_error()}")
        rb = np.swapaxes(runs[:len(elim.run_edges)].reshape(-1, 6, 6), -1, -2).copy()
        cs = np.swapaxes(coarse[:m * m].reshape(m, m, 6, 6), -1, -2).copy() if m else None
        return rb, cs

    def optimize_batch(self, problems: List[BundleProblem], options: abi.Options,
                       cap: int = 64) -> "BatchOutput":
        """vba_optimize_batch: independent windows solved together with the
        dense reduced solve, LM control per window on the device."""
        n = len(problems)
        views = (abi.Problem * max(n, 1))()
        for i, p in enumerate(problems):
            views[i] = p.view()
        poses = [np.empty_like(p.poses) for p in problems]
        lms = [np.empty_like(p.landmarks) for p in problems]
        pose_ptrs = (abi.P64 * max(n, 1))(*[a.ctypes.data_as(abi.P64) for a in poses])
        lm_ptrs = (abi.P64 * max(n, 1))(*[a.ctypes.data_as(abi.P64) for a in lms])
        stats = (abi.IterationStats * max(n * cap, 1))()
        results = (abi.Result * max(n, 1))()
        statuses = (C.c_int32 * max(n, 1))()
        st = self.lib.vba_optimize_batch(self.ctx, n, views, C.byref(options), pose_ptrs, lm_ptrs, stats, cap,
                                         results, statuses)
        outs = []
        for i in range(n):
            ok = st == abi.OK and statuses[i] == abi.OK
            k = min(results[i].num_iterations, cap)
            its = [{f[0].rstrip("_"): getattr(stats[i * cap + j], f[0]) for f in abi.IterationStats._fields_}
                   for j in range(k)] if ok else []
            outs.append(SolveOutput(statuses[i] if st == abi.OK else st, abi.Result.from_buffer_copy(results[i]),
                                    its, poses[i] if ok else None, lms[i] if ok else None))
        return BatchOutput(st, outs, self.last_error() if st != abi.OK else "")

    def prepare_batch(self, problems: List[BundleProblem], options: abi.Options) -> "PreparedBatch":
        return PreparedBatch(self, problems, options)


@dataclass
class BatchOutput:
    status: int
    windows: List[SolveOutput]
    message: str = ""


class PreparedProblem:
    """A problem laid out and resident on the device (vba_prepare). solve()
    starts from the prepared initial state each time; state() reads back the
    final state of the last solve."""

    def __init__(self, device: "Device", problem: BundleProblem, options: abi.Options):
        self.device = device
        self.problem = problem
        self.handle = C.c_void_p()
        res = abi.Result()
        view = problem.view()
        st = device.lib.vba_prepare(device.ctx, C.byref(view), C.byref(options), C.byref(self.handle), C.byref(res))
        if st != abi.OK:
            self.handle = None
            raise PrepareError(st, device.last_error(), res.error_id)

    def solve(self, options: abi.Options, cap: int = 1024) -> SolveOutput:
        stats = (abi.IterationStats * cap)()
        res = abi.Result()
        st = self.device.lib.vba_solve_prepared(self.device.ctx, self.handle, C.byref(options), stats, cap,
                                                C.byref(res))
        n = min(res.num_iterations, cap)
        out = SolveOutput(st, res, stats_list(stats, n) if st == abi.OK else [], None, None)
        if st != abi.OK:
            out.message = self.device.last_error()
        return out

    def autotune(self, options: abi.Options, lm_iterations: int = 3) -> dict:
        """Pick the fastest PCG launch schedule for this problem
        (vba_autotune); it is left set on the device context."""
        t = abi.Tuning()
        best = C.c_double()
        default = C.c_double()
        st = self.device.lib.vba_autotune(self.device.ctx, self.handle, C.byref(options), int(lm_iterations),
                                          C.byref(t), C.byref(best), C.byref(default))
        if st != abi.OK:
            raise RuntimeError(self.device.last_error())
        return {"tuning": t.as_dict(), "seconds": best.value, "default_seconds": default.value}

    def state(self):
        poses = np.empty_like(self.problem.poses)
        lms = np.empty_like(self.problem.landmarks)
        st = self.device.lib.vba_prepared_state(self.device.ctx, self.handle, poses.ctypes.data_as(abi.P64),
                                                lms.ctypes.data_as(abi.P64))
        if st != abi.OK:
            raise RuntimeError(self.device.last_error())
        return poses, lms

    def free(self):
        if self.handle:
            self.device.lib.vba_prepared_free(self.device.ctx, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PreparedBatch:
    """Windows resident on the device as one batch (vba_prepare_batch)."""

    def __init__(self, device: "Device", problems: List[BundleProblem], options: abi.Options):
        self.device = device
        self.problems = list(problems)
        n = len(self.problems)
        views = (abi.Problem * max(n, 1))()
        for i, p in enumerate(self.problems):
            views[i] = p.view()
        self.handle = C.c_void_p()
        results = (abi.Result * max(n, 1))()
        statuses = (C.c_int32 * max(n, 1))()
        st = device.lib.vba_prepare_batch(device.ctx, n, views, C.byref(options), C.byref(self.handle), results,
                                          statuses)
        if st != abi.OK:
            self.handle = None
            raise PrepareError(st, device.last_error())
        self.statuses = [int(statuses[i]) for i in range(n)]
        self.error_ids = [int(results[i].error_id) for i in range(n)]

    def solve(self, options: abi.Options, cap: int = 64) -> BatchOutput:
        n = len(self.problems)
        stats = (abi.IterationStats * max(n * cap, 1))()
        results = (abi.Result * max(n, 1))()
        statuses = (C.c_int32 * max(n, 1))(*self.statuses)
        st = self.device.lib.vba_solve_prepared_batch(self.device.ctx, self.handle, C.byref(options), stats, cap,
                                                      results, statuses)
        outs = []
        for i in range(n):
            ok = st == abi.OK and statuses[i] == abi.OK
            k = min(results[i].num_iterations, cap)
            its = [{f[0].rstrip("_"): getattr(stats[i * cap + j], f[0]) for f in abi.IterationStats._fields_}
                   for j in range(k)] if ok else []
            outs.append(SolveOutput(statuses[i] if st == abi.OK else st, abi.Result.from_buffer_copy(results[i]),
                                    its, None, None))
        return BatchOutput(st, outs, self.device.last_error() if st != abi.OK else "")

    def state(self, i: int):
        p = self.problems[i]
        poses = np.empty_like(p.poses)
        lms = np.empty_like(p.landmarks)
        st = self.device.lib.vba_prepared_batch_state(self.device.ctx, self.handle, i, poses.ctypes.data_as(abi.P64),
                                                      lms.ctypes.data_as(abi.P64))
        if st != abi.OK:
            raise RuntimeError(self.device.last_error())
        return poses, lms

    def free(self):
        if self.handle:
            self.device.lib.vba_prepared_batch_free(self.device.ctx, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class ServedResult:
    status: int
    result: abi.Result
    poses: Optional[np.ndarray]
    landmarks: Optional[np.ndarray]
    latency: float


class Server:
    """The request loop (vba_server_*): submit() returns a ticket at once,
    wait() returns the finished solve with its latency."""

    def __init__(self, **config):
        self.lib = abi.library()
        cfg = abi.ServerConfig()
        self.lib.vba_server_default_config(C.byref(cfg))
        for k, v in config.items():
            if not hasattr(cfg, k):
                raise AttributeError(k)
            setattr(cfg, k, v)
        self.handle = self.lib.vba_server_create(C.byref(cfg))
        if not self.handle:
            raise RuntimeError("vba_server_create failed: " + self.lib.vba_last_error(None).decode())
        self.shapes = {}
        self.borrowed = {}

    def submit(self, problem: BundleProblem, options: abi.Options, borrow: bool = False) -> int:
        """Queue a solve. With borrow=True the server reads the problem's
        arrays in place: keep `problem` alive and unchanged until wait()
        has returned for the ticket."""
        view = problem.view()
        fn = self.lib.vba_server_submit_borrowed if borrow else self.lib.vba_server_submit
        t = fn(self.handle, C.byref(view), C.byref(options))
        if t == 0:
            raise RuntimeError(self.lib.vba_last_error(None).decode())
        self.shapes[t] = (problem.poses.shape, problem.landmarks.shape)
        if borrow:
            self.borrowed[t] = (problem, view)
        return t

    def wait(self, ticket: int, timeout: float = 600.0) -> ServedResult:
        ps, ls = self.shapes[ticket]
        poses = np.empty(ps)
        lms = np.empty(ls)
        res = abi.Result()
        lat = C.c_double()
        st = self.lib.vba_server_wait(self.handle, ticket, timeout, poses.ctypes.data_as(abi.P64),
                                      lms.ctypes.data_as(abi.P64), C.byref(res), C.byref(lat))
        if st == abi.PENDING:
            return ServedResult(st, res, None, None, float("nan"))
        del self.shapes[ticket]
        self.borrowed.pop(ticket, None)
        ok = st == abi.OK
        return ServedResult(st, res, poses if ok else None, lms if ok else None, lat.value)

    def request_bytes(self, problem: BundleProblem, options: abi.Options) -> int:
        """Device bytes the server counts against its HBM budget for this request."""
        view = problem.view()
        return int(self.lib.vba_server_request_bytes(self.handle, C.byref(view), C.byref(options)))

    def counters(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        self.lib.vba_server_counters(self.handle, *[C.byref(x) for x in v])
        return dict(zip(("completed", "batches", "batched_windows", "queued"), (int(x.value) for x in v)))

    def close(self):
        if self.handle:
            self.lib.vba_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PrepareError(RuntimeError):
    def __init__(self, status, message, error_id=0):
        super().__init__(f"vba_prepare status {status}: {message}")
        self.status = status
        self.error_id = error_id


def check(problem: BundleProblem, options: abi.Options) -> int:
    lib = abi.library()
    view = problem.view()
    return lib.vba_check(C.byref(view), C.byref(options))


def camera_centres(poses: np.ndarray) -> np.ndarray:
    """Pose::camera_center_world (crates/core/src/geometry/pose.rs:31-33):
    C = -R^T t for world->camera (q, t)."""
    poses = np.asarray(poses, np.float64).reshape(-1, 7)
    w, x, y, z = poses[:, 0], poses[:, 1], poses[:, 2], poses[:, 3]
    R = np.empty((len(poses), 3, 3))
    R[:, 0, 0] = w * w + x * x - y * y - z * z
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (w * y + x * z)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = w * w - x * x + y * y - z * z
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (w * x + y * z)
    R[:, 2, 2] = w * w - x * x - y * y + z * z
    return -np.einsum("nji,nj->ni", R, poses[:, 4:7])
