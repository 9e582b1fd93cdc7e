"""Bundle-adjustment problems on disk (include/vba_io.h).

`Model` wraps the library's host-side model: the reference's plain-text
problem dump (`read_ba_problem` / `write_ba_problem`,
pipelines/slam/src/ba_problem_io.rs:20-149) and COLMAP sparse models, text
or binary (`read_colmap_*_model` / `write_colmap_*_model`,
crates/io/src/colmap/mod.rs:148-236). `Model.problem()` is the
`BundleProblem` the solver takes; `Model.set_state` takes its result back so
that the model can be written out again.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Tuple

import numpy as np

from . import abi
from .api import BundleProblem, _np_from


class IoError(RuntimeError):
    """A failed call of vba_io.h: `code` is its status (abi.ERR_IO,
    abi.ERR_FORMAT, abi.ERR_EXPORT, ...), the message the library's."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (status {code})")
        self.code = code
        self.message = message


def _check(status: int) -> None:
    if status != abi.OK:
        raise IoError(status, abi.library().vba_io_last_error().decode(errors="replace"))


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


class Model:
    """A problem held by the library, read from or written to disk."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    # ---- constructors -----------------------------------------------------
    @classmethod
    def read_ba_problem(cls, path) -> "Model":
        h = C.c_void_p()
        _check(abi.library().vba_model_read_ba_problem(_path(path), C.byref(h)))
        return cls(h)

    @classmethod
    def read_colmap(cls, directory, binary: bool = False) -> "Model":
        h = C.c_void_p()
        _check(abi.library().vba_model_read_colmap(_path(directory), int(bool(binary)), C.byref(h)))
        return cls(h)

    @classmethod
    def from_problem(cls, problem: BundleProblem, weights: Optional[np.ndarray] = None) -> "Model":
        view = problem.view()
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, np.float64)
            if len(w) != len(problem.obs_pose_ids):
                raise ValueError("one weight per observation")
        h = C.c_void_p()
        _check(abi.library().vba_model_from_problem(
            C.byref(view), w.ctypes.data_as(abi.P64) if w is not None else None, C.byref(h)))
        return cls(h)

    # ---- writers ----------------------------------------------------------
    def write_ba_problem(self, path) -> None:
        _check(abi.library().vba_model_write_ba_problem(self._handle(), _path(path)))

    def write_colmap(self, directory, binary: bool = False) -> None:
        _check(abi.library().vba_model_write_colmap(self._handle(), _path(directory), int(bool(binary))))

    # ---- the problem --------------------------------------------------------
    def problem(self) -> Tuple[BundleProblem, Optional[np.ndarray]]:
        """The model as a BundleProblem (copied out of the library's view)
        and its observation weights, or None when unweighted."""
        v = abi.Problem()
        w = abi.P64()
        _check(abi.library().vba_model_problem(self._handle(), C.byref(v), C.byref(w)))
        P, L = int(v.num_poses), int(v.num_landmarks)
        N, R, S = int(v.num_observations), int(v.num_rig_observations), int(v.num_sensors)
        sensors = [abi.Sensor.from_buffer_copy(v.sensors[i]) for i in range(S)]
        prob = BundleProblem(
            camera=abi.Camera.from_buffer_copy(v.camera),
            pose_ids=_np_from(v.pose_ids, P, np.uint64, (P,)),
            poses=_np_from(v.poses, P, np.float64, (P, 7)),
            pose_fixed=_np_from(v.pose_fixed, P, np.uint8, (P,)),
            landmark_ids=_np_from(v.landmark_ids, L, np.uint64, (L,)),
            landmarks=_np_from(v.landmarks, L, np.float64, (L, 3)),
            landmark_fixed=np.zeros(L, np.uint8),
            obs_pose_ids=_np_from(v.obs_pose_ids, N, np.uint64, (N,)),
            obs_landmark_ids=_np_from(v.obs_landmark_ids, N, np.uint64, (N,)),
            obs_xy=_np_from(v.obs_xy, N, np.float64, (N, 2)),
            rig_pose_ids=_np_from(v.rig_pose_ids, R, np.uint64, (R,)),
            rig_landmark_ids=_np_from(v.rig_landmark_ids, R, np.uint64, (R,)),
            rig_xy=_np_from(v.rig_xy, R, np.float64, (R, 2)),
            rig_sensor=_np_from(v.rig_sensor, R, np.uint32, (R,)),
            sensors=sensors,
            declined_features=int(v.declined_features),
        )
        weights = _np_from(w, N, np.float64, (N,)) if w else None
        return prob, weights

    def set_state(self, poses: Optional[np.ndarray] = None, landmarks: Optional[np.ndarray] = None) -> None:
        prob, _ = self.problem()
        p = l = None
        if poses is not None:
            p = np.ascontiguousarray(poses, np.float64).reshape(-1)
            if p.size != prob.poses.size:
                raise ValueError("7 values per pose of the model")
        if landmarks is not None:
            l = np.ascontiguousarray(landmarks, np.float64).reshape(-1)
            if l.size != prob.landmarks.size:
                raise ValueError("3 values per landmark of the model")
        _check(abi.library().vba_model_set_state(
            self._handle(), p.ctypes.data_as(abi.P64) if p is not None else None,
            l.ctypes.data_as(abi.P64) if l is not None else None))

    def set_fixed(self, fixed: np.ndarray) -> None:
        prob, _ = self.problem()
        f = np.ascontiguousarray(fixed, np.uint8)
        if f.size != len(prob.pose_ids):
            raise ValueError("one flag per pose of the model")
        _check(abi.library().vba_model_set_fixed(self._handle(), f.ctypes.data_as(C.POINTER(C.c_uint8))))

    def keep_local_window(self, n: int) -> None:
        _check(abi.library().vba_model_keep_local_window(self._handle(), int(n)))

    # ---- lifetime -----------------------------------------------------------
    def _handle(self):
        if not self._h:
            raise ValueError("model is closed")
        return self._h

    def close(self) -> None:
        if self._h:
            abi.library().vba_model_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
