"""The seeded workloads the solver is measured and checked on, with their LM
options, and a digest of a problem's bytes.

Every workload is built by the library's seeded generators
(include/vba_synth.h), so a (workload, seed) pair gives the same bytes on
every machine; `problem_digest` hashes those bytes, which lets a stored
trace name the exact problem it was computed on.

- circle: 50 cameras on a radius-10 m circle, shared SIMPLE_RADIAL
  [500, 320, 240, 0.01] at 640x480, 5,000 landmarks in a 4 m cube, every
  in-frame observation kept; squared loss, dense Schur + Cholesky,
  20 LM iterations.
- loop: 1,200 pinhole cameras (f=800, 1024x768) on a closed 200 m loop,
  400k landmarks in a 20 m corridor, ~3.5M observations, Huber 1 px,
  matrix-free block-Jacobi PCG (rtol 1e-4, max 100), 30 LM iterations.
- windows: local stereo windows, K uniform in [5, 40], L in [300, 8000],
  Cauchy 1 px, dense, 10 LM iterations.
- city: 5,500 cameras with per-camera OPENCV intrinsics on a city-block
  orbit, 2.5M landmarks, ~20M observations, 10 % outliers, Huber 2 px,
  matrix-free PCG (rtol 1e-4, max 100), 25 LM iterations; `city_problem`
  with fewer cameras keeps the landmark density of the full scene.
- pose-only: one free camera over 200-2,000 fixed points, 1 px noise,
  5 % outliers, Cauchy 1 px, dense, 10 LM iterations.
- submaps: 100-300 loop cameras, 20k-120k landmarks.
"""
from __future__ import annotations

import hashlib

import numpy as np

from . import abi
from .api import BundleProblem, default_options, synthesize


def circle_options(precision=abi.PRECISION_FP64):
    return default_options(linear_solver=abi.SOLVER_DENSE, max_iterations=20, robust_kernel=abi.KERNEL_NONE,
                           precision=precision)


def circle_problem(seed: int | None = None):
    kw = {} if seed is None else {"seed": seed}
    return synthesize(abi.SYNTH_CIRCLE, **kw)


def loop_options(precision=abi.PRECISION_FP64, policy=abi.PCG_MAXIT_ACCEPT_ITERATE, solver=abi.SOLVER_MATRIX_FREE,
                 max_iterations: int = 30):
    """PCG rtol 1e-4 / max 100 are the reference GPU backend's GpuBaSettings
    defaults (crates/ba-gpu/src/solver.rs:36-48); ACCEPT_ITERATE takes the
    last iterate of a solve that reaches the cap, as that backend does
    (solver.rs:497-508), REJECT fails the step as the CPU matrix-free entry
    does (implicit_schur.rs:715-736)."""
    return default_options(linear_solver=solver, max_iterations=max_iterations, max_pcg_iterations=100,
                           pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                           precision=precision, pcg_maxit_policy=policy)


def loop_problem(num_landmarks: int | None = None):
    kw = {} if num_landmarks is None else {"num_landmarks": num_landmarks}
    return synthesize(abi.SYNTH_LOOP, **kw)


def dense_loop_options(linear_solver=abi.SOLVER_DENSE, max_iterations: int = 8):
    """A large dense reduced solve: the loop scene's Huber 1 px with the
    dense Schur complement and Cholesky (LinearSolver::Dense)."""
    return default_options(linear_solver=linear_solver, max_iterations=max_iterations,
                           robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0)


def dense_loop_problem(num_cameras: int = 1500, num_landmarks: int = 60000):
    """Loop cameras at the full scene's spacing (a 250 m loop for 1,500
    cameras), for the dense reduced solve of 6 x (cameras - 1) unknowns."""
    return synthesize(abi.SYNTH_LOOP, num_cameras=num_cameras, num_landmarks=num_landmarks,
                      scene_scale=200.0 * num_cameras / 1200.0)


def window_options(precision=abi.PRECISION_FP64):
    return default_options(linear_solver=abi.SOLVER_DENSE, max_iterations=10, robust_kernel=abi.KERNEL_CAUCHY,
                           robust_scale=1.0, precision=precision)


def window_specs(count: int = 256, seed: int = 7):
    """(keyframes, landmarks, generator seed) of each window."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        k = int(rng.integers(5, 41))
        l = int(rng.integers(300, 8001))
        out.append((k, l, 1000 + i))
    return out


def window_problem(spec):
    k, l, seed = spec
    return synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=k, num_landmarks=l, seed=seed)


def window_problems(count: int = 256, seed: int = 7):
    return [window_problem(s).problem for s in window_specs(count, seed)]


CITY_CAMERAS = 5500
CITY_LANDMARKS = 2_500_000


def city_options(precision=abi.PRECISION_FP64, policy=abi.PCG_MAXIT_ACCEPT_ITERATE, max_iterations: int = 25):
    return default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_iterations=max_iterations,
                           max_pcg_iterations=100, pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER,
                           robust_scale=2.0, precision=precision, pcg_maxit_policy=policy)


def city_problem(num_cameras: int = CITY_CAMERAS):
    """The city workload; fewer cameras keep the full scene's landmarks per
    camera (a sub-problem of the same generator)."""
    lms = CITY_LANDMARKS * num_cameras // CITY_CAMERAS
    return synthesize(abi.SYNTH_CITY, num_cameras=num_cameras, num_landmarks=lms)


def pose_only_options(precision=abi.PRECISION_FP64):
    """Kernel and iteration count of pose-only requests are not fixed by the
    workload; Cauchy 1 px and 10 iterations are this build's choice."""
    return default_options(linear_solver=abi.SOLVER_DENSE, max_iterations=10, robust_kernel=abi.KERNEL_CAUCHY,
                           robust_scale=1.0, precision=precision)


def pose_only_problems(count: int, seed: int = 11):
    rng = np.random.default_rng(seed)
    return [synthesize(abi.SYNTH_POSE_ONLY, num_landmarks=int(rng.integers(200, 2001)), seed=5000 + i).problem
            for i in range(count)]


def submap_problems(count: int, seed: int = 11):
    rng = np.random.default_rng(seed + 1)
    out = []
    for i in range(count):
        cams = int(rng.integers(100, 301))
        out.append(synthesize(abi.SYNTH_LOOP, num_cameras=cams, num_landmarks=int(rng.integers(20000, 120001)),
                              seed=7000 + i, scene_scale=200.0 * cams / 1200.0).problem)
    return out


def problem_digest(p: BundleProblem) -> str:
    """SHA-256 over every array and struct of the problem, in a fixed order."""
    h = hashlib.sha256()
    h.update(bytes(p.camera))
    for s in list(p.sensors) + list(p.pairs):
        h.update(bytes(s))
    h.update(np.float64(p.stereo_baseline).tobytes())
    for a in (p.pose_ids, p.poses, p.pose_fixed, p.landmark_ids, p.landmarks, p.landmark_fixed, p.obs_pose_ids,
              p.obs_landmark_ids, p.obs_xy, p.rig_pose_ids, p.rig_landmark_ids, p.rig_xy, p.rig_sensor,
              p.stereo_pose_ids, p.stereo_landmark_ids, p.stereo_obs, p.general_pose_ids, p.general_landmark_ids,
              p.general_obs, p.general_pair):
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(np.asarray(a.shape, np.int64).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()
