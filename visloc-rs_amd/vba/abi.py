"""ctypes mirror of include/vba.h and include/vba_synth.h.

The structures below must match the C layouts field for field; the library
is libvba.so built in-tree by `make -C visloc-rs_amd`. Loading fails loudly
when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

ABI_VERSION = 2

OK = 0
DECLINED = 1
FLAG_NO_GRAPH = 1  # vba_options.flags: launch a prepared batch kernel by kernel
PENDING = 2
ERR_NO_POSES = -1
ERR_NO_LANDMARKS = -2
ERR_NO_OBSERVATIONS = -3
ERR_ALL_POSES_FIXED = -4
ERR_MISSING_POSE = -5
ERR_MISSING_LANDMARK = -6
ERR_UNSUPPORTED_CAMERA_MODEL = -7
ERR_MISSING_STEREO_BASELINE = -8
ERR_OBSERVATION_WEIGHT_COUNT = -9
ERR_INVALID_OBSERVATION_WEIGHT = -10
ERR_WEIGHTS_WITH_INTRINSICS = -12
ERR_SINGULAR_SYSTEM = -13
ERR_INELIGIBLE = -20
ERR_INVALID_CONFIGURATION = -21
ERR_LINEAR_SOLVE = -22
ERR_INVALID_ARGUMENT = -100
ERR_DEVICE = -101
ERR_OUT_OF_MEMORY = -102
ERR_IO = -110           # vba_io.h: a file could not be read or written
ERR_FORMAT = -111       # vba_io.h: malformed content
ERR_EXPORT = -112       # vba_io.h: the model cannot be written in the asked format

CAMERA_PINHOLE = 0
CAMERA_SIMPLE_PINHOLE = 1
CAMERA_SIMPLE_RADIAL = 2
CAMERA_RADIAL = 3
CAMERA_OPENCV = 4

KERNEL_NONE = 0
KERNEL_HUBER = 1
KERNEL_CAUCHY = 2

SOLVER_DENSE = 0
SOLVER_SPARSE = 1  # LinearSolver::Sparse: Cholesky only, no LU fallback
SOLVER_MATRIX_FREE = 2
SOLVER_PCG_EXPLICIT = 3

PRECISION_FP64 = 0
PRECISION_MIXED = 1

PCG_MAXIT_REJECT = 0
PCG_MAXIT_ACCEPT_ITERATE = 1

FEATURE_STEREO_OBSERVATIONS = 1 << 0
FEATURE_GENERAL_STEREO = 1 << 1
FEATURE_PRIORS = 1 << 2
FEATURE_INERTIAL = 1 << 3
FEATURE_FIXED_ROTATIONS = 1 << 4
FEATURE_OBSERVATION_WEIGHTS = 1 << 5

# vba_options.refine_intrinsics bits (BaConfig::refine_intrinsics,
# refine_distortion, refine_tangential_distortion, shared_focal)
REFINE_INTRINSICS = 1
REFINE_DISTORTION = 2
REFINE_TANGENTIAL = 4
REFINE_SHARED_FOCAL = 8

SYNTH_CIRCLE = 1
SYNTH_LOOP = 2
SYNTH_STEREO_WINDOW = 3
SYNTH_POSE_ONLY = 4
SYNTH_CITY = 5

P64 = C.POINTER(C.c_double)
PU64 = C.POINTER(C.c_uint64)
PU32 = C.POINTER(C.c_uint32)
PU8 = C.POINTER(C.c_uint8)


class Camera(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("num_params", C.c_uint32),
        ("params", C.c_double * 12),
    ]


class Sensor(C.Structure):
    _fields_ = [("camera", Camera), ("sensor_from_rig", C.c_double * 7)]


class StereoTerms(C.Structure):
    """vba_stereo_terms: rectified (u_l, v_l, u_r) and general (left x, y,
    right x, y) stereo observations."""
    _fields_ = [
        ("baseline", C.c_double),
        ("num_stereo", C.c_uint64),
        ("stereo_pose_ids", PU64),
        ("stereo_landmark_ids", PU64),
        ("stereo_obs", P64),
        ("num_general", C.c_uint64),
        ("general_pose_ids", PU64),
        ("general_landmark_ids", PU64),
        ("general_obs", P64),
        ("general_pair", PU32),
        ("num_pairs", C.c_uint64),
        ("pairs", C.POINTER(Sensor)),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("camera", Camera),
        ("num_poses", C.c_uint64),
        ("pose_ids", PU64),
        ("poses", P64),
        ("pose_fixed", PU8),
        ("num_landmarks", C.c_uint64),
        ("landmark_ids", PU64),
        ("landmarks", P64),
        ("landmark_fixed", PU8),
        ("num_observations", C.c_uint64),
        ("obs_pose_ids", PU64),
        ("obs_landmark_ids", PU64),
        ("obs_xy", P64),
        ("num_rig_observations", C.c_uint64),
        ("rig_pose_ids", PU64),
        ("rig_landmark_ids", PU64),
        ("rig_xy", P64),
        ("rig_sensor", PU32),
        ("num_sensors", C.c_uint64),
        ("sensors", C.POINTER(Sensor)),
        ("declined_features", C.c_uint32),
        ("reserved", C.c_uint32),
        ("stereo", C.POINTER(StereoTerms)),
    ]


class Options(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint32),
        ("has_initial_lambda", C.c_uint32),
        ("initial_lambda", C.c_double),
        ("lambda_increase_factor", C.c_double),
        ("lambda_decrease_factor", C.c_double),
        ("max_lambda", C.c_double),
        ("min_lambda", C.c_double),
        ("step_tolerance", C.c_double),
        ("cost_tolerance", C.c_double),
        ("has_relative_cost_tolerance", C.c_uint32),
        ("robust_kernel", C.c_uint32),
        ("relative_cost_tolerance", C.c_double),
        ("robust_scale", C.c_double),
        ("linear_solver", C.c_uint32),
        ("precision", C.c_uint32),
        ("max_pcg_iterations", C.c_uint32),
        ("pcg_maxit_policy", C.c_uint32),
        ("pcg_relative_tolerance", C.c_double),
        ("pcg_absolute_tolerance", C.c_double),
        ("refine_intrinsics", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class IterationStats(C.Structure):
    _fields_ = [
        ("iteration", C.c_uint32),
        ("step_accepted", C.c_uint32),
        ("cost_before", C.c_double),
        ("cost_after", C.c_double),
        ("max_pose_step", C.c_double),
        ("max_landmark_step", C.c_double),
        ("lambda_", C.c_double),
        ("pcg_iterations", C.c_int32),
        ("nonprojectable_after", C.c_uint32),
        ("pcg_residual_norm", C.c_double),
        ("pcg_target", C.c_double),
    ]


class Result(C.Structure):
    _fields_ = [
        ("initial_cost", C.c_double),
        ("final_cost", C.c_double),
        ("num_iterations", C.c_uint32),
        ("converged", C.c_uint32),
        ("error_id", C.c_uint64),
        ("num_observations_used", C.c_uint64),
        ("total_pcg_iterations", C.c_uint32),
        ("linearizations", C.c_uint32),
        ("device_seconds", C.c_double),
    ]


class SynthParams(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("num_cameras", C.c_uint32),
        ("seed", C.c_uint64),
        ("num_landmarks", C.c_uint64),
        ("pixel_noise", C.c_double),
        ("rotation_noise", C.c_double),
        ("translation_noise", C.c_double),
        ("landmark_noise", C.c_double),
        ("outlier_fraction", C.c_double),
        ("mean_track_length", C.c_double),
        ("min_track_length", C.c_uint32),
        ("max_track_length", C.c_uint32),
        ("scene_scale", C.c_double),
        ("focal", C.c_double),
        ("distortion", C.c_double),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("fixed_cameras", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


# Every symbol include/vba.h and include/vba_synth.h declare, with its
# signature (restype, argtypes).
class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("max_window_batch", C.c_uint32),
        ("window_streams", C.c_uint32),
        ("submap_streams", C.c_uint32),
        ("hbm_budget_bytes", C.c_uint64),
        ("batch_wait_seconds", C.c_double),
        ("build_threads", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


GNC_GEMAN_MCCLURE = 0
GNC_TRUNCATED_LEAST_SQUARES = 1


class GncOptions(C.Structure):
    """vba_gnc_options (GncConfig)."""
    _fields_ = [("kernel", C.c_uint32), ("max_outer", C.c_uint32), ("inner_iterations", C.c_uint32),
                ("has_auto_scale", C.c_uint32), ("c", C.c_double), ("anneal_factor", C.c_double),
                ("auto_scale", C.c_double), ("auto_scale_readapt", C.c_uint32), ("reserved", C.c_uint32)]


class GncResult(C.Structure):
    """vba_gnc_result (BaGncResult)."""
    _fields_ = [("initial_cost", C.c_double), ("final_cost", C.c_double), ("inlier_cost", C.c_double),
                ("inlier_scale", C.c_double), ("observation_count", C.c_uint64), ("outer_iterations", C.c_uint32),
                ("converged", C.c_uint32)]


class Tuning(C.Structure):
    """vba_tuning: the PCG launch schedule (results do not depend on it)."""
    _fields_ = [("pcg_chunk", C.c_uint32), ("pcg_lookahead", C.c_uint32), ("pcg_predict", C.c_uint32),
                ("apply_waves_per_cu", C.c_uint32)]

    def as_dict(self):
        return {"pcg_chunk": self.pcg_chunk, "pcg_lookahead": self.pcg_lookahead, "pcg_predict": self.pcg_predict,
                "apply_waves_per_cu": self.apply_waves_per_cu}


class ReducedSystem(C.Structure):
    _fields_ = [
        ("structure_id", C.c_uint64),
        ("num_nodes", C.c_uint32),
        ("num_edges", C.c_uint32),
        ("diag", P64),
        ("edges", C.POINTER(C.c_uint32)),
        ("offdiag", C.POINTER(C.c_float)),
        ("rhs", P64),
        ("cluster_size", C.c_uint32),
        ("coarse_group", C.c_uint32),
        ("cluster_inverses", P64),
        ("coarse_nodes", C.c_uint32),
        ("max_iterations", C.c_uint32),
        ("coarse_inverse", P64),
        ("tol", C.c_double),
    ]


class PointElimination(C.Structure):
    _fields_ = [
        ("structure_id", C.c_uint64),
        ("linearization_id", C.c_uint64),
        ("num_nodes", C.c_uint32),
        ("num_edges", C.c_uint32),
        ("edges", PU32),
        ("row_off", PU64),
        ("slot_off", PU64),
        ("slots", PU32),
        ("num_points", C.c_uint32),
        ("num_extra", C.c_uint32),
        ("point_off", PU64),
        ("point_nodes", PU32),
        ("hcp", P64),
        ("point_inverse", P64),
        ("has_inverse", PU8),
        ("scale", P64),
        ("extra_edge", PU32),
        ("extra_blocks", P64),
        ("num_run_edges", C.c_uint32),
        ("coarse_group", C.c_uint32),
        ("run_edges", PU32),
        ("coarse_nodes", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


SIGNATURES = {
    "vba_abi_version": (C.c_uint32, []),
    "vba_create": (C.c_void_p, [C.c_int32]),
    "vba_destroy": (None, [C.c_void_p]),
    "vba_last_error": (C.c_char_p, [C.c_void_p]),
    "vba_default_options": (None, [C.POINTER(Options)]),
    "vba_check": (C.c_int32, [C.POINTER(Problem), C.POINTER(Options)]),
    "vba_optimize": (
        C.c_int32,
        [C.c_void_p, C.POINTER(Problem), C.POINTER(Options), P64, P64, C.POINTER(IterationStats),
         C.c_uint32, C.POINTER(Result)],
    ),
    "vba_prepare": (C.c_int32, [C.c_void_p, C.POINTER(Problem), C.POINTER(Options), C.POINTER(C.c_void_p),
                                C.POINTER(Result)]),
    "vba_solve_prepared": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Options), C.POINTER(IterationStats),
                                       C.c_uint32, C.POINTER(Result)]),
    "vba_prepared_state": (C.c_int32, [C.c_void_p, C.c_void_p, P64, P64]),
    "vba_prepared_free": (None, [C.c_void_p, C.c_void_p]),
    "vba_optimize_batch": (
        C.c_int32,
        [C.c_void_p, C.c_uint32, C.POINTER(Problem), C.POINTER(Options), C.POINTER(P64), C.POINTER(P64),
         C.POINTER(IterationStats), C.c_uint32, C.POINTER(Result), C.POINTER(C.c_int32)],
    ),
    "vba_prepare_batch": (
        C.c_int32,
        [C.c_void_p, C.c_uint32, C.POINTER(Problem), C.POINTER(Options), C.POINTER(C.c_void_p), C.POINTER(Result),
         C.POINTER(C.c_int32)],
    ),
    "vba_solve_prepared_batch": (
        C.c_int32,
        [C.c_void_p, C.c_void_p, C.POINTER(Options), C.POINTER(IterationStats), C.c_uint32, C.POINTER(Result),
         C.POINTER(C.c_int32)],
    ),
    "vba_prepared_batch_state": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, P64, P64]),
    "vba_prepared_batch_free": (None, [C.c_void_p, C.c_void_p]),
    "vba_server_default_config": (None, [C.POINTER(ServerConfig)]),
    "vba_server_create": (C.c_void_p, [C.POINTER(ServerConfig)]),
    "vba_server_submit": (C.c_uint64, [C.c_void_p, C.POINTER(Problem), C.POINTER(Options)]),
    "vba_server_submit_borrowed": (C.c_uint64, [C.c_void_p, C.POINTER(Problem), C.POINTER(Options)]),
    "vba_server_wait": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_double, P64, P64, C.POINTER(Result),
                                    C.POINTER(C.c_double)]),
    "vba_server_counters": (None, [C.c_void_p, PU64, PU64, PU64, PU64]),
    "vba_server_request_bytes": (C.c_uint64, [C.c_void_p, C.POINTER(Problem), C.POINTER(Options)]),
    "vba_server_destroy": (None, [C.c_void_p]),
    "vba_dense_solve": (C.c_int32, [C.c_void_p, C.c_uint32, P64, P64, P64]),
    "vba_dense_solve_noalloc": (C.c_int32, [C.c_void_p, C.c_uint32, P64, P64, P64]),
    "vba_reduced_solve": (C.c_int32, [C.c_void_p, C.POINTER(ReducedSystem), P64, C.POINTER(C.c_uint32)]),
    "vba_reduced_eliminate": (C.c_int32, [C.c_void_p, C.POINTER(PointElimination), P64, P64]),
    "vba_optimize_weighted": (
        C.c_int32,
        [C.c_void_p, C.POINTER(Problem), C.POINTER(Options), P64, C.c_uint64, P64, P64,
         C.POINTER(IterationStats), C.c_uint32, C.POINTER(Result)],
    ),
    "vba_optimize_intrinsics": (
        C.c_int32,
        [C.c_void_p, C.POINTER(Problem), C.POINTER(Options), P64, P64, C.POINTER(Camera),
         C.POINTER(IterationStats), C.c_uint32, C.POINTER(Result)],
    ),
    "vba_default_gnc_options": (None, [C.POINTER(GncOptions)]),
    "vba_optimize_gnc": (
        C.c_int32,
        [C.c_void_p, C.POINTER(Problem), C.POINTER(Options), C.POINTER(GncOptions), P64, P64, P64,
         C.POINTER(GncResult)],
    ),
    "vba_set_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    "vba_get_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    "vba_autotune": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Options), C.c_uint32, C.POINTER(Tuning),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    # vba_io.h: problems on disk
    "vba_io_last_error": (C.c_char_p, []),
    "vba_model_read_ba_problem": (C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "vba_model_write_ba_problem": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "vba_model_read_colmap": (C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "vba_model_write_colmap": (C.c_int32, [C.c_void_p, C.c_char_p, C.c_int32]),
    "vba_model_from_problem": (C.c_int32, [C.POINTER(Problem), P64, C.POINTER(C.c_void_p)]),
    "vba_model_problem": (C.c_int32, [C.c_void_p, C.POINTER(Problem), C.POINTER(P64)]),
    "vba_model_set_state": (C.c_int32, [C.c_void_p, P64, P64]),
    "vba_model_set_fixed": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint8)]),
    "vba_model_keep_local_window": (C.c_int32, [C.c_void_p, C.c_uint32]),
    "vba_model_free": (None, [C.c_void_p]),
    "vba_profile_enable": (None, [C.c_void_p, C.c_int32]),
    "vba_profile_reset": (None, [C.c_void_p]),
    "vba_profile_query": (C.c_int32, [C.c_void_p, C.c_char_p, PU64, P64, P64]),
    "vba_synth_default": (None, [C.c_uint32, C.POINTER(SynthParams)]),
    "vba_synth_generate": (C.c_void_p, [C.POINTER(SynthParams)]),
    "vba_synth_free": (None, [C.c_void_p]),
    "vba_synth_view": (None, [C.c_void_p, C.POINTER(Problem)]),
    "vba_synth_truth": (None, [C.c_void_p, C.POINTER(P64), C.POINTER(P64)]),
    "vba_synth_scene_extent": (C.c_double, [C.c_void_p]),
}

PACKAGE_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRARY_PATH = os.path.join(PACKAGE_ROOT, "libvba.so")

_lib = None


def library() -> C.CDLL:
    """Load libvba.so (built in-tree); raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: build it with `make -C visloc-rs_amd` "
                "(or __graft_entry__.build()); there is no fallback implementation"
            )
        lib = C.CDLL(LIBRARY_PATH)
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        if lib.vba_abi_version() != ABI_VERSION:
            raise RuntimeError("libvba.so ABI version mismatch")
        _lib = lib
    return _lib
