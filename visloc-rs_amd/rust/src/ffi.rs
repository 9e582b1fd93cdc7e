//! Declarations of `include/vba.h` and `include/vba_io.h`, field for field
//! and in declaration order (`#[repr(C)]`). tests/test_rust_ffi.py checks
//! every struct's field names, offsets and size, every constant and every
//! function's name and parameter count against the C headers.
#![allow(non_camel_case_types, dead_code)]

use std::os::raw::c_char;

// ---- include/vba.h --------------------------------------------------------

pub const VBA_ABI_VERSION: u32 = 2;

pub const VBA_OK: i32 = 0;
pub const VBA_DECLINED: i32 = 1;
pub const VBA_PENDING: i32 = 2;
pub const VBA_ERR_NO_POSES: i32 = -1;
pub const VBA_ERR_NO_LANDMARKS: i32 = -2;
pub const VBA_ERR_NO_OBSERVATIONS: i32 = -3;
pub const VBA_ERR_ALL_POSES_FIXED: i32 = -4;
pub const VBA_ERR_MISSING_POSE: i32 = -5;
pub const VBA_ERR_MISSING_LANDMARK: i32 = -6;
pub const VBA_ERR_UNSUPPORTED_CAMERA_MODEL: i32 = -7;
pub const VBA_ERR_MISSING_STEREO_BASELINE: i32 = -8;
pub const VBA_ERR_OBSERVATION_WEIGHT_COUNT: i32 = -9;
pub const VBA_ERR_INVALID_OBSERVATION_WEIGHT: i32 = -10;
pub const VBA_ERR_WEIGHTS_WITH_INTRINSICS: i32 = -12;
pub const VBA_ERR_SINGULAR_SYSTEM: i32 = -13;
pub const VBA_ERR_INELIGIBLE: i32 = -20;
pub const VBA_ERR_INVALID_CONFIGURATION: i32 = -21;
pub const VBA_ERR_LINEAR_SOLVE: i32 = -22;
pub const VBA_ERR_INVALID_ARGUMENT: i32 = -100;
pub const VBA_ERR_DEVICE: i32 = -101;
pub const VBA_ERR_OUT_OF_MEMORY: i32 = -102;

pub const VBA_CAMERA_PINHOLE: i32 = 0;
pub const VBA_CAMERA_SIMPLE_PINHOLE: i32 = 1;
pub const VBA_CAMERA_SIMPLE_RADIAL: i32 = 2;
pub const VBA_CAMERA_RADIAL: i32 = 3;
pub const VBA_CAMERA_OPENCV: i32 = 4;

pub const VBA_KERNEL_NONE: u32 = 0;
pub const VBA_KERNEL_HUBER: u32 = 1;
pub const VBA_KERNEL_CAUCHY: u32 = 2;

pub const VBA_SOLVER_DENSE: u32 = 0;
pub const VBA_SOLVER_SPARSE: u32 = 1;
pub const VBA_SOLVER_MATRIX_FREE: u32 = 2;
pub const VBA_SOLVER_PCG_EXPLICIT: u32 = 3;

pub const VBA_PRECISION_FP64: u32 = 0;
pub const VBA_PRECISION_MIXED: u32 = 1;

pub const VBA_PCG_MAXIT_REJECT: u32 = 0;
pub const VBA_PCG_MAXIT_ACCEPT_ITERATE: u32 = 1;

pub const VBA_FEATURE_STEREO_OBSERVATIONS: u32 = 1 << 0;
pub const VBA_FEATURE_GENERAL_STEREO: u32 = 1 << 1;
pub const VBA_FEATURE_PRIORS: u32 = 1 << 2;
pub const VBA_FEATURE_INERTIAL: u32 = 1 << 3;
pub const VBA_FEATURE_FIXED_ROTATIONS: u32 = 1 << 4;
pub const VBA_FEATURE_OBSERVATION_WEIGHTS: u32 = 1 << 5;

pub const VBA_FLAG_NO_GRAPH: u32 = 1;

pub const VBA_REFINE_INTRINSICS: u32 = 1 << 0;
pub const VBA_REFINE_DISTORTION: u32 = 1 << 1;
pub const VBA_REFINE_TANGENTIAL: u32 = 1 << 2;
pub const VBA_REFINE_SHARED_FOCAL: u32 = 1 << 3;

pub const VBA_GNC_GEMAN_MCCLURE: u32 = 0;
pub const VBA_GNC_TRUNCATED_LEAST_SQUARES: u32 = 1;

pub const VBA_DENSE_NOALLOC_MAX_N: u32 = 1024;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vba_camera {
    pub model: i32,
    pub width: u32,
    pub height: u32,
    pub num_params: u32,
    pub params: [f64; 12],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vba_sensor {
    pub camera: vba_camera,
    pub sensor_from_rig: [f64; 7],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vba_stereo_terms {
    pub baseline: f64,
    pub num_stereo: u64,
    pub stereo_pose_ids: *const u64,
    pub stereo_landmark_ids: *const u64,
    pub stereo_obs: *const f64,
    pub num_general: u64,
    pub general_pose_ids: *const u64,
    pub general_landmark_ids: *const u64,
    pub general_obs: *const f64,
    pub general_pair: *const u32,
    pub num_pairs: u64,
    pub pairs: *const vba_sensor,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vba_problem {
    pub camera: vba_camera,
    pub num_poses: u64,
    pub pose_ids: *const u64,
    pub poses: *const f64,
    pub pose_fixed: *const u8,
    pub num_landmarks: u64,
    pub landmark_ids: *const u64,
    pub landmarks: *const f64,
    pub landmark_fixed: *const u8,
    pub num_observations: u64,
    pub obs_pose_ids: *const u64,
    pub obs_landmark_ids: *const u64,
    pub obs_xy: *const f64,
    pub num_rig_observations: u64,
    pub rig_pose_ids: *const u64,
    pub rig_landmark_ids: *const u64,
    pub rig_xy: *const f64,
    pub rig_sensor: *const u32,
    pub num_sensors: u64,
    pub sensors: *const vba_sensor,
    pub declined_features: u32,
    pub reserved: u32,
    pub stereo: *const vba_stereo_terms,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vba_options {
    pub max_iterations: u32,
    pub has_initial_lambda: u32,
    pub initial_lambda: f64,
    pub lambda_increase_factor: f64,
    pub lambda_decrease_factor: f64,
    pub max_lambda: f64,
    pub min_lambda: f64,
    pub step_tolerance: f64,
    pub cost_tolerance: f64,
    pub has_relative_cost_tolerance: u32,
    pub robust_kernel: u32,
    pub relative_cost_tolerance: f64,
    pub robust_scale: f64,
    pub linear_solver: u32,
    pub precision: u32,
    pub max_pcg_iterations: u32,
    pub pcg_maxit_policy: u32,
    pub pcg_relative_tolerance: f64,
    pub pcg_absolute_tolerance: f64,
    pub refine_intrinsics: u32,
    pub flags: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vba_iteration_stats {
    pub iteration: u32,
    pub step_accepted: u32,
    pub cost_before: f64,
    pub cost_after: f64,
    pub max_pose_step: f64,
    pub max_landmark_step: f64,
    pub lambda: f64,
    pub pcg_iterations: i32,
    pub nonprojectable_after: u32,
    pub pcg_residual_norm: f64,
    pub pcg_target: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vba_result {
    pub initial_cost: f64,
    pub final_cost: f64,
    pub num_iterations: u32,
    pub converged: u32,
    pub error_id: u64,
    pub num_observations_used: u64,
    pub total_pcg_iterations: u32,
    pub linearizations: u32,
    pub device_seconds: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vba_gnc_options {
    pub kernel: u32,
    pub max_outer: u32,
    pub inner_iterations: u32,
    pub has_auto_scale: u32,
    pub c: f64,
    pub anneal_factor: f64,
    pub auto_scale: f64,
    pub auto_scale_readapt: u32,
    pub reserved: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vba_gnc_result {
    pub initial_cost: f64,
    pub final_cost: f64,
    pub inlier_cost: f64,
    pub inlier_scale: f64,
    pub observation_count: u64,
    pub outer_iterations: u32,
    pub converged: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vba_server_config {
    pub device: i32,
    pub max_window_batch: u32,
    pub window_streams: u32,
    pub submap_streams: u32,
    pub hbm_budget_bytes: u64,
    pub batch_wait_seconds: f64,
    pub build_threads: u32,
    pub reserved: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vba_tuning {
    pub pcg_chunk: u32,
    pub pcg_lookahead: u32,
    pub pcg_predict: u32,
    pub apply_waves_per_cu: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vba_reduced_system {
    pub structure_id: u64,
    pub num_nodes: u32,
    pub num_edges: u32,
    pub diag: *const f64,
    pub edges: *const u32,
    pub offdiag: *const f32,
    pub rhs: *const f64,
    pub cluster_size: u32,
    pub coarse_group: u32,
    pub cluster_inverses: *const f64,
    pub coarse_nodes: u32,
    pub max_iterations: u32,
    pub coarse_inverse: *const f64,
    pub tol: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vba_point_elimination {
    pub structure_id: u64,
    pub linearization_id: u64,
    pub num_nodes: u32,
    pub num_edges: u32,
    pub edges: *const u32,
    pub row_off: *const u64,
    pub slot_off: *const u64,
    pub slots: *const u32,
    pub num_points: u32,
    pub num_extra: u32,
    pub point_off: *const u64,
    pub point_nodes: *const u32,
    pub hcp: *const f64,
    pub point_inverse: *const f64,
    pub has_inverse: *const u8,
    pub scale: *const f64,
    pub extra_edge: *const u32,
    pub extra_blocks: *const f64,
    pub num_run_edges: u32,
    pub coarse_group: u32,
    pub run_edges: *const u32,
    pub coarse_nodes: u32,
    pub reserved: u32,
}

/// Opaque handles.
#[repr(C)]
pub struct vba_context {
    _private: [u8; 0],
}
#[repr(C)]
pub struct vba_prepared {
    _private: [u8; 0],
}
#[repr(C)]
pub struct vba_prepared_batch {
    _private: [u8; 0],
}
#[repr(C)]
pub struct vba_server {
    _private: [u8; 0],
}

#[link(name = "vba")]
extern "C" {
    pub fn vba_abi_version() -> u32;
    pub fn vba_create(device: i32) -> *mut vba_context;
    pub fn vba_destroy(ctx: *mut vba_context);
    pub fn vba_last_error(ctx: *const vba_context) -> *const c_char;
    pub fn vba_default_options(options: *mut vba_options);
    pub fn vba_check(problem: *const vba_problem, options: *const vba_options) -> i32;
    pub fn vba_optimize(ctx: *mut vba_context, problem: *const vba_problem, options: *const vba_options,
                        poses_out: *mut f64, landmarks_out: *mut f64, stats: *mut vba_iteration_stats,
                        stats_capacity: u32, result: *mut vba_result) -> i32;
    pub fn vba_optimize_intrinsics(ctx: *mut vba_context, problem: *const vba_problem, options: *const vba_options,
                                   poses_out: *mut f64, landmarks_out: *mut f64, camera_out: *mut vba_camera,
                                   stats: *mut vba_iteration_stats, stats_capacity: u32,
                                   result: *mut vba_result) -> i32;
    pub fn vba_optimize_weighted(ctx: *mut vba_context, problem: *const vba_problem, options: *const vba_options,
                                 weights: *const f64, num_weights: u64, poses_out: *mut f64,
                                 landmarks_out: *mut f64, stats: *mut vba_iteration_stats, stats_capacity: u32,
                                 result: *mut vba_result) -> i32;
    pub fn vba_default_gnc_options(gnc: *mut vba_gnc_options);
    pub fn vba_optimize_gnc(ctx: *mut vba_context, problem: *const vba_problem, options: *const vba_options,
                            gnc: *const vba_gnc_options, poses_out: *mut f64, landmarks_out: *mut f64,
                            weights_out: *mut f64, result: *mut vba_gnc_result) -> i32;
    pub fn vba_prepare(ctx: *mut vba_context, problem: *const vba_problem, options: *const vba_options,
                       out: *mut *mut vba_prepared, result: *mut vba_result) -> i32;
    pub fn vba_solve_prepared(ctx: *mut vba_context, prepared: *mut vba_prepared, options: *const vba_options,
                              stats: *mut vba_iteration_stats, stats_capacity: u32, result: *mut vba_result)
                              -> i32;
    pub fn vba_prepared_state(ctx: *mut vba_context, prepared: *const vba_prepared, poses_out: *mut f64,
                              landmarks_out: *mut f64) -> i32;
    pub fn vba_prepared_free(ctx: *mut vba_context, prepared: *mut vba_prepared);
    pub fn vba_optimize_batch(ctx: *mut vba_context, count: u32, problems: *const vba_problem,
                              options: *const vba_options, poses_out: *const *mut f64,
                              landmarks_out: *const *mut f64, stats: *mut vba_iteration_stats,
                              stats_capacity: u32, results: *mut vba_result, statuses: *mut i32) -> i32;
    pub fn vba_prepare_batch(ctx: *mut vba_context, count: u32, problems: *const vba_problem,
                             options: *const vba_options, out: *mut *mut vba_prepared_batch,
                             results: *mut vba_result, statuses: *mut i32) -> i32;
    pub fn vba_solve_prepared_batch(ctx: *mut vba_context, batch: *mut vba_prepared_batch,
                                    options: *const vba_options, stats: *mut vba_iteration_stats,
                                    stats_capacity: u32, results: *mut vba_result, statuses: *mut i32) -> i32;
    pub fn vba_prepared_batch_state(ctx: *mut vba_context, batch: *const vba_prepared_batch, window: u32,
                                    poses_out: *mut f64, landmarks_out: *mut f64) -> i32;
    pub fn vba_prepared_batch_free(ctx: *mut vba_context, batch: *mut vba_prepared_batch);
    pub fn vba_server_default_config(config: *mut vba_server_config);
    pub fn vba_server_create(config: *const vba_server_config) -> *mut vba_server;
    pub fn vba_server_submit(server: *mut vba_server, problem: *const vba_problem, options: *const vba_options)
                             -> u64;
    pub fn vba_server_submit_borrowed(server: *mut vba_server, problem: *const vba_problem,
                                      options: *const vba_options) -> u64;
    pub fn vba_server_wait(server: *mut vba_server, ticket: u64, timeout_seconds: f64, poses_out: *mut f64,
                           landmarks_out: *mut f64, result: *mut vba_result, latency_seconds: *mut f64) -> i32;
    pub fn vba_server_counters(server: *mut vba_server, completed: *mut u64, batches: *mut u64,
                               batched_windows: *mut u64, queued: *mut u64);
    pub fn vba_server_request_bytes(server: *const vba_server, problem: *const vba_problem,
                                    options: *const vba_options) -> u64;
    pub fn vba_server_destroy(server: *mut vba_server);
    pub fn vba_profile_enable(ctx: *mut vba_context, enable: i32);
    pub fn vba_profile_reset(ctx: *mut vba_context);
    pub fn vba_profile_query(ctx: *mut vba_context, kernel: *const c_char, launches: *mut u64, seconds: *mut f64,
                             bytes: *mut f64) -> i32;
    pub fn vba_set_tuning(ctx: *mut vba_context, tuning: *const vba_tuning) -> i32;
    pub fn vba_get_tuning(ctx: *const vba_context, tuning: *mut vba_tuning) -> i32;
    pub fn vba_autotune(ctx: *mut vba_context, prepared: *mut vba_prepared, options: *const vba_options,
                        lm_iterations: u32, best: *mut vba_tuning, best_seconds: *mut f64,
                        default_seconds: *mut f64) -> i32;
    pub fn vba_dense_solve(ctx: *mut vba_context, n: u32, h: *const f64, b: *const f64, x: *mut f64) -> i32;
    pub fn vba_dense_solve_noalloc(ctx: *mut vba_context, n: u32, h: *const f64, b: *const f64, x: *mut f64)
                                   -> i32;
    pub fn vba_reduced_solve(ctx: *mut vba_context, system: *const vba_reduced_system, step: *mut f64,
                             iterations: *mut u32) -> i32;
    pub fn vba_reduced_eliminate(ctx: *mut vba_context, elimination: *const vba_point_elimination,
                                 run_blocks: *mut f64, coarse_sums: *mut f64) -> i32;
}

// ---- include/vba_io.h -----------------------------------------------------

pub const VBA_ERR_IO: i32 = -110;
pub const VBA_ERR_FORMAT: i32 = -111;
pub const VBA_ERR_EXPORT: i32 = -112;

#[repr(C)]
pub struct vba_model {
    _private: [u8; 0],
}

#[link(name = "vba")]
extern "C" {
    pub fn vba_io_last_error() -> *const c_char;
    pub fn vba_model_read_ba_problem(path: *const c_char, out: *mut *mut vba_model) -> i32;
    pub fn vba_model_write_ba_problem(model: *const vba_model, path: *const c_char) -> i32;
    pub fn vba_model_read_colmap(dir: *const c_char, binary: i32, out: *mut *mut vba_model) -> i32;
    pub fn vba_model_write_colmap(model: *const vba_model, dir: *const c_char, binary: i32) -> i32;
    pub fn vba_model_from_problem(problem: *const vba_problem, weights: *const f64, out: *mut *mut vba_model)
                                  -> i32;
    pub fn vba_model_problem(model: *const vba_model, out: *mut vba_problem, weights: *mut *const f64) -> i32;
    pub fn vba_model_set_state(model: *mut vba_model, poses: *const f64, landmarks: *const f64) -> i32;
    pub fn vba_model_set_fixed(model: *mut vba_model, fixed: *const u8) -> i32;
    pub fn vba_model_keep_local_window(model: *mut vba_model, n: u32) -> i32;
    pub fn vba_model_free(model: *mut vba_model);
}
