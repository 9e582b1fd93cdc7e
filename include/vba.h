/*
 * vba.h — C ABI of the MI355X bundle-adjustment library (libvba.so).
 *
 * The library is a drop-in for the Levenberg–Marquardt Schur-complement
 * bundle-adjustment path of visloc-rs: the `BaAccelerator` trait
 * (pipelines/slam/src/ba_accel.rs:22-32) whose reference implementation is
 * `GpuBundleAdjuster` (crates/ba-gpu/src/solver.rs:729-742), with the LM
 * semantics of `BundleAdjustment::optimize` (pipelines/slam/src/bundle/
 * optimize.rs:8-28 → lm.rs:162-1017) and of the matrix-free entry
 * `optimize_matrix_free` (optimize.rs:124-139).
 *
 * Conventions (same as the reference's native modules,
 * native/descriptor_gemm/descriptor_gemm.cu:103-128):
 *   - every entry point is `extern "C"`, takes plain pointers and sizes;
 *   - a context is created once per device and destroyed by the caller;
 *   - errors are negative return codes plus a message from vba_last_error;
 *   - the library copies all inputs to the device and keeps no pointer to
 *     caller memory after a call returns.
 *
 * Ids are mapped to dense indices in ascending id order, exactly as the
 * reference builds its variable layout from `BTreeMap` keys
 * (lm.rs:234-249; solver.rs:256-280).
 */
#ifndef VBA_H
#define VBA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VBA_ABI_VERSION 2u

/* ---------------------------------------------------------------------------
 * Status codes. VBA_DECLINED is the `None` of BaAccelerator::optimize
 * (ba_accel.rs:24-31): the caller falls back to its CPU solver. Negative codes
 * map one-to-one onto `BaError` (pipelines/slam/src/bundle/mod.rs:1021-1056);
 * the offending id of MissingPose / MissingLandmark is returned in
 * vba_result.error_id.
 * ------------------------------------------------------------------------- */
enum vba_status {
    VBA_OK = 0,
    VBA_DECLINED = 1,
    VBA_PENDING = 2,                       /* vba_server_wait: the request is still running */
    VBA_ERR_NO_POSES = -1,                 /* BaError::NoPoses */
    VBA_ERR_NO_LANDMARKS = -2,             /* BaError::NoLandmarks */
    VBA_ERR_NO_OBSERVATIONS = -3,          /* BaError::NoObservations */
    VBA_ERR_ALL_POSES_FIXED = -4,          /* BaError::AllPosesFixed */
    VBA_ERR_MISSING_POSE = -5,             /* BaError::MissingPose(id) */
    VBA_ERR_MISSING_LANDMARK = -6,         /* BaError::MissingLandmark(id) */
    VBA_ERR_UNSUPPORTED_CAMERA_MODEL = -7, /* BaError::UnsupportedCameraModel */
    VBA_ERR_MISSING_STEREO_BASELINE = -8,  /* BaError::MissingStereoBaseline */
    VBA_ERR_OBSERVATION_WEIGHT_COUNT = -9, /* BaError::ObservationWeightCount */
    VBA_ERR_INVALID_OBSERVATION_WEIGHT = -10, /* BaError::InvalidObservationWeight(index in error_id) */
    VBA_ERR_WEIGHTS_WITH_INTRINSICS = -12, /* BaError::ObservationWeightsWithIntrinsicsRefinement */
    VBA_ERR_SINGULAR_SYSTEM = -13,         /* BaError::SingularSystem */
    /* Matrix-free entry errors (MatrixFreeBaError, mod.rs:893-910). */
    VBA_ERR_INELIGIBLE = -20,              /* MatrixFreeBaError::Ineligible */
    VBA_ERR_INVALID_CONFIGURATION = -21,   /* MatrixFreeBaError::InvalidConfiguration */
    VBA_ERR_LINEAR_SOLVE = -22,            /* MatrixFreeBaError::LinearSolve */
    /* Library errors with no BaError counterpart. */
    VBA_ERR_INVALID_ARGUMENT = -100,
    VBA_ERR_DEVICE = -101,
    VBA_ERR_OUT_OF_MEMORY = -102
};

/* Camera models: the COLMAP pinhole family that `BundleAdjustment` admits on
 * this path (crates/core/src/types/camera.rs:15-41). Parameter layouts:
 *   PINHOLE        [fx, fy, cx, cy (, k1, k2)]
 *   SIMPLE_PINHOLE [f, cx, cy]
 *   SIMPLE_RADIAL  [f, cx, cy, k]
 *   RADIAL         [f, cx, cy, k1, k2]
 *   OPENCV         [fx, fy, cx, cy, k1, k2, p1, p2]                        */
enum vba_camera_model {
    VBA_CAMERA_PINHOLE = 0,
    VBA_CAMERA_SIMPLE_PINHOLE = 1,
    VBA_CAMERA_SIMPLE_RADIAL = 2,
    VBA_CAMERA_RADIAL = 3,
    VBA_CAMERA_OPENCV = 4
};

/* Robust kernels, `RobustKernel` (pipelines/slam/src/pose_graph.rs:87-140). */
enum vba_robust_kernel {
    VBA_KERNEL_NONE = 0,
    VBA_KERNEL_HUBER = 1, /* robust_scale = delta */
    VBA_KERNEL_CAUCHY = 2 /* robust_scale = c */
};

/* Reduced-camera solvers. DENSE is `LinearSolver::Dense` (dense Schur +
 * Cholesky with LU fallback, normal_equations.rs:1403-1504); MATRIX_FREE is
 * the implicit-Schur block-Jacobi PCG of `optimize_matrix_free`
 * (implicit_schur.rs:133-824); PCG_EXPLICIT is the same PCG on an explicitly
 * assembled block-sparse Schur matrix. */
enum vba_linear_solver {
    VBA_SOLVER_DENSE = 0,
    /* LinearSolver::Sparse (normal_equations.rs:1483-1535): the reduced
     * system by Cholesky only, a failed factor is SingularSystem (no LU
     * fallback); factored densely on the device */
    VBA_SOLVER_SPARSE = 1,
    VBA_SOLVER_MATRIX_FREE = 2,
    VBA_SOLVER_PCG_EXPLICIT = 3
};

/* Arithmetic of the device solve. FP64 is the plain double-precision solver.
 * MIXED keeps residual/Jacobian storage, Schur products and PCG matvecs in
 * FP32, the 3x3 / 6x6 block inverses, cost sums, the outer PCG residual and
 * the LM step in FP64 (iterative refinement of the step). */
enum vba_precision {
    VBA_PRECISION_FP64 = 0,
    VBA_PRECISION_MIXED = 1
};

/* What a PCG solve that reaches max_pcg_iterations without meeting its
 * target does. REJECT is the CPU matrix-free semantics
 * (implicit_schur.rs:663-690: MaxIterations → failed solve → LM reject with
 * lambda increase); ACCEPT_ITERATE is the reference GPU backend's
 * (crates/ba-gpu/src/solver.rs:497-508: use the last iterate). */
enum vba_pcg_maxit_policy {
    VBA_PCG_MAXIT_REJECT = 0,
    VBA_PCG_MAXIT_ACCEPT_ITERATE = 1
};

/* Terms of `BundleAdjustment` that this ABI does not carry. A caller sets the
 * bit when its problem has them; the library then declines (as
 * GpuBundleAdjuster::check does, solver.rs:212-251). Stereo observations
 * are carried by vba_problem.stereo; the two stereo bits are for callers
 * that do not pass them. */
enum vba_declined_feature {
    VBA_FEATURE_STEREO_OBSERVATIONS = 1u << 0,
    VBA_FEATURE_GENERAL_STEREO = 1u << 1,
    VBA_FEATURE_PRIORS = 1u << 2,          /* gravity / position / pairwise */
    VBA_FEATURE_INERTIAL = 1u << 3,        /* velocities, biases, IMU, nav prior */
    VBA_FEATURE_FIXED_ROTATIONS = 1u << 4, /* fixed_pose_rotations */
    VBA_FEATURE_OBSERVATION_WEIGHTS = 1u << 5
};

typedef struct vba_camera {
    int32_t model;       /* enum vba_camera_model */
    uint32_t width;
    uint32_t height;
    uint32_t num_params;
    double params[12];
} vba_camera;

/* A rig sensor: its lens and the fixed extrinsic T_sensor<-rig
 * (BaRigObservation::{camera, sensor_from_rig}, bundle/mod.rs:244-251). */
typedef struct vba_sensor {
    vba_camera camera;
    double sensor_from_rig[7]; /* qw qx qy qz tx ty tz */
} vba_sensor;

/* Stereo terms of a `BundleAdjustment`: the rectified stereo observations
 * with their baseline (BaStereoObservation, bundle/mod.rs:207-221;
 * stereo_baseline, problem.rs:233-235) and the calibrated general stereo
 * observations (BaGeneralStereoObservation, bundle/mod.rs:223-237). Both need
 * a PINHOLE or SIMPLE_PINHOLE `camera` (problem.rs:733-746) and project with
 * its intrinsics only; a general pair's right camera also contributes its
 * pinhole intrinsics only (sqrt_qr.rs:465-515). */
typedef struct vba_stereo_terms {
    double baseline;                    /* stereo_baseline; NaN or <= 0: none */
    uint64_t num_stereo;
    const uint64_t* stereo_pose_ids;
    const uint64_t* stereo_landmark_ids;
    const double* stereo_obs;           /* 3 per observation: u_l, v_l, u_r */
    uint64_t num_general;
    const uint64_t* general_pose_ids;
    const uint64_t* general_landmark_ids;
    const double* general_obs;          /* 4 per observation: left x, y, right x, y */
    const uint32_t* general_pair;       /* index into pairs */
    uint64_t num_pairs;
    const vba_sensor* pairs;            /* right_camera, left_to_right = T_right<-left */
} vba_stereo_terms;

/* Caller-owned view of a `BundleAdjustment` (bundle/mod.rs:529-622).
 * Poses are T_world->camera (T_rig<-world for rig observations) as a unit
 * quaternion (w, x, y, z) and a translation, 7 doubles each. */
typedef struct vba_problem {
    vba_camera camera;               /* BundleAdjustment::camera */
    uint64_t num_poses;
    const uint64_t* pose_ids;        /* unique */
    const double* poses;             /* 7 per pose */
    const uint8_t* pose_fixed;       /* nullable; nonzero = in fixed_poses */
    uint64_t num_landmarks;
    const uint64_t* landmark_ids;    /* unique */
    const double* landmarks;         /* 3 per landmark */
    const uint8_t* landmark_fixed;   /* nullable; nonzero = in fixed_landmarks */
    uint64_t num_observations;       /* BaObservation list (mod.rs:193-199) */
    const uint64_t* obs_pose_ids;
    const uint64_t* obs_landmark_ids;
    const double* obs_xy;            /* 2 per observation */
    uint64_t num_rig_observations;   /* BaRigObservation list (mod.rs:244-251) */
    const uint64_t* rig_pose_ids;
    const uint64_t* rig_landmark_ids;
    const double* rig_xy;            /* 2 per rig observation */
    const uint32_t* rig_sensor;      /* index into sensors */
    uint64_t num_sensors;
    const vba_sensor* sensors;
    uint32_t declined_features;      /* enum vba_declined_feature bits */
    uint32_t reserved;
    const vba_stereo_terms* stereo;  /* nullable; observations are flattened as the
                                        reference does: mono, rectified stereo,
                                        general stereo, rig */
} vba_problem;

/* `BaConfig` (bundle/mod.rs:625-752) plus the PCG settings of
 * `MatrixFreeBaOptions` (mod.rs:780-798) and `GpuBaSettings`
 * (crates/ba-gpu/src/solver.rs:24-50). vba_default_options gives
 * BaConfig::default() and MatrixFreeBaOptions::default(). */
typedef struct vba_options {
    uint32_t max_iterations;
    uint32_t has_initial_lambda;     /* 0 = None: pure Gauss-Newton */
    double initial_lambda;
    double lambda_increase_factor;
    double lambda_decrease_factor;
    double max_lambda;
    double min_lambda;
    double step_tolerance;
    double cost_tolerance;
    uint32_t has_relative_cost_tolerance;
    uint32_t robust_kernel;          /* enum vba_robust_kernel */
    double relative_cost_tolerance;
    double robust_scale;
    uint32_t linear_solver;          /* enum vba_linear_solver */
    uint32_t precision;              /* enum vba_precision */
    uint32_t max_pcg_iterations;
    uint32_t pcg_maxit_policy;       /* enum vba_pcg_maxit_policy */
    double pcg_relative_tolerance;
    double pcg_absolute_tolerance;
    uint32_t refine_intrinsics;      /* enum vba_refine bits; vba_optimize declines them,
                                        vba_optimize_intrinsics runs the joint solve */
    uint32_t flags;                  /* enum vba_flags, 0 by default */
} vba_options;

/* vba_options.flags. A prepared batch of windows (vba_solve_prepared_batch)
 * records its LM iterations as one HIP graph on its first solve and replays
 * it on later solves with the same options; VBA_FLAG_NO_GRAPH launches the
 * kernels one by one instead, with the same results. */
enum vba_flags { VBA_FLAG_NO_GRAPH = 1 };

/* `BaIterationStats` (mod.rs:754-763) plus the matrix-free diagnostics of
 * `MatrixFreeBaIterationStats` (mod.rs:810-818). */
typedef struct vba_iteration_stats {
    uint32_t iteration;
    uint32_t step_accepted;
    double cost_before;
    double cost_after;
    double max_pose_step;
    double max_landmark_step;
    double lambda;                   /* post-update lambda, as the reference records */
    int32_t pcg_iterations;          /* -1 when the solver is not PCG */
    uint32_t nonprojectable_after;
    double pcg_residual_norm;
    double pcg_target;
} vba_iteration_stats;

/* `BaResult` (mod.rs:765-771) plus counters the harness reports. */
typedef struct vba_result {
    double initial_cost;
    double final_cost;
    uint32_t num_iterations;         /* entries written to the stats array (capped) */
    uint32_t converged;
    uint64_t error_id;               /* id for MISSING_POSE / MISSING_LANDMARK */
    uint64_t num_observations_used;  /* N of the obs/s metric */
    uint32_t total_pcg_iterations;
    uint32_t linearizations;
    double device_seconds;           /* device time of the solve (HIP events) */
} vba_result;

typedef struct vba_context vba_context;

/* ---- lifecycle (the reference's *_abi_version/_create/_destroy/_last_error) */
uint32_t vba_abi_version(void);
/* Create a context on a HIP device ordinal. NULL on failure (see
 * vba_last_error(NULL)). */
struct vba_context* vba_create(int32_t device);
void vba_destroy(struct vba_context* ctx);
/* Last error message of ctx, or of the calling thread when ctx is NULL. */
const char* vba_last_error(const struct vba_context* ctx);

/* BaConfig::default() + MatrixFreeBaOptions::default()
 * (mod.rs:729-752, 789-798). */
void vba_default_options(vba_options* options);

/* GpuBundleAdjuster::check (solver.rs:212-251), widened to the lens models
 * and kernels this library implements: returns VBA_OK, or VBA_DECLINED with
 * the reason in vba_last_error(NULL). */
int32_t vba_check(const vba_problem* problem, const vba_options* options);

/* BaAccelerator::optimize (ba_accel.rs:22-32; solver.rs:552-690, 729-742).
 * On VBA_OK the final poses (7 per pose) and landmarks (3 per landmark), in
 * the caller's input order, are written to poses_out / landmarks_out (which
 * may alias the input arrays), and up to stats_capacity iteration records to
 * stats. On an error nothing is written to poses_out / landmarks_out, as the
 * reference leaves `ba` untouched when validation fails. */
int32_t vba_optimize(vba_context* ctx, const vba_problem* problem,
                     const vba_options* options, double* poses_out,
                     double* landmarks_out, vba_iteration_stats* stats,
                     uint32_t stats_capacity, vba_result* result);

/* vba_options.refine_intrinsics: BaConfig::refine_intrinsics,
 * refine_distortion, refine_tangential_distortion and shared_focal
 * (bundle/mod.rs:655-700). */
enum vba_refine {
    VBA_REFINE_INTRINSICS = 1u << 0,   /* (fx, fy, cx, cy) */
    VBA_REFINE_DISTORTION = 1u << 1,   /* + (k1, k2), monocular problems only */
    VBA_REFINE_TANGENTIAL = 1u << 2,   /* + (p1, p2) with DISTORTION; the camera becomes OPENCV */
    VBA_REFINE_SHARED_FOCAL = 1u << 3  /* fx == fy, starting from their mean */
};

/* BundleAdjustment::optimize with refine_intrinsics set (optimize.rs:8-28):
 * for a PINHOLE or OPENCV camera with intrinsics and no general stereo or
 * rig observations, the joint pose / landmark / camera solve
 * (optimize_joint_intrinsics, joint_intrinsics.rs:24-353): the 4, 6 or 8
 * camera parameters border the dense reduced camera system (the Schur
 * complement also eliminates the landmarks from the camera block), always
 * with the dense solver, and the camera is refined even when every pose and
 * landmark is fixed. Otherwise, as the reference does, the pose / structure
 * solve of vba_optimize with refine_intrinsics cleared. camera_out receives
 * the refined camera: its model (OPENCV when TANGENTIAL promoted it), its
 * parameter count (at least 6 with DISTORTION, 8 with TANGENTIAL) and
 * values; it is the input camera when the camera is not refined. */
int32_t vba_optimize_intrinsics(vba_context* ctx, const vba_problem* problem, const vba_options* options,
                                double* poses_out, double* landmarks_out, vba_camera* camera_out,
                                vba_iteration_stats* stats, uint32_t stats_capacity, vba_result* result);

/* BundleAdjustment::optimize_with_observation_weights (optimize.rs:43-53):
 * `weights` holds num_weights values, one per observation in the reference's
 * flattened order (monocular, rectified stereo, general stereo, rig).
 * Each multiplies its observation's robust-kernel weight in the normal
 * equations (normal_equations.rs:458, 627) and its cost term
 * (problem.rs:521-523, 583). Checked first, as validate_observation_weights
 * does (optimize.rs:628-650): VBA_ERR_OBSERVATION_WEIGHT_COUNT, then
 * VBA_ERR_INVALID_OBSERVATION_WEIGHT for the first negative or non-finite
 * weight (its index in result.error_id), then
 * VBA_ERR_WEIGHTS_WITH_INTRINSICS. The reference solves weighted problems
 * with its normal-equation path (the matrix-free entry does not take weights,
 * optimize.rs:87-103); here every linear solver takes them. */
int32_t vba_optimize_weighted(vba_context* ctx, const vba_problem* problem, const vba_options* options,
                              const double* weights, uint64_t num_weights, double* poses_out,
                              double* landmarks_out, vba_iteration_stats* stats, uint32_t stats_capacity,
                              vba_result* result);

/* Graduated non-convexity, BundleAdjustment::optimize_gnc (lm.rs:28-142;
 * pipelines/slam/src/gnc.rs): weighted LM solves with no M-estimator,
 * re-weighted between them by the surrogate's Black–Rangarajan weights while
 * the control parameter anneals. Fields mirror `GncConfig` (gnc.rs:52-100;
 * defaults gnc.rs:103-115 via vba_default_gnc_options). */
enum vba_gnc_kernel {
    VBA_GNC_GEMAN_MCCLURE = 0,
    VBA_GNC_TRUNCATED_LEAST_SQUARES = 1
};
typedef struct vba_gnc_options {
    uint32_t kernel;             /* enum vba_gnc_kernel */
    uint32_t max_outer;
    uint32_t inner_iterations;
    uint32_t has_auto_scale;     /* auto_scale: Some(k) when nonzero */
    double c;
    double anneal_factor;
    double auto_scale;           /* k of median + k 1.4826 MAD */
    uint32_t auto_scale_readapt;
    uint32_t reserved;
} vba_gnc_options;
/* BaGncResult (bundle/mod.rs:973-1010). */
typedef struct vba_gnc_result {
    double initial_cost;
    double final_cost;
    double inlier_cost;
    double inlier_scale;
    uint64_t observation_count;
    uint32_t outer_iterations;
    uint32_t converged;
} vba_gnc_result;
void vba_default_gnc_options(vba_gnc_options* gnc);
/* `options` is the inner LM's BaConfig (its kernel and iteration count are
 * replaced as the reference does). On VBA_OK the final state is written as by
 * vba_optimize, and weights_out (nullable) receives the final weights, one per
 * observation in the flattened order (monocular, rectified stereo, general
 * stereo, rig), NaN where an observation cannot be evaluated. Errors of an inner solve are returned as they are. */
int32_t vba_optimize_gnc(vba_context* ctx, const vba_problem* problem, const vba_options* options,
                         const vba_gnc_options* gnc, double* poses_out, double* landmarks_out,
                         double* weights_out, vba_gnc_result* result);

/* ---- device-resident problems ------------------------------------------
 * No counterpart in the reference, whose GPU backend uploads on every call
 * (solver.rs:253-385). vba_optimize is vba_prepare + vba_solve_prepared +
 * vba_prepared_state + vba_prepared_free. Callers that solve one problem
 * repeatedly (benchmarks, the serving loop) prepare it once: validation,
 * layout and the host->device copy happen in vba_prepare; every
 * vba_solve_prepared starts from the prepared initial state and leaves its
 * final state on the device. Status codes are those of vba_optimize. */
typedef struct vba_prepared vba_prepared;
int32_t vba_prepare(vba_context* ctx, const vba_problem* problem, const vba_options* options,
                    vba_prepared** out, vba_result* result);
int32_t vba_solve_prepared(vba_context* ctx, vba_prepared* prepared, const vba_options* options,
                           vba_iteration_stats* stats, uint32_t stats_capacity,
                           vba_result* result);
/* Final state of the last solve (the initial state before any), in the
 * caller's input order. */
int32_t vba_prepared_state(vba_context* ctx, const vba_prepared* prepared, double* poses_out,
                           double* landmarks_out);
void vba_prepared_free(vba_context* ctx, vba_prepared* prepared);

/* ---- batched local windows ----------------------------------------------
 * No counterpart in the reference, whose local BA calls
 * BaAccelerator::optimize once per window (refinement.rs:1067-1072). Solves
 * `count` independent problems together: one device-resident batch whose LM
 * control (lambda, accept/reject, convergence) runs per window on the
 * device, with no host round trip per iteration. Window i gets the results
 * vba_optimize gives it alone with the dense solver: the same accept/reject
 * sequence and the same arithmetic. `options` applies to every window and
 * must select VBA_SOLVER_DENSE; a window has at most 256 variable poses.
 * statuses[i] is window i's status (the code vba_optimize would return);
 * windows with an error are skipped and their outputs left untouched.
 * poses_out[i] / landmarks_out[i] (either array, or an entry, may be NULL)
 * receive window i's final state in its input order; stats holds
 * count x stats_capacity records, window i's at i * stats_capacity.
 * Returns VBA_OK when the batch ran (see statuses), else the error that
 * stopped the whole batch. */
int32_t vba_optimize_batch(vba_context* ctx, uint32_t count, const vba_problem* problems,
                           const vba_options* options, double* const* poses_out,
                           double* const* landmarks_out, vba_iteration_stats* stats,
                           uint32_t stats_capacity, vba_result* results, int32_t* statuses);

/* The batch kept on the device, as vba_prepare does for one problem:
 * vba_optimize_batch is vba_prepare_batch + vba_solve_prepared_batch +
 * vba_prepared_batch_state per window + vba_prepared_batch_free. Every
 * solve starts from the prepared initial state. results may be NULL in
 * vba_prepare_batch (it then only reports statuses). */
typedef struct vba_prepared_batch vba_prepared_batch;
int32_t vba_prepare_batch(vba_context* ctx, uint32_t count, const vba_problem* problems,
                          const vba_options* options, vba_prepared_batch** out,
                          vba_result* results, int32_t* statuses);
int32_t vba_solve_prepared_batch(vba_context* ctx, vba_prepared_batch* batch,
                                 const vba_options* options, vba_iteration_stats* stats,
                                 uint32_t stats_capacity, vba_result* results, int32_t* statuses);
/* Final state of window `window` (the caller's index) after the last solve. */
int32_t vba_prepared_batch_state(vba_context* ctx, const vba_prepared_batch* batch, uint32_t window,
                                 double* poses_out, double* landmarks_out);
void vba_prepared_batch_free(vba_context* ctx, vba_prepared_batch* batch);

/* ---- serving: continuous batching of independent solves -----------------
 * No counterpart in the reference. A server owns one device and a request
 * loop: vba_server_submit copies a problem and returns at once; lane threads
 * admit queued requests as streams and device memory free up and solve them
 * while others run. Dense requests with at most 256 variable poses
 * (pose-only refinements, local windows) are gathered, per equal options,
 * into device batches of up to max_window_batch windows; other requests
 * (matrix-free, large dense) are solved one per submap stream. Admission
 * keeps the estimated device footprint of running work within
 * hbm_budget_bytes. A result equals the one vba_optimize gives the request
 * alone. */
typedef struct vba_server_config {
    int32_t device;
    uint32_t max_window_batch;    /* windows per device batch */
    uint32_t window_streams;      /* window batches in flight at once */
    uint32_t submap_streams;      /* single-problem solves in flight at once (with window_streams, 16 lanes at most) */
    uint64_t hbm_budget_bytes;    /* admission limit on resident work */
    double batch_wait_seconds;    /* how long a window batch waits to fill */
    uint32_t build_threads;       /* host threads laying out a window batch */
    uint32_t reserved;
} vba_server_config;
typedef struct vba_server vba_server;
void vba_server_default_config(vba_server_config* config);
/* NULL on failure (vba_last_error(NULL)). */
vba_server* vba_server_create(const vba_server_config* config);
/* Queue a solve (the problem is copied). Returns a ticket > 0, or 0 on
 * failure (vba_last_error(NULL)). */
uint64_t vba_server_submit(vba_server* server, const vba_problem* problem, const vba_options* options);
/* As vba_server_submit without the copy: the caller keeps every array of
 * `problem` valid and unchanged until the ticket is retired. */
uint64_t vba_server_submit_borrowed(vba_server* server, const vba_problem* problem,
                                    const vba_options* options);
/* Wait up to timeout_seconds for a ticket. VBA_PENDING while it runs;
 * otherwise the solve's status (the ticket is then retired), with outputs as
 * vba_optimize writes them and the latency from submission to completion. */
int32_t vba_server_wait(vba_server* server, uint64_t ticket, double timeout_seconds, double* poses_out,
                        double* landmarks_out, vba_result* result, double* latency_seconds);
void vba_server_counters(vba_server* server, uint64_t* completed, uint64_t* batches,
                         uint64_t* batched_windows, uint64_t* queued);
/* The device bytes the server counts against hbm_budget_bytes while it runs
 * this request (an upper estimate of its buffers); 0 for a null argument.
 * A batch holds the sum of its windows'. Work is admitted while the running
 * total plus the request stays within the budget; a request larger than the
 * budget runs only when nothing else does. */
uint64_t vba_server_request_bytes(const vba_server* server, const vba_problem* problem,
                                  const vba_options* options);
/* Finishes the queued requests, then stops. */
void vba_server_destroy(vba_server* server);

/* ---- per-kernel device time ----------------------------------------------
 * While enabled, HIP events time the launches of the kernel families
 * "linearize", "schur_apply", "pcg_pose", "pcg_step", "cost",
 * "explicit_assemble" and "explicit_apply": every launch for enable = 1,
 * every enable-th launch of each family for enable > 1 (sampling keeps the
 * events' host cost out of the solve), none for enable = 0. The matrix-free
 * PCG loop's "schur_apply", "pcg_pose" and "pcg_step" launches carry the
 * events on their dispatch (the kernel's own start and end); the others are
 * bracketed by events on the context's stream, which also count the wait
 * before the dispatch. vba_profile_query returns the recorded launches that did
 * work, their summed device seconds and the algorithmic bytes they moved
 * (the minimum HBM traffic of the data layout, summed over launches). */
void vba_profile_enable(vba_context* ctx, int32_t enable);
void vba_profile_reset(vba_context* ctx);
int32_t vba_profile_query(vba_context* ctx, const char* kernel, uint64_t* launches,
                          double* seconds, double* bytes);

/* ---- PCG launch schedule and its autotuner --------------------------------
 * The reference's GPU backend dispatches PCG in chunks of 10 iterations and
 * reads its done flag between chunks (crates/ba-gpu/src/solver.rs:493-510;
 * VISLOC_BA_GPU_PCG_CHUNK); the chunk is the first field below, its default
 * the reference's. The device decides when PCG stops and launches after the stop do
 * nothing, so no field here changes a result, only how much host-device
 * waiting and how many empty launches a solve costs:
 * - pcg_chunk: iterations enqueued per read of the device PCG state (>= 1;
 *   default 10);
 * - pcg_lookahead: chunks enqueued beyond the one whose state is awaited
 *   (0..2; default 0), so that the device does not idle while the host reads;
 * - pcg_predict: nonzero: a solve's first chunk runs to the iteration count
 *   of the previous PCG solve of the same prepared problem (default 0).
 * The schedule is per context and applies to the matrix-free and explicit
 * PCG solvers. */
typedef struct vba_tuning {
    uint32_t pcg_chunk;
    uint32_t pcg_lookahead;
    uint32_t pcg_predict;
    /* 0: the PCG product launches one wave per landmark tile; n (1..32): a
     * persistent product of n waves per compute unit looping over the tiles
     * (the same results bit for bit); mixed precision only, the FP64
     * product always takes one wave per tile */
    uint32_t apply_waves_per_cu;
} vba_tuning;
/* NULL restores the defaults. VBA_ERR_INVALID_ARGUMENT for a field out of
 * range. */
int32_t vba_set_tuning(vba_context* ctx, const vba_tuning* tuning);
int32_t vba_get_tuning(const vba_context* ctx, vba_tuning* tuning);
/* Times each schedule of a fixed candidate set on `prepared` (a PCG solver in
 * `options`, max_iterations replaced by lm_iterations when that is nonzero;
 * each candidate solved three times, the fastest kept), checks that every candidate
 * gives the default schedule's result bit for bit (VBA_ERR_DEVICE
 * otherwise), sets the fastest on the context and returns it in *best with
 * its seconds and the default schedule's. */
int32_t vba_autotune(vba_context* ctx, vba_prepared* prepared, const vba_options* options,
                     uint32_t lm_iterations, vba_tuning* best, double* best_seconds,
                     double* default_seconds);

/* ---- dense normal equations ----------------------------------------------
 * `solve_normal_equations` (pipelines/slam/src/pose_graph.rs:2259-2270), the
 * dense solve behind LinearSolver::Dense (normal_equations.rs:1478), the
 * joint-intrinsics solve (joint_intrinsics.rs:205-208) and the pose graphs
 * (pose_graph.rs:619, sim3_pose_graph.rs:286): H x = b by Cholesky, and by
 * LU with partial pivoting when H is not positive definite. H is n x n
 * column-major (nalgebra's DMatrix layout); Cholesky reads its lower
 * triangle. VBA_OK with x, or VBA_ERR_SINGULAR_SYSTEM
 * (PoseGraphError::SingularSystem) when LU finds H singular. */
int32_t vba_dense_solve(vba_context* ctx, uint32_t n, const double* h, const double* b, double* x);

/* vba_dense_solve without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, layout, results and errors as
 * vba_dense_solve, for n <= VBA_DENSE_NOALLOC_MAX_N unknowns
 * (VBA_ERR_INVALID_ARGUMENT beyond). Every device buffer of the solve lives
 * in one fixed device array of the library (about 16 MiB for the maximum n,
 * held on each GPU the library runs on until the process ends) and the
 * context's pinned status word, both set up by vba_create; the call
 * allocates nothing. All calls on one GPU share that array: calls that
 * overlap on one GPU, from any context or thread, race. */
#define VBA_DENSE_NOALLOC_MAX_N 1024u
int32_t vba_dense_solve_noalloc(vba_context* ctx, uint32_t n, const double* h, const double* b, double* x);

/* ---- reduced camera solve ------------------------------------------------
 * The iterative reduced-camera solve of the COLMAP-port rig solver,
 * `ReducedSolveAccelerator::solve` (pipelines/slam/src/ba_accel.rs:46-157;
 * caller colmap_incremental/rig_ba_solver.rs:3072-3101): the same
 * preconditioned conjugate gradients as its CPU pcg_core (:1406-1490), with
 * the same preconditioner and stopping rule, in FP64 arithmetic on the FP32
 * blocks the caller stores. Fields mirror `ReducedCameraSystem`:
 * - diag: the damped 6x6 diagonal block of each free node, column-major
 *   (nalgebra's Matrix6 layout);
 * - edges: edge e couples nodes (edges[2e], edges[2e+1]) = (a, b), a < b,
 *   sorted; offdiag[36e ..] is its (row b, column a) block, column-major;
 * - rhs: 6 per node; tol: stop once |r| <= tol |rhs|; max_iterations: the
 *   iteration cap (the reference passes 500);
 * - cluster_size > 1 with cluster_inverses: `clusters`, the dense inverses
 *   of the system restricted to runs of cluster_size consecutive nodes
 *   (run c starts at node c * cluster_size; the last may be shorter), each
 *   column-major (6m)^2, concatenated; 0 or 1: block-Jacobi;
 * - coarse_group > 0 with coarse_nodes and coarse_inverse: `coarse`, the
 *   additive coarse level over the leading coarse_nodes nodes in aggregates
 *   of coarse_group, its inverse column-major (6M)^2 with
 *   M = ceil(coarse_nodes / coarse_group);
 * - structure_id: identifies `edges` across calls (0: unknown); while it
 *   stays the same the library reuses the edge index it derived.
 * The point elimination hook (`ReducedSolveAccelerator::eliminate`) is not
 * offered: blocks always come in offdiag. */
typedef struct vba_reduced_system {
    uint64_t structure_id;
    uint32_t num_nodes;
    uint32_t num_edges;
    const double* diag;
    const uint32_t* edges;
    const float* offdiag;
    const double* rhs;
    uint32_t cluster_size;
    uint32_t coarse_group;
    const double* cluster_inverses;
    uint32_t coarse_nodes;
    uint32_t max_iterations;
    const double* coarse_inverse;
    double tol;
} vba_reduced_system;

/* `ReducedSolveAccelerator::solve`: VBA_OK with step[6 * num_nodes] and the
 * iteration count, or VBA_DECLINED (`None`, the caller solves on the CPU)
 * when the system is outside what the device solve supports (unsorted edges,
 * clusters longer than 256 nodes, more than 512 coarse aggregates, non-finite
 * input) or when the iteration breaks down as pcg_core returns None (a
 * diagonal block not positive definite, p^T A p not positive and finite).
 * offdiag == NULL with edges: the blocks the last vba_reduced_eliminate of
 * the same (nonzero) structure_id formed on this context (declined when
 * there are none). vba_last_error(ctx) says which. */
int32_t vba_reduced_solve(vba_context* ctx, const vba_reduced_system* system, double* step,
                          uint32_t* iterations);

/* ---- point elimination ---------------------------------------------------
 * `ReducedSolveAccelerator::eliminate` (pipelines/slam/src/ba_accel.rs:71-157;
 * caller colmap_incremental/rig_ba_solver.rs:2733-2855): the off-diagonal
 * blocks of the reduced camera system formed on the device,
 *   block(e = (s, t)) = -sum_p hcp[p, j] V_p^-1 hcp[p, k]^T + extra(e)
 * over the points p whose k-th and j-th nodes are s and t (ascending p), the
 * (row t, column s) block as in vba_reduced_system. Fields mirror
 * `PointElimination`:
 * - edges / row_off: the system's sorted edges (s, t), s < t, and per node
 *   the first of its edges (s, _); num_nodes + 1 offsets;
 * - slot_off / slots: per node, the (point, k) pairs of the points observing
 *   it, ascending by point (2 values per slot);
 * - point_off / point_nodes: point p's nodes, ascending, at
 *   point_nodes[point_off[p] .. point_off[p + 1]];
 * - hcp: J_node^T J_point per (point, k), 6x3 column-major (18 values);
 * - point_inverse: the damped point inverses, 3x3 column-major (9 values),
 *   has_inverse[p] == 0 for `None` (the point is left out);
 * - scale: the per-node Cauchy-Schwarz scales (6 per node); the device sums
 *   in FP64 and does not need them, they are accepted for the interface;
 * - extra_edge / extra_blocks: direct couplings added to the eliminated
 *   blocks (36 values each, column-major);
 * - run_edges: edges whose final blocks are returned in run_blocks (36 each,
 *   column-major, FP64), in that order;
 * - coarse_group > 0: also return coarse_sums[36 (g m + h)], g <= h, the sum
 *   of the blocks of edges from aggregate g to h (m = ceil(coarse_nodes /
 *   coarse_group) aggregates of the leading coarse_nodes nodes), summed in
 *   edge order; entries g > h are zero.
 * structure_id (nonzero) covers every field but hcp, point_inverse,
 * has_inverse, scale and the extra couplings; while it stays the same the
 * device keeps its pair index. The blocks stay on the device for the next
 * vba_reduced_solve of the same structure with offdiag == NULL.
 * VBA_OK, or VBA_DECLINED (`None`: the caller eliminates on the CPU) for
 * inconsistent input; vba_last_error(ctx) says why. */
typedef struct vba_point_elimination {
    uint64_t structure_id;
    uint64_t linearization_id;
    uint32_t num_nodes;
    uint32_t num_edges;
    const uint32_t* edges;
    const uint64_t* row_off;
    const uint64_t* slot_off;
    const uint32_t* slots;
    uint32_t num_points;
    uint32_t num_extra;
    const uint64_t* point_off;
    const uint32_t* point_nodes;
    const double* hcp;
    const double* point_inverse;
    const uint8_t* has_inverse;
    const double* scale;
    const uint32_t* extra_edge;
    const double* extra_blocks;
    uint32_t num_run_edges;
    uint32_t coarse_group;
    const uint32_t* run_edges;
    uint32_t coarse_nodes;
    uint32_t reserved;
} vba_point_elimination;

int32_t vba_reduced_eliminate(vba_context* ctx, const vba_point_elimination* elimination, double* run_blocks,
                              double* coarse_sums);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VBA_H */
