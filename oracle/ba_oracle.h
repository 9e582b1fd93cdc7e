/*
 * ba_oracle.h — C API of the CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain-C++ restatement of the
 * reference CPU bundle adjuster (visloc-rs pipelines/slam/src/bundle/) used
 * to check the HIP library. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline; the product library never links or calls it.
 *
 * It shares the problem/options/result structs of include/vba.h so the same
 * caller-side buffers drive both.
 */
#ifndef BA_ORACLE_H
#define BA_ORACLE_H

#include "../include/vba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* BundleAdjustment::optimize / optimize_matrix_free restated in FP64
 * (lm.rs:162-1017). Same contract as vba_optimize. `threads` > 1 computes the
 * per-observation contributions on that many threads and scatters them
 * serially in observation order, so the result is bitwise equal to
 * threads = 1 (the reference's `BaConfig::parallel`, bundle/mod.rs:22-62). */
int32_t oracle_optimize(const vba_problem* problem, const vba_options* options,
                        double* poses_out, double* landmarks_out,
                        vba_iteration_stats* stats, uint32_t stats_capacity,
                        vba_result* result, int32_t threads);

/* BundleAdjustment::optimize_with_observation_weights (optimize.rs:43-53,
 * 628-650): the external weight multiplies each observation's robust weight
 * (normal_equations.rs:458, 627) and its cost term (problem.rs:521-523, 583).
 * Same contract as vba_optimize_weighted. */
/* BundleAdjustment::optimize with BaConfig::refine_intrinsics
 * (optimize.rs:8-28): the joint pose / landmark / intrinsics solve
 * (joint_intrinsics.rs:24-353, dense camera system) where it applies, else
 * the pose / structure solve. options->refine_intrinsics holds the bits
 * VBA_REFINE_*; camera_out receives the refined camera. */
int32_t oracle_optimize_intrinsics(const vba_problem* problem, const vba_options* options, double* poses_out,
                                   double* landmarks_out, vba_camera* camera_out, vba_iteration_stats* stats,
                                   uint32_t stats_capacity, vba_result* result, int32_t threads);
int32_t oracle_optimize_weighted(const vba_problem* problem, const vba_options* options,
                                 const double* weights, uint64_t num_weights, double* poses_out,
                                 double* landmarks_out, vba_iteration_stats* stats,
                                 uint32_t stats_capacity, vba_result* result, int32_t threads);

/* GncConfig::default (gnc.rs:103-115). */
void oracle_default_gnc_options(vba_gnc_options* gnc);

/* BundleAdjustment::optimize_gnc (lm.rs:28-142) with GncState and
 * estimate_scale_mad (gnc.rs:118-286). Same contract as vba_optimize_gnc. */
int32_t oracle_optimize_gnc(const vba_problem* problem, const vba_options* options,
                            const vba_gnc_options* gnc, double* poses_out, double* landmarks_out,
                            double* weights_out, vba_gnc_result* result, int32_t threads);

/* GncState::new(gnc, max_residual_sq), annealed anneal_steps times, then
 * weight(s[i]) for each i; returns is_terminal() and the final mu. */
int32_t oracle_gnc_state_weights(const vba_gnc_options* gnc, double max_residual_sq, uint32_t anneal_steps,
                                 const double* s, uint32_t n, double* weights, double* mu);

/* estimate_scale_mad (gnc.rs:252-273): 1 and *out, or 0 for None. */
int32_t oracle_estimate_scale_mad(const double* squared, uint32_t n, double k, double* out);

/* Camera::project (crates/core/src/types/camera.rs:351-433). Returns 1 and
 * writes uv, or 0 for None. */
int32_t oracle_camera_project(const vba_camera* camera, const double xc[3], double uv[2]);

/* mono_project_with_jacobian with MonoProjection::for_camera
 * (bundle/jacobians.rs:29-113). J is 2x3 row-major. Returns 1 or 0 (None). */
int32_t oracle_mono_project(const vba_camera* camera, const double xc[3], double uv[2],
                            double j[6]);

/* RobustKernel::{cost, weight} (pose_graph.rs:102-139). */
void oracle_robust(uint32_t kernel, double scale, double s, double* cost, double* weight);

/* Per-observation linearization at the given state, in input order: mono
 * observations first, then rig observations (normal_equations.rs:418-484,
 * 617-651; sqrt_qr.rs:960-1001). For each: residual r[2], J_pose 2x6
 * row-major [12], J_landmark 2x3 [6], IRLS weight w and valid (0 when the
 * observation is skipped). Returns VBA_OK or a validation error, and
 * VBA_ERR_INVALID_ARGUMENT for a problem with stereo observations (3 or 4
 * rows each). */
int32_t oracle_linearize(const vba_problem* problem, uint32_t kernel, double scale,
                         double* r, double* j_pose, double* j_landmark, double* w,
                         uint8_t* valid);

/* robust_cost_weighted (problem.rs:497-711) and
 * nonprojectable_observation_count (problem.rs:252-303). */
int32_t oracle_cost(const vba_problem* problem, uint32_t kernel, double scale, double* cost,
                    uint64_t* nonprojectable);

/* SE3::exp / SE3::log / Pose::camera_center_world
 * (crates/core/src/geometry/se3.rs:57-82, pose.rs:31-33). Poses are
 * (qw qx qy qz tx ty tz). */
void oracle_se3_exp(const double xi[6], double pose[7]);
void oracle_se3_log(const double pose[7], double xi[6]);
void oracle_se3_compose(const double a[7], const double b[7], double out[7]);
void oracle_camera_center(const double pose[7], double center[3]);

/* A normal-equation system in the pose-block form of `NormalEquationsBa`
 * (normal_equations.rs:10-18, 78-86): P diagonal 6x6 blocks (row-major),
 * the pose gradient b_p, and per landmark its 3x3 h_ll, b_l and the list of
 * (pose, 6x3 cross block) entries cross_ptr[l]..cross_ptr[l+1]. */
typedef struct oracle_normal_system {
    uint32_t num_poses;
    const double* diag;        /* 36 per pose */
    const double* b_p;         /* 6 per pose */
    uint32_t num_landmarks;
    const double* h_ll;        /* 9 per landmark */
    const double* b_l;         /* 3 per landmark */
    const uint32_t* cross_ptr; /* num_landmarks + 1 */
    const uint32_t* cross_pose;
    const double* cross;       /* 18 per entry, 6x3 row-major */
} oracle_normal_system;

/* Dense Schur reduction of solve_step (normal_equations.rs:1403-1471):
 * S (6P x 6P row-major) and the reduced right-hand side. */
void oracle_ns_reduce_dense(const oracle_normal_system* ns, double lambda, double* s,
                            double* rhs);
/* solve_step with LinearSolver::Dense (normal_equations.rs:1314-1592).
 * Returns VBA_OK or VBA_ERR_SINGULAR_SYSTEM. */
int32_t oracle_ns_solve_dense(const oracle_normal_system* ns, double lambda, double* delta_p,
                              double* delta_l);
/* ImplicitSchurOperator (implicit_schur.rs:133-297, 390-473). Returns 0 or a
 * negative ImplicitSchurError code (see ba_oracle.cpp). */
int32_t oracle_ns_operator(const oracle_normal_system* ns, double lambda, const double* x,
                           double* y, const double* r, double* z, double* rhs);
/* solve_matrix_free_step: PCG + complete_delta (matrix_free.rs:1106-1240).
 * accept_at_cap != 0 takes the iterate at the iteration cap (the GPU
 * backend's rule, ba-gpu/src/solver.rs:497-508) instead of MaxIterations. */
int32_t oracle_ns_solve_pcg(const oracle_normal_system* ns, double lambda,
                            uint32_t max_iterations, double rtol, double atol, int32_t accept_at_cap,
                            double* delta_p, double* delta_l, uint32_t* iterations,
                            double* residual_norm, double* target);
/* solve_normal_equations (pose_graph.rs:2259-2270): Cholesky, LU fallback.
 * A is n x n row-major. Returns 1 on success. */
int32_t oracle_solve_normal_equations(uint32_t n, const double* a, const double* b, double* x);
/* nalgebra's dense Cholesky (row-major in and out; the lower triangle of l is
 * the factor, the rest is what the factor leaves there); 1 on success. */
int32_t oracle_cholesky(uint32_t n, const double* a, double* l, int32_t threads);

/* ReducedSolveAccelerator::eliminate's blocks as the CPU rig solver forms
 * them (eliminate_by_output, rig_ba_solver.rs:2551-2660): blocks[36 E] the
 * final (row t, column s) blocks, column-major; run_blocks and coarse_sums
 * as vba_reduced_eliminate returns them (either may be NULL). */
int32_t oracle_reduced_eliminate(const vba_point_elimination* elimination, double* blocks, double* run_blocks,
                                 double* coarse_sums);

#ifdef __cplusplus
}
#endif

#endif
