// kernels.hpp — device data layout and kernel launchers (kernels.hip).
//
// Per-edge data. For an observation of landmark X from pose T with IRLS
// weight w the reference builds J_pose = J_pi [R | -R [X]x] and
// J_lm = J_pi R (pipelines/slam/src/bundle/normal_equations.rs:440-452; rig
// form sqrt_qr.rs:989-999), so J_pose = J_lm [I | -[X]x]. Every block the
// Schur path needs then follows from the symmetric 3x3 G = w J_lm^T J_lm,
// the 3-vector h = w J_lm^T r and X:
//   U_p  += [[G, -G[X]x], [[X]x G, -[X]x G [X]x]]     g_p += [h; X x h]
//   V_l  += G                                          g_l += h
//   W^T x_p = G (rho + omega x X)                      W u = [G u; X x (G u)]
// An edge is a (pose, landmark) pair; observations sharing it (a rig's
// sensors) are summed into one G, h, which is the grouping the reference's
// block-Jacobi preconditioner applies (implicit_schur.rs:241-258).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/vba.h"
#include "geom.hpp"
#include "problem.hpp"

namespace vba {

constexpr int kBlock = 256;        // threads per tile work-group
constexpr int kSmallBlock = 1024;  // single work-group kernels
constexpr int kPosesPerWave = 10;  // PCG step kernel: 6 lanes per pose

// Scalars exchanged with the host once per LM iteration.
struct StepScalars {
    double cost_sum, cost_comp;
    double max_pose_step, max_landmark_step;
    uint64_t nonprojectable;
    int32_t precond_failed;
    int32_t pad;
};

// PCG state kept on the device (implicit_schur.rs:522-736 control flow).
// rho holds r.z of the last two iterations by iteration parity, so the
// kernels of one iteration read the previous value while the first
// work-group of the operator kernel records the new one.
struct PcgState {
    double rhs_norm, target, rnorm, true_norm, lambda, beta;
    double rho[2];
    int32_t iter, status, max_iter, stop;
};
enum PcgStatus : int32_t {
    PCG_RUNNING = 0,
    PCG_CONVERGED = 1,   // recursive residual <= target
    PCG_MAXIT = 2,       // iteration budget reached above target
    PCG_NONPOSITIVE = -1,
    PCG_NONFINITE = -2
};

// Per-window LM state of a batch (device side of the LM control,
// lm.rs:804-1008, one thread per window).
struct WindowState {
    double lambda, current_cost, cost_after, max_pose_step, max_landmark_step, initial_cost;
    uint64_t current_nonproj, nonproj_after;
    int32_t active, relin, accept, converged, iterations, solve_status;
};

// All device pointers of one problem.
struct DevProblem {
    int32_t num_pose_slots, num_lm_slots, num_edges, num_obs;
    int32_t num_cam_tiles, num_lm_tiles, any_rig, kernel_kind;
    int32_t num_cus, pad0;  // compute units of the device (persistent grids)
    double kernel_scale;
    // state
    double* poses;        // 7 per slot
    double* lms;          // 3 per slot
    double* poses_trial;
    double* lms_trial;
    const DevSensor* sensors;
    const int32_t* pose_var;
    const int32_t* lm_var;
    // structure
    const double* obs_uv;
    const double* obs_w;  // external weight per observation, c-order (null: all 1)
    double* obs_s;        // k_cost writes each observation's squared residual here (NaN: not evaluable)
    const uint16_t* obs_sensor;
    const uint8_t* obs_kind;    // ObsKind per observation, c-order (null: all pixel residuals)
    const double* obs_extra;    // 2 per observation, c-order: rectified (u_r, baseline), general (u_r, v_r)
    const int32_t* edge_lm;
    const int32_t* edge_obs;
    const int32_t* perm_cl;   // c-order edge -> storage position
    const int32_t* s_pose;    // pose at each storage position
    const int32_t* lm_store;  // storage position of each l-order edge
    const int32_t* lm_edge_ptr;
    const CamTile* cam_tiles;
    const int32_t* cam_tile_ptr;
    const LmTile* lm_tiles;
    // linearization: per edge (storage order) the upper triangle of G and h
    double* Gl64;    // 6 per edge
    float* Gl32;     // FP32 copy for the mixed-precision operator
    double* hl64;    // 3 per edge
    double* tile27;  // per camera tile partials (27)
    double* U;       // 21 per pose slot (upper 6x6, row-major)
    double* g;       // 6 per pose slot
    double* V;       // 6 per landmark slot (upper 3x3)
    double* gl;      // 3 per landmark slot
    double* Vinv64;  // 6 per landmark slot
    double* Vm;      // 6 per landmark: V summed from the FP32 G (mixed precision)
    double* VinvM;   // its damped inverse, the one the mixed operator uses
    double* Minv;    // 36 per pose slot (block-Jacobi inverse)
    // Schur operator structure. The edges of a landmark tile, sorted by pose
    // (tile_perm: tile-local l-order index at each sorted position; s_lm:
    // tile-local landmark), form camera runs; a run leaves one partial in
    // its slot (run_slot at the run's last sorted position), the slots of
    // pose p being pose_run_ptr[p] .. pose_run_ptr[p + 1] - 1 in ascending
    // storage order, and a pose sums its slots in that order.
    const uint16_t* tile_perm;
    const uint8_t* s_lm;
    const int32_t* big_tiles;  // tiles of one landmark with more than kTileEdges edges
    int32_t num_big_tiles, pad1;
    const int32_t* pose_run_ptr;  // num_pose_slots + 1
    const int32_t* run_slot;      // per storage position: slot of the run ending there, else -1
    const int32_t* g32_pos;       // per storage position: its index in the lane-major FP32 G
    // the landmark tiles' run tables (see Model::t_lm): local landmark by
    // position, run-start mask, each run's pose and slot
    const uint8_t* t_lm;
    const uint64_t* t_head;
    const int32_t* t_rpose;
    const int32_t* t_rslot;
    double* part64;  // 6 per run slot
    double* part21;  // 21 per run slot (block-Jacobi blocks)
    // PCG vectors, 6 per pose slot
    double *b, *x, *r, *z, *d, *q;
    double* bpart;  // per work-group partials of the pose kernels (2 each)
    double* cpart;  // per work-group (r.z, r.r) of the step kernel
    int32_t num_bpart, num_cpart;
    PcgState* pcg;
    // step outputs
    double* delta_l;       // 3 per landmark slot
    double* tile_scalar;   // per tile partial scalars (cost sum/comp, count, max)
    StepScalars* scalars;
    // Batches of independent windows (null for a single problem): the window
    // of every slot and tile, each window's lambda and its control flags.
    int32_t num_windows;
    const int32_t* pose_win;
    const int32_t* lm_win;
    const int32_t* cam_tile_win;
    const int32_t* lm_tile_win;
    const WindowState* wstate;  // lambda, active (iterating), relin (state changed)
    // joint pose / landmark / camera solve (joint_k > 0 camera parameters)
    int32_t joint_k, joint_shared;
    DevSensor* sensors_trial;  // the trial camera (null outside the joint solve)
    double* joint_E;           // per storage position: sum of w J_l^T J_k (3 x k)
    double* joint_tile;        // per camera tile: J_p^T J_k (6 x k), J_k^T J_k (upper), J_k^T r
    double* joint_hpk;         // per pose slot: 6 x k
    double* joint_kk;          // J_k^T J_k (upper) and J_k^T r over all observations
    double* joint_Wk;          // per landmark slot: 3 x k
    double* joint_part;        // per work-group partials of the camera-block Schur terms
    double* joint_rhs;         // k
    double* joint_delta;       // k
    double* joint_gl;          // per landmark slot: g_l + W_lk delta_k
};

struct WindowLayout {
    int32_t tile_begin, tile_end;  // camera tiles
    int32_t pose_begin, pose_end;  // pose slots
    int32_t lm_begin, lm_end;      // landmark slots
    int32_t var_begin, var_end;    // variable pose indices
    int64_t s_offset;              // dense S offset (doubles)
    int64_t pad;
};

// ---- launchers (all asynchronous on `stream`) ----
// kernels.hip
void launch_linearize(const DevProblem& P, bool store32, hipStream_t stream);
void launch_camera_finalize(const DevProblem& P, hipStream_t stream);
void launch_landmark_reduce(const DevProblem& P, hipStream_t stream);
void launch_landmark_prepare(const DevProblem& P, double lambda, hipStream_t stream);
void launch_pose_update(const DevProblem& P, hipStream_t stream);
// Cost and non-projectable count of (poses, lms) or of the trial state.
void launch_cost(const DevProblem& P, bool trial, hipStream_t stream);
void launch_scalars_finalize(const DevProblem& P, hipStream_t stream);

// schur.hip: the Schur operator over landmark tiles and the PCG vector work.
// Block-Jacobi inverse blocks of U + lambda I - W V^-1 W^T into Minv.
void launch_precondition(const DevProblem& P, double lambda, hipStream_t stream);
// b = -g_p + sum_e W_e V^-1 g_l.
void launch_rhs(const DevProblem& P, hipStream_t stream);
// Operator partials of d = v1 + beta v2 (v2 may be null) into the run
// partials part64; fp32 reads the FP32 copies of G and V^-1 (arithmetic
// stays FP64). With pcg_iter >= 0 the launch is PCG
// iteration pcg_iter: each work-group first derives the iteration's
// scalars (beta, stop) from the step kernel's partials. resident_waves > 0
// runs the persistent form (that many waves loop over the tiles) with the
// same results.
// Events bound to one kernel dispatch (hipExtLaunchKernel): they take the
// kernel's own start and end times, so their interval leaves out the time
// the stream spends getting to the dispatch. Null events time nothing.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
void launch_schur_apply(const DevProblem& P, bool fp32, const double* v1, const double* v2, int32_t pcg_iter,
                        int32_t fresh, hipStream_t stream, int32_t resident_waves = 0,
                        const LaunchEvents& ev = {});
// PCG: r = b, z = M^-1 r, x = d = 0 and the state (implicit_schur.rs:522-560).
void launch_pcg_init(const DevProblem& P, double rtol, double atol, int32_t max_iter, double lambda,
                     hipStream_t stream);
// d = z + beta d, q = S d from the run partials, and d.q partials.
void launch_pcg_pose(const DevProblem& P, hipStream_t stream, const LaunchEvents& ev = {});
// alpha, x += alpha d, r -= alpha q, z = M^-1 r and (r.z, r.r) partials.
void launch_pcg_step(const DevProblem& P, int32_t pcg_iter, hipStream_t stream, const LaunchEvents& ev = {});
// The scalars of iteration pcg_iter (status, residual norm) without its work.
void launch_pcg_status(const DevProblem& P, int32_t pcg_iter, int32_t fresh, hipStream_t stream);
// True residual b - S x in FP64 (state: true_norm).
void launch_pcg_true_residual(const DevProblem& P, hipStream_t stream);
// Continue the recursion from the residual of the last true-residual check:
// r = b - S x, z = M^-1 r, d = 0 (FP64 refinement of the mixed solve).
void launch_pcg_restart(const DevProblem& P, hipStream_t stream);
// delta_l = -V^-1 (g_l + sum_e W_e^T x_p), trial landmarks, step maxima.
void launch_backsub(const DevProblem& P, hipStream_t stream);

// r_t.r_t (and with restart r_t.z) from the pose kernels' partials into the
// PCG state; with restart the recursion continues from r_t.
void launch_pcg_true_finish(const DevProblem& P, int32_t restart, hipStream_t stream);

// Dense reduced system (LinearSolver::Dense).
struct DensePairs {
    int32_t num_blocks;
    const int32_t* block_pq;   // 2 per block: pose slots (var(p) <= var(q))
    const int32_t* block_ptr;  // num_blocks + 1 term ranges
    const int32_t* terms;      // 3 per term: landmark slot, l-order edge of p, of q
    const int32_t* var_slot;   // pose slot of each variable pose
};

// Explicit block-sparse reduced camera matrix (explicit_schur.hip):
// S = U + lambda I - sum_l W V^-1 W^T as 6x6 blocks between variable poses
// that share a landmark, plus every diagonal block. Each upper block (p, q)
// sums its term list (the DensePairs lists of the dense solver) in list
// order. S is stored in full block rows (both S_pq and S_qp = S_pq^T),
// columns ascending, each block row-major.
struct ExplicitDev {
    int32_t num_blocks, num_full;
    const int32_t* block_pos;  // 2 per upper block: full positions of S_pq and S_qp (-1 on the diagonal)
    const int32_t* row_ptr;    // num_pose_slots + 1
    const int32_t* col;        // pose slot of each full position
    const int32_t* diag_pos;   // full position of S_pp per pose slot (-1: fixed pose)
    double* S64;               // 36 per full position
    float* S32;                // FP32 copy for the mixed-precision matvec (null in FP64)
};
// S64 (and S32) from the term lists, U and lambda, then the block-Jacobi
// inverses of the diagonal blocks into Minv (implicit_schur.rs:241-287).
void launch_explicit_assemble(const DevProblem& P, const DensePairs& D, const ExplicitDev& X, double lambda,
                              hipStream_t stream);
// PCG iteration pcg_iter on the explicit matrix: d_new = z + beta d_old,
// q = S d_new and d.q partials; the scalars come from the step kernel's
// partials as for the implicit operator. fp32 reads S32.
void launch_pcg_explicit(const DevProblem& P, const ExplicitDev& X, bool fp32, const double* d_old, double* d_new,
                         int32_t pcg_iter, int32_t fresh, hipStream_t stream);
// True residual b - S64 x and its norm (launch_pcg_true_finish); with
// restart also r = r_t, z = M^-1 r_t and d = 0 (d is P.d).
void launch_explicit_true_residual(const DevProblem& P, const ExplicitDev& X, int32_t restart, hipStream_t stream);

// Joint camera block (joint_k = 4, 6 or 8): the camera terms after
// linearization; the border rows of S and the camera right-hand side once
// the pose blocks are assembled (damping, Schur terms, the shared-focal
// fold); after the solve, the landmark gradient shifted by the camera step
// (back-substitution reads joint_gl) and the trial camera.
void launch_joint_linearize(const DevProblem& P, hipStream_t stream);
void launch_joint_border(const DevProblem& P, double* S, int32_t n, double lambda, hipStream_t stream);
void launch_joint_step(const DevProblem& P, hipStream_t stream);
// Doubles of joint_part a problem needs.
size_t joint_part_size(const DevProblem& P);

void launch_dense_assemble(const DevProblem& P, const DensePairs& D, double* S, int32_t n,
                           double lambda, hipStream_t stream);
// Solve S x = b (Cholesky, then an LU fallback with allow_lu, else singular)
// for the variable poses; writes x and sets status[0] (0 ok, 1 singular);
// status holds two words (the second records a failed pivot). S is
// overwritten; `work` (n * n + n) receives a copy of S for the LU fallback.
// Systems of more than kDenseOneGroupMax unknowns are factored across the
// device (two launches per 64 columns), smaller ones by one work-group.
// Returns the error of the copy or of a launch.
// The same solve for a plain system: S row-major n x n (overwritten), b and x
// n long, always with the LU fallback.
constexpr int32_t kDenseOneGroupMax = 256;
// The LU fallback runs in one work-group, reading and writing the trailing
// matrix once per pivot (n^3 / 3 x 16 bytes through one CU); beyond this size
// it would run for minutes, so a failed Cholesky is reported singular.
constexpr int32_t kLuFallbackMax = 3072;
hipError_t launch_dense_solve_plain(double* S, double* work, const double* b, double* x, int32_t n,
                                    int32_t* status, hipStream_t stream);
hipError_t launch_dense_solve(const DevProblem& P, const DensePairs& D, double* S, double* work,
                              int32_t n, int32_t* status, hipStream_t stream, bool allow_lu = true);

// ---- batches of windows (one work-group per window where noted) ----
struct BatchDev {
    int32_t num_windows;
    const WindowLayout* layout;
    WindowState* state;
    vba_iteration_stats* stats;  // num_windows x stats_capacity
    int32_t stats_capacity;
    double* S;      // dense systems, WindowLayout::s_offset
    double* work;   // per window: n*n + n doubles at s_offset + s_offset/... (see solver)
    const int64_t* work_offset;
};
// Assemble every active window's dense S (lambda per window).
void launch_batch_dense_assemble(const DevProblem& P, const DensePairs& D, const BatchDev& B, hipStream_t stream);
// Zero the dense systems of active windows.
void launch_batch_dense_zero(const DevProblem& P, const BatchDev& B, hipStream_t stream);
// One work-group per window: Cholesky (LU fallback) solve into P.x.
void launch_batch_dense_solve(const DevProblem& P, const DensePairs& D, const BatchDev& B, hipStream_t stream);
// Per-window cost / count / step maxima of the trial (or current) state.
void launch_batch_scalars(const DevProblem& P, const BatchDev& B, hipStream_t stream);
// Initial LM state from the current-state cost.
void launch_batch_init(const DevProblem& P, const BatchDev& B, const vba_options& o, hipStream_t stream);
// Accept / reject, lambda update, convergence, stats (one thread per window).
void launch_batch_control(const DevProblem& P, const BatchDev& B, const vba_options& o, int32_t iteration,
                          hipStream_t stream);
// Copy the trial state of accepted windows into the current state.
void launch_batch_commit(const DevProblem& P, const BatchDev& B, hipStream_t stream);

}  // namespace vba
