//! `visloc-ba-mi355x`: visloc-rs's bundle-adjustment accelerator hooks on
//! libvba, the HIP (gfx950) Levenberg–Marquardt Schur-complement solver
//! declared in `include/vba.h`.
//!
//! - [`VbaAccelerator`] implements `BaAccelerator`
//!   (pipelines/slam/src/ba_accel.rs:22-32) in the place of the wgpu
//!   `GpuBundleAdjuster` (crates/ba-gpu/src/solver.rs:729-742): `None` hands
//!   the call back to the CPU solver.
//! - It implements `ReducedSolveAccelerator` (ba_accel.rs:134-157) in
//!   `reduced.rs`: the COLMAP-port rig solver's reduced-camera PCG and its
//!   point elimination.
//! - [`VbaAccelerator::optimize_weighted`] and
//!   [`VbaAccelerator::optimize_gnc`] run the weighted and GNC solves
//!   (optimize.rs:43-53, lm.rs:28-142), which the reference keeps on the CPU.
//! - [`replay::replay`] reads a `VISLOC_SFM_BA_DUMP` problem, optionally keeps
//!   a local window, solves it and writes the model
//!   (crates/ba-gpu/examples/ba_gpu_replay.rs:43-141).
//!
//! Registration is the wgpu backend's (examples/colmap.rs:443):
//!
//! ```ignore
//! if let Some(acc) = visloc_ba_mi355x::VbaAccelerator::new(0) {
//!     visloc_slam::set_ba_accelerator(Box::new(acc));
//! }
//! if let Some(acc) = visloc_ba_mi355x::VbaAccelerator::new(0) {
//!     visloc_slam::set_reduced_solve_accelerator(Box::new(acc));
//! }
//! ```

pub mod ffi;
mod problem;
mod reduced;
pub mod replay;

use std::ffi::CStr;
use std::sync::{Mutex, MutexGuard};

use visloc_slam::gnc::{GncConfig, GncKernel};
use visloc_slam::{
    BaAccelerator, BaConfig, BaError, BaGncResult, BaIterationStats, BaResult, BaScope, BundleAdjustment,
    LinearSolver, RobustKernel,
};

pub use problem::ProblemView;

/// One libvba context (one HIP device and stream). Calls through one
/// accelerator are serialized, as the library serializes calls on one
/// context.
pub struct VbaAccelerator {
    ctx: Mutex<Context>,
}

pub(crate) struct Context(*mut ffi::vba_context);

// The context is used under the mutex only; libvba contexts may move
// between threads.
unsafe impl Send for Context {}

impl VbaAccelerator {
    /// A context on HIP device `device`; `None` when there is no such device
    /// or the library was built for another ABI.
    pub fn new(device: i32) -> Option<Self> {
        if unsafe { ffi::vba_abi_version() } != ffi::VBA_ABI_VERSION {
            return None;
        }
        let ctx = unsafe { ffi::vba_create(device) };
        (!ctx.is_null()).then(|| Self { ctx: Mutex::new(Context(ctx)) })
    }

    pub(crate) fn lock(&self) -> Option<MutexGuard<'_, Context>> {
        self.ctx.lock().ok()
    }

    /// The library's message for the last failed call on this context.
    pub fn last_error(&self) -> String {
        match self.lock() {
            Some(g) => unsafe { c_string(ffi::vba_last_error(g.0)) },
            None => String::from("context lock poisoned"),
        }
    }

    /// `optimize_with_observation_weights` (optimize.rs:43-53) on the device;
    /// `None`: run the CPU solver. Weights are in the flattened observation
    /// order (monocular, rectified stereo, general stereo, rig).
    pub fn optimize_weighted(&self, ba: &mut BundleAdjustment, config: &BaConfig, weights: &[f64])
                             -> Option<Result<BaResult, BaError>> {
        let view = ProblemView::new(ba)?;
        let options = options_from(config);
        let mut poses = vec![0.0f64; 7 * view.num_poses()];
        let mut landmarks = vec![0.0f64; 3 * view.num_landmarks()];
        let mut stats = vec![ffi::vba_iteration_stats::default(); config.max_iterations.max(1)];
        let mut res = ffi::vba_result::default();
        let st = {
            let g = self.lock()?;
            unsafe {
                ffi::vba_optimize_weighted(g.0, view.as_ffi(), &options, weights.as_ptr(), weights.len() as u64,
                                           poses.as_mut_ptr(), landmarks.as_mut_ptr(), stats.as_mut_ptr(),
                                           stats.len() as u32, &mut res)
            }
        };
        match st {
            ffi::VBA_OK => {
                view.write_back(ba, &poses, &landmarks);
                Some(Ok(result_from(&res, &stats)))
            }
            ffi::VBA_ERR_OBSERVATION_WEIGHT_COUNT => Some(Err(BaError::ObservationWeightCount {
                expected: view.num_weighted_observations(),
                actual: weights.len(),
            })),
            ffi::VBA_ERR_INVALID_OBSERVATION_WEIGHT => {
                Some(Err(BaError::InvalidObservationWeight(res.error_id as usize)))
            }
            ffi::VBA_ERR_WEIGHTS_WITH_INTRINSICS => Some(Err(BaError::ObservationWeightsWithIntrinsicsRefinement)),
            _ => error_or_decline(st, &res).map(Err),
        }
    }

    /// `BundleAdjustment::optimize_gnc` (lm.rs:28-142) on the device; `None`:
    /// run the CPU solver.
    pub fn optimize_gnc(&self, ba: &mut BundleAdjustment, config: &BaConfig, gnc: &GncConfig)
                        -> Option<Result<BaGncResult, BaError>> {
        let view = ProblemView::new(ba)?;
        let options = options_from(config);
        let mut g_opts = ffi::vba_gnc_options::default();
        unsafe { ffi::vba_default_gnc_options(&mut g_opts) };
        g_opts.kernel = match gnc.kernel {
            GncKernel::GemanMcClure => ffi::VBA_GNC_GEMAN_MCCLURE,
            GncKernel::TruncatedLeastSquares => ffi::VBA_GNC_TRUNCATED_LEAST_SQUARES,
        };
        g_opts.c = gnc.c;
        g_opts.anneal_factor = gnc.anneal_factor;
        g_opts.max_outer = u32::try_from(gnc.max_outer).ok()?;
        g_opts.inner_iterations = u32::try_from(gnc.inner_iterations).ok()?;
        g_opts.has_auto_scale = gnc.auto_scale.is_some() as u32;
        g_opts.auto_scale = gnc.auto_scale.unwrap_or(0.0);
        g_opts.auto_scale_readapt = gnc.auto_scale_readapt as u32;
        let mut poses = vec![0.0f64; 7 * view.num_poses()];
        let mut landmarks = vec![0.0f64; 3 * view.num_landmarks()];
        let mut weights = vec![0.0f64; view.num_all_observations()];
        let mut res = ffi::vba_gnc_result::default();
        let st = {
            let g = self.lock()?;
            unsafe {
                ffi::vba_optimize_gnc(g.0, view.as_ffi(), &options, &g_opts, poses.as_mut_ptr(),
                                      landmarks.as_mut_ptr(), weights.as_mut_ptr(), &mut res)
            }
        };
        match st {
            ffi::VBA_OK => {}
            // vba_gnc_result carries no offending id: the CPU solver reports it
            ffi::VBA_ERR_MISSING_POSE | ffi::VBA_ERR_MISSING_LANDMARK => return None,
            _ => return error_or_decline(st, &ffi::vba_result::default()).map(Err),
        }
        view.write_back(ba, &poses, &landmarks);
        weights.truncate(res.observation_count as usize);
        Some(Ok(BaGncResult {
            initial_cost: res.initial_cost,
            final_cost: res.final_cost,
            inlier_cost: res.inlier_cost,
            inlier_scale: res.inlier_scale,
            observation_count: res.observation_count as usize,
            outer_iterations: res.outer_iterations as usize,
            converged: res.converged != 0,
            observation_weights: weights,
        }))
    }
}

impl Drop for VbaAccelerator {
    fn drop(&mut self) {
        if let Ok(g) = self.ctx.get_mut() {
            unsafe { ffi::vba_destroy(g.0) };
        }
    }
}

impl BaAccelerator for VbaAccelerator {
    fn optimize(&self, ba: &mut BundleAdjustment, config: &BaConfig, _scope: BaScope)
                -> Option<Result<BaResult, BaError>> {
        let view = ProblemView::new(ba)?;
        let options = options_from(config);
        let mut poses = vec![0.0f64; 7 * view.num_poses()];
        let mut landmarks = vec![0.0f64; 3 * view.num_landmarks()];
        let mut stats = vec![ffi::vba_iteration_stats::default(); config.max_iterations.max(1)];
        let mut res = ffi::vba_result::default();
        let mut camera_out = view.as_ffi().camera;
        let st = {
            let g = self.lock()?;
            if config.refine_intrinsics {
                // the joint pose / landmark / camera solve, which also returns the camera
                unsafe {
                    ffi::vba_optimize_intrinsics(g.0, view.as_ffi(), &options, poses.as_mut_ptr(),
                                                 landmarks.as_mut_ptr(), &mut camera_out, stats.as_mut_ptr(),
                                                 stats.len() as u32, &mut res)
                }
            } else {
                unsafe {
                    ffi::vba_optimize(g.0, view.as_ffi(), &options, poses.as_mut_ptr(), landmarks.as_mut_ptr(),
                                      stats.as_mut_ptr(), stats.len() as u32, &mut res)
                }
            }
        };
        if st != ffi::VBA_OK {
            return error_or_decline(st, &res).map(Err);
        }
        if config.refine_intrinsics {
            problem::set_camera(&mut ba.camera, &camera_out);
        }
        // In place: the reference accelerator leaves the accepted state in `ba`
        // (solver.rs:634-637).
        view.write_back(ba, &poses, &landmarks);
        Some(Ok(result_from(&res, &stats)))
    }
}

/// `BaConfig` (bundle/mod.rs:625-752) as `vba_options`, on top of
/// `vba_default_options` (which holds `MatrixFreeBaOptions::default()` for
/// the PCG settings and the CPU matrix-free entry's at-cap rule).
pub fn options_from(config: &BaConfig) -> ffi::vba_options {
    let mut o = ffi::vba_options::default();
    unsafe { ffi::vba_default_options(&mut o) };
    o.max_iterations = u32::try_from(config.max_iterations).unwrap_or(u32::MAX);
    o.has_initial_lambda = config.initial_lambda.is_some() as u32;
    o.initial_lambda = config.initial_lambda.unwrap_or(0.0);
    o.lambda_increase_factor = config.lambda_increase_factor;
    o.lambda_decrease_factor = config.lambda_decrease_factor;
    o.max_lambda = config.max_lambda;
    o.min_lambda = config.min_lambda;
    o.step_tolerance = config.step_tolerance;
    o.cost_tolerance = config.cost_tolerance;
    o.has_relative_cost_tolerance = config.relative_cost_tolerance.is_some() as u32;
    o.relative_cost_tolerance = config.relative_cost_tolerance.unwrap_or(0.0);
    let (kernel, scale) = match config.robust_kernel {
        RobustKernel::None => (ffi::VBA_KERNEL_NONE, 0.0),
        RobustKernel::Huber { delta } => (ffi::VBA_KERNEL_HUBER, delta),
        RobustKernel::Cauchy { c } => (ffi::VBA_KERNEL_CAUCHY, c),
    };
    o.robust_kernel = kernel;
    o.robust_scale = scale;
    // matrix_free_ba selects the matrix-free entry (optimize.rs:70-103);
    // otherwise LinearSolver picks dense Cholesky/LU or the sparse Cholesky
    // (no LU fallback, as block_cholesky.rs has none)
    o.linear_solver = if config.matrix_free_ba {
        ffi::VBA_SOLVER_MATRIX_FREE
    } else {
        match config.linear_solver {
            LinearSolver::Dense => ffi::VBA_SOLVER_DENSE,
            LinearSolver::Sparse => ffi::VBA_SOLVER_SPARSE,
        }
    };
    // the other three flags act only with refine_intrinsics (mod.rs:655-700)
    o.refine_intrinsics = if config.refine_intrinsics {
        ffi::VBA_REFINE_INTRINSICS
            | if config.refine_distortion { ffi::VBA_REFINE_DISTORTION } else { 0 }
            | if config.refine_tangential_distortion { ffi::VBA_REFINE_TANGENTIAL } else { 0 }
            | if config.shared_focal { ffi::VBA_REFINE_SHARED_FOCAL } else { 0 }
    } else {
        0
    };
    o
}

/// `BaResult` (mod.rs:765-771) from the library's result and stats.
pub fn result_from(res: &ffi::vba_result, stats: &[ffi::vba_iteration_stats]) -> BaResult {
    let n = (res.num_iterations as usize).min(stats.len());
    let iterations = stats[..n]
        .iter()
        .map(|s| BaIterationStats {
            iteration: s.iteration as usize,
            cost_before: s.cost_before,
            cost_after: s.cost_after,
            max_pose_step: s.max_pose_step,
            max_landmark_step: s.max_landmark_step,
            lambda: s.lambda,
            step_accepted: s.step_accepted != 0,
        })
        .collect();
    BaResult { initial_cost: res.initial_cost, final_cost: res.final_cost, iterations, converged: res.converged != 0 }
}

/// A status other than `VBA_OK` as the accelerator's answer: `None` (the CPU
/// solver runs) for declines, device trouble and the matrix-free entry's
/// eligibility errors, which the reference answers by falling back to
/// `optimize` (optimize.rs:70-103); `BaError` for the rest (mod.rs:1021-1056).
pub fn error_or_decline(status: i32, res: &ffi::vba_result) -> Option<BaError> {
    match status {
        ffi::VBA_ERR_NO_POSES => Some(BaError::NoPoses),
        ffi::VBA_ERR_NO_LANDMARKS => Some(BaError::NoLandmarks),
        ffi::VBA_ERR_NO_OBSERVATIONS => Some(BaError::NoObservations),
        ffi::VBA_ERR_ALL_POSES_FIXED => Some(BaError::AllPosesFixed),
        ffi::VBA_ERR_MISSING_POSE => Some(BaError::MissingPose(res.error_id)),
        ffi::VBA_ERR_MISSING_LANDMARK => Some(BaError::MissingLandmark(res.error_id)),
        ffi::VBA_ERR_UNSUPPORTED_CAMERA_MODEL => Some(BaError::UnsupportedCameraModel),
        ffi::VBA_ERR_MISSING_STEREO_BASELINE => Some(BaError::MissingStereoBaseline),
        ffi::VBA_ERR_SINGULAR_SYSTEM => Some(BaError::SingularSystem),
        _ => None,
    }
}

pub(crate) unsafe fn c_string(p: *const std::os::raw::c_char) -> String {
    if p.is_null() {
        String::new()
    } else {
        CStr::from_ptr(p).to_string_lossy().into_owned()
    }
}
