//! A `VISLOC_SFM_BA_DUMP` problem (pipelines/slam/src/ba_problem_io.rs) read,
//! windowed and solved through `include/vba_io.h`, without a
//! `BundleAdjustment` in between: the library counterpart of
//! crates/ba-gpu/examples/ba_gpu_replay.rs (examples/vba_replay.rs is the
//! example itself).
use std::ffi::CString;
use std::path::Path;
use std::time::Instant;

use crate::{c_string, ffi, VbaAccelerator};

/// What a replay reports (ba_gpu_replay.rs:19-29).
#[derive(Debug, Clone)]
pub struct ReplayReport {
    pub poses: usize,
    pub fixed_poses: usize,
    pub landmarks: usize,
    pub observations: usize,
    pub seconds: f64,
    pub result: ffi::vba_result,
    pub iterations: Vec<ffi::vba_iteration_stats>,
}

impl ReplayReport {
    pub fn accepted(&self) -> usize {
        self.iterations.iter().filter(|s| s.step_accepted != 0).count()
    }
}

struct Model(*mut ffi::vba_model);

impl Drop for Model {
    fn drop(&mut self) {
        unsafe { ffi::vba_model_free(self.0) };
    }
}

fn io_error() -> String {
    unsafe { c_string(ffi::vba_io_last_error()) }
}

fn c_path(p: &Path) -> Result<CString, String> {
    CString::new(p.as_os_str().as_encoded_bytes()).map_err(|e| e.to_string())
}

/// Read `dump`, keep the local window of its last `local` poses
/// (ba_gpu_replay.rs:43-66) when given, solve it with `options` and, when
/// `out` is given, write the solved model there as COLMAP binary.
pub fn replay(acc: &VbaAccelerator, dump: &Path, local: Option<u32>, options: &ffi::vba_options,
              out: Option<&Path>) -> Result<ReplayReport, String> {
    let path = c_path(dump)?;
    let mut raw = std::ptr::null_mut();
    if unsafe { ffi::vba_model_read_ba_problem(path.as_ptr(), &mut raw) } != ffi::VBA_OK {
        return Err(io_error());
    }
    let model = Model(raw);
    if let Some(n) = local {
        if unsafe { ffi::vba_model_keep_local_window(model.0, n) } != ffi::VBA_OK {
            return Err(io_error());
        }
    }
    let mut view = std::mem::MaybeUninit::<ffi::vba_problem>::zeroed();
    let mut weights = std::ptr::null();
    if unsafe { ffi::vba_model_problem(model.0, view.as_mut_ptr(), &mut weights) } != ffi::VBA_OK {
        return Err(io_error());
    }
    let view = unsafe { view.assume_init() };
    let np = view.num_poses as usize;
    let fixed = if view.pose_fixed.is_null() {
        0
    } else {
        unsafe { std::slice::from_raw_parts(view.pose_fixed, np) }.iter().filter(|&&f| f != 0).count()
    };
    let mut poses = vec![0.0f64; 7 * np];
    let mut landmarks = vec![0.0f64; 3 * view.num_landmarks as usize];
    let mut stats = vec![ffi::vba_iteration_stats::default(); (options.max_iterations as usize).max(1)];
    let mut res = ffi::vba_result::default();
    let t = Instant::now();
    let st = {
        let g = acc.lock().ok_or("context lock poisoned")?;
        unsafe {
            if weights.is_null() {
                ffi::vba_optimize(g.0, &view, options, poses.as_mut_ptr(), landmarks.as_mut_ptr(),
                                  stats.as_mut_ptr(), stats.len() as u32, &mut res)
            } else {
                // a dump holds monocular observations only, one weight each
                ffi::vba_optimize_weighted(g.0, &view, options, weights, view.num_observations, poses.as_mut_ptr(),
                                           landmarks.as_mut_ptr(), stats.as_mut_ptr(), stats.len() as u32,
                                           &mut res)
            }
        }
    };
    let seconds = t.elapsed().as_secs_f64();
    if st != ffi::VBA_OK {
        return Err(format!("solve status {st}: {}", acc.last_error()));
    }
    stats.truncate((res.num_iterations as usize).min(stats.len()));
    if let Some(dir) = out {
        if unsafe { ffi::vba_model_set_state(model.0, poses.as_ptr(), landmarks.as_ptr()) } != ffi::VBA_OK {
            return Err(io_error());
        }
        let dir = c_path(dir)?;
        if unsafe { ffi::vba_model_write_colmap(model.0, dir.as_ptr(), 1) } != ffi::VBA_OK {
            return Err(io_error());
        }
    }
    Ok(ReplayReport {
        poses: np,
        fixed_poses: fixed,
        landmarks: view.num_landmarks as usize,
        observations: view.num_observations as usize,
        seconds,
        result: res,
        iterations: stats,
    })
}
