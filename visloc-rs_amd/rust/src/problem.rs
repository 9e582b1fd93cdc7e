//! A `BundleAdjustment` (pipelines/slam/src/bundle/mod.rs:529-622) as a
//! `vba_problem`: id arrays in the BTreeMap order the reference's variable
//! layout uses, the state as SoA arrays, and the observation lists in the
//! reference's flattened order (monocular, rectified stereo, general stereo,
//! rig). The view owns every array it points into; the arrays are heap
//! buffers, so the view may move.
use std::collections::HashMap;

use nalgebra::{Point3, Quaternion, UnitQuaternion, Vector3};
use visloc_core::geometry::SE3;
use visloc_core::types::{Camera, CameraModel};
use visloc_slam::BundleAdjustment;

use crate::ffi;

pub struct ProblemView {
    pose_ids: Vec<u64>,
    poses: Vec<f64>,
    pose_fixed: Vec<u8>,
    landmark_ids: Vec<u64>,
    landmarks: Vec<f64>,
    landmark_fixed: Vec<u8>,
    obs_pose: Vec<u64>,
    obs_landmark: Vec<u64>,
    obs_xy: Vec<f64>,
    rig_pose: Vec<u64>,
    rig_landmark: Vec<u64>,
    rig_xy: Vec<f64>,
    rig_sensor: Vec<u32>,
    sensors: Vec<ffi::vba_sensor>,
    stereo_pose: Vec<u64>,
    stereo_landmark: Vec<u64>,
    stereo_obs: Vec<f64>,
    general_pose: Vec<u64>,
    general_landmark: Vec<u64>,
    general_obs: Vec<f64>,
    general_pair: Vec<u32>,
    pairs: Vec<ffi::vba_sensor>,
    stereo: Box<ffi::vba_stereo_terms>,
    problem: ffi::vba_problem,
}

/// The library's camera for a reference camera; `None` for a model the
/// library does not project (the call is then declined).
pub fn camera(c: &Camera) -> Option<ffi::vba_camera> {
    let model = match c.model {
        CameraModel::Pinhole => ffi::VBA_CAMERA_PINHOLE,
        CameraModel::SimplePinhole => ffi::VBA_CAMERA_SIMPLE_PINHOLE,
        CameraModel::SimpleRadial => ffi::VBA_CAMERA_SIMPLE_RADIAL,
        CameraModel::Radial => ffi::VBA_CAMERA_RADIAL,
        CameraModel::OpenCv => ffi::VBA_CAMERA_OPENCV,
        _ => return None,
    };
    if c.params.len() > 12 {
        return None;
    }
    let mut params = [0.0; 12];
    params[..c.params.len()].copy_from_slice(&c.params);
    Some(ffi::vba_camera { model, width: c.width, height: c.height, num_params: c.params.len() as u32, params })
}

/// The refined camera of the joint intrinsics solve written back
/// (`VBA_REFINE_TANGENTIAL` turns the camera into OPENCV).
pub fn set_camera(c: &mut Camera, out: &ffi::vba_camera) {
    if out.model == ffi::VBA_CAMERA_OPENCV {
        c.model = CameraModel::OpenCv;
    }
    c.params = out.params[..(out.num_params as usize).min(12)].to_vec();
}

fn se3_array(t: &SE3) -> [f64; 7] {
    let q = t.rotation.quaternion();
    [q.w, q.i, q.j, q.k, t.translation.x, t.translation.y, t.translation.z]
}

/// Sensors (a lens and an extrinsic) shared by many observations are stored
/// once; the key is the bit pattern of everything the library reads.
struct SensorTable {
    sensors: Vec<ffi::vba_sensor>,
    index: HashMap<Vec<u64>, u32>,
}

impl SensorTable {
    fn new() -> Self {
        Self { sensors: Vec::new(), index: HashMap::new() }
    }

    fn add(&mut self, cam: &Camera, extrinsic: &SE3) -> Option<u32> {
        let sensor = ffi::vba_sensor { camera: camera(cam)?, sensor_from_rig: se3_array(extrinsic) };
        let mut key = vec![sensor.camera.model as u64, sensor.camera.width as u64, sensor.camera.height as u64,
                           sensor.camera.num_params as u64];
        key.extend(sensor.camera.params.iter().map(|v| v.to_bits()));
        key.extend(sensor.sensor_from_rig.iter().map(|v| v.to_bits()));
        if let Some(&k) = self.index.get(&key) {
            return Some(k);
        }
        let k = u32::try_from(self.sensors.len()).ok()?;
        self.sensors.push(sensor);
        self.index.insert(key, k);
        Some(k)
    }
}

impl ProblemView {
    pub fn new(ba: &BundleAdjustment) -> Option<Self> {
        let mut declined = 0u32;
        if ba.gravity_prior.is_some()
            || ba.per_pose_gravity_prior.is_some()
            || ba.position_prior.is_some()
            || !ba.pairwise_pose_factors.is_empty()
        {
            declined |= ffi::VBA_FEATURE_PRIORS;
        }
        if !ba.imu_factors.is_empty()
            || !ba.velocities.is_empty()
            || !ba.biases.is_empty()
            || !ba.bias_random_walk_factors.is_empty()
            || ba.navigation_state_prior.is_some()
        {
            declined |= ffi::VBA_FEATURE_INERTIAL;
        }
        if !ba.fixed_pose_rotations.is_empty() {
            declined |= ffi::VBA_FEATURE_FIXED_ROTATIONS;
        }
        let pose_ids: Vec<u64> = ba.poses.keys().copied().collect();
        let poses: Vec<f64> = ba.poses.values().flat_map(|p| se3_array(&p.world_to_camera)).collect();
        let pose_fixed: Vec<u8> = pose_ids.iter().map(|id| ba.fixed_poses.contains(id) as u8).collect();
        let landmark_ids: Vec<u64> = ba.landmarks.keys().copied().collect();
        let landmarks: Vec<f64> = ba.landmarks.values().flat_map(|p| [p.x, p.y, p.z]).collect();
        let landmark_fixed: Vec<u8> = landmark_ids.iter().map(|id| ba.fixed_landmarks.contains(id) as u8).collect();
        let obs_pose: Vec<u64> = ba.observations.iter().map(|o| o.keyframe_id).collect();
        let obs_landmark: Vec<u64> = ba.observations.iter().map(|o| o.landmark_id).collect();
        let obs_xy: Vec<f64> = ba.observations.iter().flat_map(|o| [o.xy.x, o.xy.y]).collect();
        let mut rig_table = SensorTable::new();
        let mut rig_sensor = Vec::with_capacity(ba.rig_observations.len());
        for o in &ba.rig_observations {
            rig_sensor.push(rig_table.add(&o.camera, &o.sensor_from_rig)?);
        }
        let rig_pose: Vec<u64> = ba.rig_observations.iter().map(|o| o.keyframe_id).collect();
        let rig_landmark: Vec<u64> = ba.rig_observations.iter().map(|o| o.landmark_id).collect();
        let rig_xy: Vec<f64> = ba.rig_observations.iter().flat_map(|o| [o.xy.x, o.xy.y]).collect();
        let stereo_pose: Vec<u64> = ba.stereo_observations.iter().map(|o| o.keyframe_id).collect();
        let stereo_landmark: Vec<u64> = ba.stereo_observations.iter().map(|o| o.landmark_id).collect();
        let stereo_obs: Vec<f64> =
            ba.stereo_observations.iter().flat_map(|o| [o.xy.x, o.xy.y, o.u_right]).collect();
        let mut pair_table = SensorTable::new();
        let mut general_pair = Vec::with_capacity(ba.general_stereo_observations.len());
        for o in &ba.general_stereo_observations {
            general_pair.push(pair_table.add(&o.right_camera, &o.left_to_right)?);
        }
        let general_pose: Vec<u64> = ba.general_stereo_observations.iter().map(|o| o.keyframe_id).collect();
        let general_landmark: Vec<u64> = ba.general_stereo_observations.iter().map(|o| o.landmark_id).collect();
        let general_obs: Vec<f64> = ba
            .general_stereo_observations
            .iter()
            .flat_map(|o| [o.xy_left.x, o.xy_left.y, o.xy_right.x, o.xy_right.y])
            .collect();
        let sensors = rig_table.sensors;
        let pairs = pair_table.sensors;
        let stereo = Box::new(ffi::vba_stereo_terms {
            baseline: ba.stereo_baseline.unwrap_or(f64::NAN),
            num_stereo: stereo_pose.len() as u64,
            stereo_pose_ids: stereo_pose.as_ptr(),
            stereo_landmark_ids: stereo_landmark.as_ptr(),
            stereo_obs: stereo_obs.as_ptr(),
            num_general: general_pose.len() as u64,
            general_pose_ids: general_pose.as_ptr(),
            general_landmark_ids: general_landmark.as_ptr(),
            general_obs: general_obs.as_ptr(),
            general_pair: general_pair.as_ptr(),
            num_pairs: pairs.len() as u64,
            pairs: pairs.as_ptr(),
        });
        let problem = ffi::vba_problem {
            camera: camera(&ba.camera)?,
            num_poses: pose_ids.len() as u64,
            pose_ids: pose_ids.as_ptr(),
            poses: poses.as_ptr(),
            pose_fixed: pose_fixed.as_ptr(),
            num_landmarks: landmark_ids.len() as u64,
            landmark_ids: landmark_ids.as_ptr(),
            landmarks: landmarks.as_ptr(),
            landmark_fixed: landmark_fixed.as_ptr(),
            num_observations: obs_pose.len() as u64,
            obs_pose_ids: obs_pose.as_ptr(),
            obs_landmark_ids: obs_landmark.as_ptr(),
            obs_xy: obs_xy.as_ptr(),
            num_rig_observations: rig_pose.len() as u64,
            rig_pose_ids: rig_pose.as_ptr(),
            rig_landmark_ids: rig_landmark.as_ptr(),
            rig_xy: rig_xy.as_ptr(),
            rig_sensor: rig_sensor.as_ptr(),
            num_sensors: sensors.len() as u64,
            sensors: sensors.as_ptr(),
            declined_features: declined,
            reserved: 0,
            stereo: &*stereo,
        };
        Some(Self {
            pose_ids,
            poses,
            pose_fixed,
            landmark_ids,
            landmarks,
            landmark_fixed,
            obs_pose,
            obs_landmark,
            obs_xy,
            rig_pose,
            rig_landmark,
            rig_xy,
            rig_sensor,
            sensors,
            stereo_pose,
            stereo_landmark,
            stereo_obs,
            general_pose,
            general_landmark,
            general_obs,
            general_pair,
            pairs,
            stereo,
            problem,
        })
    }

    pub fn as_ffi(&self) -> &ffi::vba_problem {
        &self.problem
    }

    pub fn num_poses(&self) -> usize {
        self.pose_ids.len()
    }

    pub fn num_landmarks(&self) -> usize {
        self.landmark_ids.len()
    }

    /// Observations a weight vector covers (optimize.rs:43-53): monocular,
    /// rectified stereo, general stereo and rig.
    pub fn num_weighted_observations(&self) -> usize {
        self.num_all_observations()
    }

    pub fn num_all_observations(&self) -> usize {
        self.obs_pose.len() + self.stereo_pose.len() + self.general_pose.len() + self.rig_pose.len()
    }

    /// The solved state into `ba`, in the view's (BTreeMap) order.
    pub fn write_back(&self, ba: &mut BundleAdjustment, poses: &[f64], landmarks: &[f64]) {
        for (k, p) in ba.poses.values_mut().enumerate() {
            let v = &poses[7 * k..7 * k + 7];
            p.world_to_camera.rotation = UnitQuaternion::new_unchecked(Quaternion::new(v[0], v[1], v[2], v[3]));
            p.world_to_camera.translation = Vector3::new(v[4], v[5], v[6]);
        }
        for (k, x) in ba.landmarks.values_mut().enumerate() {
            *x = Point3::new(landmarks[3 * k], landmarks[3 * k + 1], landmarks[3 * k + 2]);
        }
    }
}
