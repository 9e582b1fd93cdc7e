// problem.hpp — host-side layout of one bundle-adjustment problem.
//
// The caller's id-keyed problem (vba_problem, the view of `BundleAdjustment`,
// pipelines/slam/src/bundle/mod.rs:529-622) becomes dense slots in ascending
// id order (the reference's BTreeMap order, lm.rs:234-249) and a
// structure-of-arrays layout sized for streaming from HBM:
//
//   observations   sorted by (pose slot, landmark slot, input index): the
//                  camera-sorted "c-order";
//   edges          runs of observations with the same (pose, landmark) in
//                  c-order (a stereo rig's two sensors seeing one landmark
//                  from one rig pose form one edge);
//   l-order        the same edges sorted by (landmark slot, pose slot), with
//                  the landmark CSR lm_edge_ptr and perm_cl (c -> l position);
//   camera tiles   slices of one pose's c-order edges, the work unit of every
//                  camera-keyed kernel (segmented reduction, no atomics);
//   landmark tiles runs of whole landmarks, the work unit of every
//                  landmark-keyed kernel.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vba.h"
#include "geom.hpp"

namespace vba {

// The linear solvers that form the reduced camera matrix densely (Dense, and
// Sparse, whose block Cholesky the device runs as a dense factor).
inline bool dense_family(uint32_t linear_solver) {
    return linear_solver == VBA_SOLVER_DENSE || linear_solver == VBA_SOLVER_SPARSE;
}
// Variable poses of a single dense reduced solve: its two n x n FP64 buffers
// (S and the LU copy) stay below 40 GB.
constexpr uint32_t kDenseMaxPoses = 8192;

// A lens plus the fixed extrinsic T_sensor<-rig as a rotation matrix and a
// translation. Sensor 0 is the shared camera of monocular and stereo
// observations with the identity extrinsic; then come the rig sensors, then
// the right cameras of general stereo pairs (pinhole lens, T_right<-left),
// whose `left` is the sensor of their left camera.
struct DevSensor {
    Lens<double> lens;
    double R[9];
    double t[3];
    int32_t identity;
    int32_t left;
};

struct CamTile {
    int32_t pose;   // pose slot
    int32_t begin;  // c-order edge range
    int32_t end;
    int32_t pad;
};
struct LmTile {
    int32_t lm_begin;  // landmark slot range
    int32_t lm_end;
    int32_t e_begin;   // its l-order edge range (lm_edge_ptr at lm_begin, lm_end)
    int32_t e_end;
    int32_t g_chunk;   // its first chunk of the lane-major FP32 G copy
};

// The FP32 copy of G that the mixed-precision operator streams is stored
// lane-major: a tile's sorted positions in chunks of kTileEdges (one chunk
// unless the tile is a landmark with more edges), chunk c holding position
// s = 4 i + k (lane i, slot k) at floats 1536 c + 384 k + 4 i (G0..G3) and
// 1536 c + 384 k + 256 + 2 i (G4, G5). A wave's load of one slot is then 1 KiB
// plus 512 bytes of consecutive memory. A position's chunk-relative index
// pp = 256 c + s is g32_pos.
constexpr int kG32Chunk = 1536;

// Observation kinds: 2-row pixel residuals (monocular and rig), rectified
// stereo (u_l, v_l, u_r) and general stereo (u_l, v_l, u_r, v_r).
enum ObsKind : uint8_t { OBS_PIXEL = 0, OBS_RECTIFIED = 1, OBS_GENERAL = 2 };

// Landmark tiles: runs of whole landmarks, each the work of one wave in the
// operator kernel: at most kTileEdges edges (four per lane), kTileLandmarks
// landmarks (one per lane) and kTileRuns camera runs (distinct poses, one
// LDS accumulator each). A landmark with more than kTileRuns edges forms a
// tile of its own, in which every edge is a run.
constexpr int kTileEdges = 256;
constexpr int kTileLandmarks = 64;
constexpr int kTileRuns = 64;

struct Model {
    uint32_t num_pose_slots = 0, num_lm_slots = 0;
    uint32_t P = 0, L = 0;  // variable poses / landmarks
    std::vector<uint32_t> pose_slot_of_input, lm_slot_of_input;
    std::vector<int32_t> pose_var, lm_var;  // variable index per slot, -1 = fixed
    std::vector<double> poses;              // 7 per slot (qw qx qy qz tx ty tz)
    std::vector<double> lms;                // 3 per slot
    std::vector<DevSensor> sensors;
    bool any_rig = false;
    bool any_stereo = false;  // rectified or general stereo observations
    bool any_fixed_pose = false;

    uint64_t num_obs = 0;
    std::vector<double> obs_uv;         // 2 per observation, c-order (the left pixel of a stereo one)
    std::vector<uint16_t> obs_sensor;   // c-order (empty unless any_rig or any_stereo)
    std::vector<uint8_t> obs_kind;      // c-order ObsKind (empty unless any_stereo)
    std::vector<double> obs_extra;      // 2 per observation, c-order (empty unless any_stereo):
                                        // rectified (u_r, baseline), general (u_r, v_r)
    std::vector<uint32_t> obs_input;    // c-order -> flattened input index
                                        // (mono, rectified stereo, general stereo, rig)

    uint64_t num_edges = 0;
    std::vector<int32_t> edge_pose, edge_lm;  // c-order
    std::vector<int32_t> edge_obs;            // num_edges + 1: first observation
    std::vector<int32_t> cam_edge_ptr;        // num_pose_slots + 1
    std::vector<int32_t> lm_edge_ptr;         // num_lm_slots + 1 (l-order)
    std::vector<int32_t> l_pose;              // l-order pose slot
    std::vector<int32_t> perm_cl;             // c-order edge -> storage position (below)

    std::vector<CamTile> cam_tiles;
    std::vector<int32_t> cam_tile_ptr;  // num_pose_slots + 1
    std::vector<LmTile> lm_tiles;
    int32_t max_lm_tile_edges = 0;
    // Camera runs of the landmark tiles: tile_perm[e] = tile-local l-order
    // index at sorted position e - (tile's first edge), sorting the tile's
    // edges by pose; a run is a maximal group of equal poses in that order.
    // Per-edge device data (G, h) is stored in this sorted order: s_pose is
    // the pose and s_lm the tile-local landmark at each storage position,
    // lm_store the storage position of each l-order edge. A run leaves its
    // partial in a slot: the slots of pose p are pose_run_ptr[p] ..
    // pose_run_ptr[p + 1] - 1, its runs in ascending storage order, and
    // run_slot gives the slot at the storage position of each run's last edge
    // (-1 at other positions).
    std::vector<uint16_t> tile_perm;
    std::vector<uint8_t> s_lm;
    std::vector<int32_t> s_pose, lm_store;
    std::vector<int32_t> big_tiles;  // tiles of one landmark with more than kTileEdges edges
    std::vector<int32_t> pose_run_ptr, run_slot;
    uint64_t num_runs = 0;
    std::vector<int32_t> g32_pos;  // per storage position (lane-major FP32 G)
    // The operator's view of a landmark tile t, read by tile index alone
    // (zero for tiles of one landmark with more than kTileRuns edges):
    // t_lm[kTileEdges t + s] the tile-local landmark at sorted position s;
    // t_head[4 t + w] bit b set when position 64 w + b starts a camera run;
    // t_rpose / t_rslot[kTileRuns t + r] the pose and the partial slot of the
    // tile's run r, runs numbered in position order.
    std::vector<uint8_t> t_lm;
    std::vector<uint64_t> t_head;
    std::vector<int32_t> t_rpose, t_rslot;
    std::vector<double> obs_w;     // external observation weights, c-order (empty: none)
    int32_t num_g32_chunks = 0;
};

// Observations of a problem, all kinds: the length of its flattened
// observation order (mono, rectified stereo, general stereo, rig).
inline uint64_t observation_count(const vba_problem& p) {
    return p.num_observations + p.num_rig_observations +
           (p.stereo ? p.stereo->num_stereo + p.stereo->num_general : 0);
}

// Camera::intrinsics (camera.rs:241-271): (fx, fy, cx, cy), false for a
// model without them.
bool camera_intrinsics(const vba_camera& camera, double* fx, double* fy, double* cx, double* cy);

// Lens of a vba_camera, resolved as MonoProjection::for_camera does
// (jacobians.rs:29-58). Returns false for an unsupported model.
bool lens_from_camera(const vba_camera& camera, Lens<double>* lens);

// Validate and lay out a problem. Status codes and their order follow
// optimize_weighted_backend (lm.rs:168-229); *error_id receives the missing
// id for VBA_ERR_MISSING_POSE / _LANDMARK. `message` explains a failure.
int32_t build_model(const vba_problem& problem, Model* model, uint64_t* error_id,
                    std::string* message);

// vba_check: VBA_OK or VBA_DECLINED with a reason.
// Options only: kernel, solver, precision and PCG policy enums in range and no
// intrinsics refinement. VBA_OK or VBA_DECLINED with the reason.
int32_t check_options(const vba_options& options, std::string* reason);
int32_t check_support(const vba_problem& problem, const vba_options& options, std::string* reason);

}  // namespace vba
