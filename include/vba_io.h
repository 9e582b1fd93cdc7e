/*
 * vba_io.h — bundle-adjustment problems on disk, for libvba.so.
 *
 * One host-side model type, `vba_model`, is read from and written to the
 * reference's two on-disk forms of this path's input:
 *   - the plain-text problem dump of pipelines/slam/src/ba_problem_io.rs
 *     (what the incremental SfM writes under VISLOC_SFM_BA_DUMP,
 *     refinement.rs:450-461, and what ba_replay / ba_gpu_replay load);
 *   - COLMAP sparse models, text (cameras.txt, images.txt, points3D.txt) and
 *     binary (cameras.bin, images.bin, points3D.bin), as read and written by
 *     crates/io/src/colmap/mod.rs.
 * vba_model_problem gives the model's `vba_problem` view for vba_optimize
 * (and the rest of vba.h); vba_model_set_state takes the solved poses and
 * landmarks back, so that the model can be written out again.
 *
 * Conventions are those of vba.h: plain pointers and sizes, status codes,
 * and a message for the last failure (per calling thread here, since a
 * failed read has no model to hold it).
 */
#ifndef VBA_IO_H
#define VBA_IO_H

#include <stdint.h>

#include "vba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of this header, beside vba.h's. */
enum vba_io_status {
    VBA_ERR_IO = -110,           /* ColmapError::Io / io::Error: a file could not be read or written */
    VBA_ERR_FORMAT = -111,       /* malformed content: ColmapError::{ParseInt, ParseFloat,
                                    InvalidLine, InvalidBinary}; io::ErrorKind::InvalidData */
    VBA_ERR_EXPORT = -112        /* ColmapError::InvalidExportInput: the model cannot be
                                    written in the asked format */
};

typedef struct vba_model vba_model;

/* Message of the calling thread's last failed call of this header. */
const char* vba_io_last_error(void);

/* read_ba_problem (pipelines/slam/src/ba_problem_io.rs:79-149): a pinhole
 * camera with id 1, poses (quaternions normalised as
 * UnitQuaternion::from_quaternion does) with their fixed flags, landmarks,
 * and observations in file order with their optional weights. A weighted
 * problem sets VBA_FEATURE_OBSERVATION_WEIGHTS in the view's
 * declined_features, as the SfM sends weighted solves to its CPU solver
 * (refinement.rs:469-473). */
int32_t vba_model_read_ba_problem(const char* path, vba_model** out);

/* write_ba_problem (ba_problem_io.rs:20-72): floats in the shortest text that
 * reads back to the same double, so a reload is exact. Needs a PINHOLE
 * camera; writes the monocular observations only, as the reference does. */
int32_t vba_model_write_ba_problem(const vba_model* model, const char* path);

/* read_colmap_text_model / read_colmap_binary_model (crates/io/src/colmap/
 * mod.rs:148-196; parsers :989-1295): binary != 0 reads the .bin files.
 * Poses are the images' world-to-camera poses; observations are the image
 * keypoints with a POINT3D_ID >= 0, by image id, then keypoint index; every
 * camera of the model is kept. The gauge is the SfM's global-BA gauge
 * (refinement.rs:316-352): the image with the lowest id and the image whose
 * camera centre is farthest from it are fixed. */
int32_t vba_model_read_colmap(const char* dir, int32_t binary, vba_model** out);

/* write_colmap_text_model / write_colmap_binary_model (mod.rs:198-236;
 * formatters :243-330, :900-987): the reference's records (images named
 * image_<id>.jpg, points white with zero error, each track sorted by
 * (image id, keypoint index)). The directory is created if missing; nothing
 * is written when a record cannot be encoded. */
int32_t vba_model_write_colmap(const vba_model* model, const char* dir, int32_t binary);

/* A model holding a copy of `problem` (its monocular part: camera, poses
 * with fixed flags, landmarks, observations) and of `weights` (one per
 * observation, or NULL): the `BundleAdjustment` that write_ba_problem
 * takes. Rig observations are refused (VBA_ERR_INVALID_ARGUMENT). */
int32_t vba_model_from_problem(const vba_problem* problem, const double* weights, vba_model** out);

/* The model as a bundle-adjustment problem. The view (and *weights, one per
 * monocular observation or NULL when unweighted) stays valid until the
 * model is changed or freed. When the observations use one camera they are
 * monocular observations on `camera`; when they use several they are rig
 * observations with one sensor per camera used (identity sensor_from_rig),
 * in ascending camera id.
 * VBA_ERR_UNSUPPORTED_CAMERA_MODEL when an observed camera is not in the
 * pinhole family of vba.h (BaError::UnsupportedCameraModel). */
int32_t vba_model_problem(const vba_model* model, vba_problem* out, const double** weights);

/* Replace the poses (7 per pose) and landmarks (3 per landmark), in the
 * view's order: the output of vba_optimize. Either may be NULL. */
int32_t vba_model_set_state(vba_model* model, const double* poses, const double* landmarks);

/* Replace the fixed flags, one per pose in the view's order (nonzero =
 * fixed). */
int32_t vba_model_set_fixed(vba_model* model, const uint8_t* fixed);

/* The local-window shape of ba_gpu_replay --local N
 * (crates/ba-gpu/examples/ba_gpu_replay.rs:43-66): the last n poses by id
 * stay variable, only the landmarks they observe and the observations of
 * those landmarks are kept, then only the poses those observations use,
 * every pose outside the n fixed. */
int32_t vba_model_keep_local_window(vba_model* model, uint32_t n);

void vba_model_free(vba_model* model);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VBA_IO_H */
