/*
 * vba_synth.h — seeded synthetic bundle-adjustment problems (libvba.so).
 *
 * The reference has no generator for these workloads; its own tests use a
 * sine-hash PRNG (pipelines/slam/src/bundle/parallel_ba_tests.rs:7-105) and
 * an LCG (crates/ba-gpu/src/reduced.rs:1030-1041). These generators use
 * SplitMix64 with 53-bit uniforms and Box–Muller normals, so a (kind, seed,
 * sizes) triple gives the same bytes on every machine.
 *
 * A generated problem is held by the library; vba_synth_view fills a
 * vba_problem whose pointers stay valid until vba_synth_free.
 */
#ifndef VBA_SYNTH_H
#define VBA_SYNTH_H

#include "vba.h"

#ifdef __cplusplus
extern "C" {
#endif

enum vba_synth_kind {
    /* Cameras evenly on a horizontal circle facing its centre, landmarks
     * uniform in a cube at the centre, shared camera, every in-frame
     * observation kept (50 cameras, radius 10 m, SIMPLE_RADIAL f=500 k=0.01
     * at 640x480, 5,000 landmarks in a 4 m cube at full size). */
    VBA_SYNTH_CIRCLE = 1,
    /* Forward-facing cameras along a closed loop, landmarks in a corridor
     * around it, each seen by a run of consecutive cameras (1,200 cameras on
     * a 200 m loop, PINHOLE f=800 at 1024x768, 400k landmarks in a 20 m
     * corridor at full size). */
    VBA_SYNTH_LOOP = 2,
    /* One local window of a stereo rig: keyframes along a short path, two
     * rig sensors (left at the rig origin, right 0.11 m to its right), each
     * landmark seen by 2..10 keyframes in both sensors, outliers replaced by
     * uniform pixels (PINHOLE f=460 at 752x480). */
    VBA_SYNTH_STEREO_WINDOW = 3,
    /* Pose-only refinement: one variable camera, fixed 3D points. */
    VBA_SYNTH_POSE_ONLY = 4,
    /* City-block orbit at 1.5 m height, landmarks on facade planes and the
     * ground, per-camera OPENCV intrinsics carried as rig sensors. */
    VBA_SYNTH_CITY = 5
};

typedef struct vba_synth_params {
    uint32_t kind;               /* enum vba_synth_kind */
    uint32_t num_cameras;        /* keyframes */
    uint64_t seed;
    uint64_t num_landmarks;
    double pixel_noise;          /* sigma, px */
    double rotation_noise;       /* sigma per axis, rad */
    double translation_noise;    /* sigma per axis of the camera centre, m */
    double landmark_noise;       /* sigma per axis, m */
    double outlier_fraction;     /* observations replaced by uniform pixels */
    double mean_track_length;    /* LOOP / CITY: mean run length */
    uint32_t min_track_length;
    uint32_t max_track_length;
    double scene_scale;          /* CIRCLE radius, LOOP length, CITY block side (m) */
    double focal;                /* focal length, px (CITY: mean) */
    double distortion;           /* CIRCLE: k of SIMPLE_RADIAL; CITY: sigma of k1 k2 p1 p2 */
    uint32_t width, height;
    uint32_t fixed_cameras;      /* first this many poses are fixed (gauge) */
    uint32_t reserved;
} vba_synth_params;

typedef struct vba_synth_problem vba_synth_problem;

/* Full-size parameters of the workload `kind` (the sizes in the comments of
 * enum vba_synth_kind). */
void vba_synth_default(uint32_t kind, vba_synth_params* params);
/* Generate; NULL on invalid parameters (message in vba_last_error(NULL)). */
vba_synth_problem* vba_synth_generate(const vba_synth_params* params);
void vba_synth_free(vba_synth_problem* problem);
/* Fill a problem view over the generated (perturbed, noisy) data. */
void vba_synth_view(const vba_synth_problem* problem, vba_problem* view);
/* Ground truth: poses (7 per pose) and landmarks (3 per landmark) before
 * perturbation, in the order of the view. Either pointer may be NULL. */
void vba_synth_truth(const vba_synth_problem* problem, const double** poses,
                     const double** landmarks);
/* Diagonal of the bounding box of the true landmarks: the scene extent the
 * camera-centre tolerance is expressed in. */
double vba_synth_scene_extent(const vba_synth_problem* problem);

#ifdef __cplusplus
}
#endif

#endif
