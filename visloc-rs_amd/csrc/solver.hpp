// solver.hpp — the device Levenberg–Marquardt driver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vba.h"
#include "kernels.hpp"
#include "problem.hpp"

namespace vba {

// A device allocation that only grows.
struct DevBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Device time of one kernel family, from HIP events recorded around its
// launches on the context's stream while profiling is enabled.
struct KernelProfile {
    std::string name;
    uint64_t launches = 0;
    double seconds = 0.0;
    double bytes = 0.0;  // algorithmic bytes moved, summed over launches
};

struct ReducedState;  // reduced.hip

struct Context {
    int device = 0;
    int num_cus = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::mutex mu;
    std::string last_error;
    // pinned staging for the per-iteration scalars
    StepScalars* host_scalars = nullptr;
    PcgState* host_pcg = nullptr;
    int32_t* host_status = nullptr;
    // profiling: events around every `profile_stride`-th launch of each
    // kernel family (0: off)
    int32_t profile_stride = 0;
    std::vector<KernelProfile> kernels;
    std::vector<uint64_t> profile_seen;  // launches per family since enabling
    struct Pending {
        hipEvent_t a, b;
        int kernel;
        double bytes;
        int32_t pcg_iter;  // PCG iteration of the launch (-1: not a PCG launch)
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    // device buffers of the reduced-camera solve (reduced.hip), kept between calls
    ReducedState* reduced = nullptr;
    // PCG launch schedule (vba_tuning) and the pinned ring its chunk reads
    // land in, one event per slot
    vba_tuning tuning{10, 0, 0, 0};
    static constexpr int kPcgRing = 4;
    PcgState* host_pcg_ring = nullptr;
    hipEvent_t pcg_ev[kPcgRing] = {};
    uint32_t pcg_ring_next = 0;
};

// One problem laid out on the device: structure, initial state and working
// buffers. Solves start from the initial state; the final state stays on the
// device until read back.
struct Prepared {
    Model m;
    std::vector<DevBuffer> bufs;
    uint64_t alloc_gen = 0;  // counts (re)allocations of bufs: device pointers change with it
    bool dense_built = false;
    std::vector<int32_t> dense_block_pq, dense_block_ptr, dense_terms, dense_var_slot;
    // explicit reduced camera matrix (over the dense term lists): full block
    // rows, the full positions of each upper block and of each diagonal
    bool explicit_built = false;
    std::vector<int32_t> x_row_ptr, x_col, x_block_pos, x_diag_pos;
    int final_slot = -1;  // buffer holding the final poses (-1: none yet)
    int final_lm_slot = -1;
    int32_t last_pcg = 0;  // iterations of the last PCG solve (the schedule's prediction)
    Lens<double> joint_lens{};  // the camera after a joint camera solve
};

// Matrix-free entry validation (optimize.rs:382-626) on a built model.
int32_t validate_matrix_free(const Model& m, const vba_options& o, std::string* message);

// Upload a built model (structure + initial state).
int32_t prepare(Context& ctx, Prepared& prep, std::string* message);

// The LM loop of optimize_weighted_backend (lm.rs:162-1017) on the device,
// from the prepared initial state.
// A joint pose / landmark / camera solve (optimize_joint_intrinsics,
// joint_intrinsics.rs:24-353): k camera parameters (4, 6 or 8) border the
// dense reduced system; shared: fx and fy move together. The model's
// sensor 0 holds the starting camera.
struct JointSpec {
    int32_t k;
    int32_t shared;
};
int32_t run(Context& ctx, Prepared& prep, const vba_options& o, vba_iteration_stats* stats,
            uint32_t stats_capacity, vba_result* result, const JointSpec* joint = nullptr);

// solve_normal_equations (pose_graph.rs:2259-2270) on the device: H x = b
// for a column-major n x n H, Cholesky with the LU fallback.
// VBA_ERR_SINGULAR_SYSTEM when both fail; x is then all zero.
int32_t dense_solve(Context& ctx, uint32_t n, const double* h, const double* b, double* x, std::string* message);

// The same solve with every device buffer in the library's fixed device
// array (scratch.hip): no device allocation. n <= VBA_DENSE_NOALLOC_MAX_N.
int32_t dense_solve_noalloc(Context& ctx, uint32_t n, const double* h, const double* b, double* x,
                            std::string* message);
// Looks up the device address of that array on `device` (current) once.
hipError_t scratch_setup(int device);

// Final (or, before any solve, initial) poses / landmarks in input order.
int32_t download_state(Context& ctx, const Prepared& prep, double* poses_out, double* landmarks_out);

// External observation weights, one per observation in the flattened input
// order (monocular, then rig), for the following solves; NULL removes them.
int32_t set_weights(Context& ctx, Prepared& prep, const double* weights);

// Make the last solve's final state the prepared initial state, so that the
// next solve continues from it.
int32_t adopt_final_state(Context& ctx, Prepared& prep);

// Squared reprojection residual of every observation at the prepared initial
// state, in the flattened input order; NaN where it cannot be evaluated.
int32_t squared_residuals(Context& ctx, Prepared& prep, const vba_options& o, double* out);

// Independent windows laid out as one device-resident problem with disjoint
// slots; the LM control runs per window on the device.
struct PreparedBatch : Prepared {
    uint32_t count = 0;                   // the caller's windows
    std::vector<int32_t> index;           // batch window -> caller index
    std::vector<int32_t> window_of;       // caller index -> batch window, -1 if not in the batch
    std::vector<WindowLayout> layout;
    std::vector<int64_t> work_offset;     // per window: dense solve workspace
    int64_t s_total = 0, work_total = 0;  // doubles
    int32_t num_dense_blocks = 0;
    std::vector<std::vector<uint32_t>> pose_slot_of_input, lm_slot_of_input;  // per batch window
    std::vector<uint64_t> num_obs;                                           // per batch window
    bool solved = false;
    // the LM iterations recorded as a HIP graph, with what it was recorded for
    hipGraphExec_t graph = nullptr;
    std::vector<uint8_t> graph_key;
};

// Lay out the valid windows (valid[i] = 1) of `models` and upload them.
// A window with more than 256 variable poses gets VBA_ERR_INVALID_ARGUMENT
// in statuses[i] and stays out of the batch.
int32_t prepare_batch(Context& ctx, std::vector<Model>& models, const std::vector<int32_t>& valid,
                      PreparedBatch* batch, int32_t* statuses, std::string* message);
// Solve every window of the batch from its prepared initial state; results,
// stats (stats_capacity per caller index) and statuses by caller index.
// With `graph` (and o.flags without VBA_FLAG_NO_GRAPH) the LM iterations are
// recorded as a HIP graph on the first solve and replayed while the options,
// the stats capacity and the device buffers stay the same.
int32_t run_batch(Context& ctx, PreparedBatch& batch, const vba_options& o, vba_iteration_stats* stats,
                  uint32_t stats_capacity, vba_result* results, int32_t* statuses, bool graph = false);
void release(PreparedBatch& batch);
// Final (before any solve: initial) state of the caller's window i.
int32_t download_batch_window(Context& ctx, const PreparedBatch& batch, uint32_t i, double* poses_out,
                              double* landmarks_out);

void release(Prepared& prep);
void release_context(Context& ctx);

}  // namespace vba
