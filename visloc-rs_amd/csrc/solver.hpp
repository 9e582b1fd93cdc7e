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

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::mutex mu;
    std::string last_error;
    // pinned staging for the per-iteration scalars
    StepScalars* host_scalars = nullptr;
    PcgState* host_pcg = nullptr;
    int32_t* host_status = nullptr;
    // profiling
    bool profile = false;
    std::vector<KernelProfile> kernels;
    struct Pending {
        hipEvent_t a, b;
        int kernel;
        double bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
};

// One problem laid out on the device: structure, initial state and working
// buffers. Solves start from the initial state; the final state stays on the
// device until read back.
struct Prepared {
    Model m;
    std::vector<DevBuffer> bufs;
    bool dense_built = false;
    std::vector<int32_t> dense_block_pq, dense_block_ptr, dense_terms, dense_var_slot;
    int final_slot = -1;  // buffer holding the final poses (-1: none yet)
    int final_lm_slot = -1;
};

// Matrix-free entry validation (optimize.rs:382-626) on a built model.
int32_t validate_matrix_free(const Model& m, const vba_options& o, std::string* message);

// Upload a built model (structure + initial state).
int32_t prepare(Context& ctx, Prepared& prep, std::string* message);

// The LM loop of optimize_weighted_backend (lm.rs:162-1017) on the device,
// from the prepared initial state.
int32_t run(Context& ctx, Prepared& prep, const vba_options& o, vba_iteration_stats* stats,
            uint32_t stats_capacity, vba_result* result);

// Final (or, before any solve, initial) poses / landmarks in input order.
int32_t download_state(Context& ctx, const Prepared& prep, double* poses_out, double* landmarks_out);

// Solve independent windows together (one device-resident batch, LM control
// on the device). models[i] is window i, valid[i] = 0 skips it; outputs are
// indexed by window; statuses[i] receives each window's status.
int32_t run_batch(Context& ctx, std::vector<Model>& models, const std::vector<int32_t>& valid,
                  const vba_options& o, double* const* poses_out, double* const* landmarks_out,
                  vba_iteration_stats* stats, uint32_t stats_capacity, vba_result* results, int32_t* statuses);

void release(Prepared& prep);
void release_context(Context& ctx);

}  // namespace vba
