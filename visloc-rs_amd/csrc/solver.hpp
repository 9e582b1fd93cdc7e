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

// A device allocation that only grows; kept by the context between calls so
// repeated solves of similar size allocate nothing.
struct DevBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct Workspace {
    std::vector<DevBuffer> bufs;
    StepScalars* host_scalars = nullptr;  // pinned
    PcgState* host_pcg = nullptr;         // pinned
    int32_t* host_status = nullptr;       // pinned
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::mutex mu;
    std::string last_error;
    Workspace ws;
};

// Matrix-free entry validation (optimize.rs:382-626) on a built model.
int32_t validate_matrix_free(const Model& m, const vba_options& o, std::string* message);

// Run the LM loop of optimize_weighted_backend (lm.rs:162-1017) on the
// device. Writes final poses / landmarks in the caller's input order.
int32_t solve(Context& ctx, const Model& m, const vba_options& o, double* poses_out,
              double* landmarks_out, vba_iteration_stats* stats, uint32_t stats_capacity,
              vba_result* result);

void release_workspace(Workspace& ws);

}  // namespace vba
