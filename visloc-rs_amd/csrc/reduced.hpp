// reduced.hpp — the device reduced-camera solve (reduced.hip).
#pragma once

#include <cstdint>
#include <string>

#include "../../include/vba.h"

namespace vba {

struct Context;

// ReducedSolveAccelerator::solve (pipelines/slam/src/ba_accel.rs:138-145):
// VBA_OK with the step and the iteration count, VBA_DECLINED (the trait's
// None) for a system outside what the device solve supports or when the
// iteration breaks down, as pcg_core returns None.
int32_t reduced_solve(Context& ctx, const vba_reduced_system& system, double* step, uint32_t* iterations,
                      std::string* message);
// ReducedSolveAccelerator::eliminate (ba_accel.rs:150-156): VBA_OK with the
// run blocks and aggregate sums (the blocks stay on the device for the next
// reduced_solve of the structure without offdiag), VBA_DECLINED for input it
// does not take.
int32_t reduced_eliminate(Context& ctx, const vba_point_elimination& elimination, double* run_blocks,
                          double* coarse_sums, std::string* message);
void release_reduced(Context& ctx);

}  // namespace vba
