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
void release_reduced(Context& ctx);

}  // namespace vba
