// capi_types.hpp — the opaque handle types of include/vba.h.
#pragma once

#include "solver.hpp"

struct vba_context : vba::Context {};
struct vba_prepared : vba::Prepared {};
struct vba_prepared_batch : vba::PreparedBatch {};
