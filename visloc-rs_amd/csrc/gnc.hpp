// gnc.hpp — graduated non-convexity (gnc.cpp).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/vba.h"

namespace vba {

struct Context;
struct Prepared;

// GncState (pipelines/slam/src/gnc.rs:118-227): the surrogate family, the
// squared inlier scale and the control parameter mu.
struct GncState {
    uint32_t kernel;
    double c2, factor, mu;
    GncState(const vba_gnc_options& g, double max_residual_sq);
    void set_inlier_scale(double c) { c2 = c * c; }
    double weight(double s) const;
    bool is_terminal() const;
    bool anneal();
};

// estimate_scale_mad (gnc.rs:252-273): false for fewer than two finite
// residuals.
bool estimate_scale_mad(const std::vector<double>& squared, double k, double* out);

// optimize_gnc (lm.rs:28-142) on a prepared problem; its final state is left
// as the prepared state.
int32_t optimize_gnc(Context& ctx, Prepared& prep, const vba_options& options, const vba_gnc_options& gnc,
                     double* weights_out, vba_gnc_result* result);

}  // namespace vba
