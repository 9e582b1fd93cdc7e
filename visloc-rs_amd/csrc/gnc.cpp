// gnc.cpp — graduated non-convexity around the device LM solve:
// BundleAdjustment::optimize_gnc (pipelines/slam/src/bundle/lm.rs:28-142)
// with GncState and estimate_scale_mad (pipelines/slam/src/gnc.rs:118-286).
//
// The outer loop runs on the host. Each level reads the squared
// reprojection residuals from the device, turns them into the surrogate's
// Black–Rangarajan weights, uploads the weights and runs a short weighted LM
// with no M-estimator on the device, which continues from the previous
// level's state.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "gnc.hpp"
#include "solver.hpp"

namespace vba {

namespace {

constexpr double kTlsMuTerminal = 200.0;  // gnc.rs:128-131

}  // namespace

GncState::GncState(const vba_gnc_options& g, double max_residual_sq)
    : kernel(g.kernel), c2(g.c * g.c), factor(g.anneal_factor) {
    const double s_max = std::max(max_residual_sq, 0.0);
    if (kernel == VBA_GNC_GEMAN_MCCLURE) {
        mu = std::max(2.0 * s_max / c2, 1.0);
    } else {
        const double denom = 2.0 * s_max - c2;
        mu = denom <= 0.0 ? kTlsMuTerminal
                          : std::min(std::max(c2 / denom, std::numeric_limits<double>::min()), kTlsMuTerminal);
    }
}

double GncState::weight(double s) const {
    s = std::max(s, 0.0);
    if (kernel == VBA_GNC_GEMAN_MCCLURE) {
        const double mc2 = mu * c2;
        const double r = mc2 / (s + mc2);
        return r * r;
    }
    const double lo = (mu / (mu + 1.0)) * c2;
    const double hi = ((mu + 1.0) / mu) * c2;
    if (s <= lo) return 1.0;
    if (s >= hi) return 0.0;
    return std::sqrt(c2 * mu * (mu + 1.0) / s) - mu;
}

bool GncState::is_terminal() const {
    return kernel == VBA_GNC_GEMAN_MCCLURE ? mu <= 1.0 : mu >= kTlsMuTerminal;
}

bool GncState::anneal() {
    if (kernel == VBA_GNC_GEMAN_MCCLURE) mu = std::max(mu / factor, 1.0);
    else mu = std::min(mu * factor, kTlsMuTerminal);
    return is_terminal();
}

namespace {

// median_in_place: sort by f64::total_cmp (no NaN here), middle element or
// the mean of the two middle ones.
double median_in_place(std::vector<double>& v) {
    std::sort(v.begin(), v.end(), [](double a, double b) {
        return a < b || (a == b && std::signbit(a) && !std::signbit(b));
    });
    const size_t n = v.size();
    return n % 2 == 1 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

// The weighted robust cost with no M-estimator (robust_cost_weighted with
// RobustKernel::None, problem.rs:497-583) from the residuals: observations
// that cannot be evaluated (NaN) are skipped.
double weighted_l2(const std::vector<double>& s, const std::vector<double>& w) {
    double total = 0.0;
    for (size_t i = 0; i < s.size(); ++i)
        if (!std::isnan(s[i])) total += w[i] * s[i];
    return total;
}

}  // namespace

bool estimate_scale_mad(const std::vector<double>& squared, double k, double* out) {
    std::vector<double> norms;
    norms.reserve(squared.size());
    for (double s : squared)
        if (std::isfinite(s) && s >= 0.0) norms.push_back(std::sqrt(s));
    if (norms.size() < 2) return false;
    const double median = median_in_place(norms);
    std::vector<double> dev;
    dev.reserve(norms.size());
    for (double r : norms) dev.push_back(std::fabs(r - median));
    const double mad = median_in_place(dev);
    *out = median + k * (1.4826 * mad);
    return true;
}

int32_t optimize_gnc(Context& ctx, Prepared& pr, const vba_options& options, const vba_gnc_options& gnc,
                     double* weights_out, vba_gnc_result* res) {
    const size_t n = pr.m.num_obs;
    std::vector<double> s(n), w(n, 1.0);
    int32_t st = squared_residuals(ctx, pr, options, s.data());
    if (st != VBA_OK) return st;
    const double initial_cost = weighted_l2(s, w);
    double s_max = 0.0;
    for (double v : s)
        if (std::isfinite(v)) s_max = std::max(s_max, v);
    vba_gnc_options eff = gnc;
    if (gnc.has_auto_scale) {
        double est;
        if (estimate_scale_mad(s, gnc.auto_scale, &est)) eff.c = std::max(est, gnc.c);
    }
    double inlier_scale = eff.c;
    GncState state(eff, s_max);
    vba_options inner = options;
    inner.robust_kernel = VBA_KERNEL_NONE;
    inner.max_iterations = std::max<uint32_t>(gnc.inner_iterations, 1);
    bool converged = false;
    uint32_t outer = 0;
    for (uint32_t it = 0; it < std::max<uint32_t>(gnc.max_outer, 1); ++it) {
        ++outer;
        const bool terminal = state.is_terminal();
        if (it > 0 && (st = squared_residuals(ctx, pr, options, s.data())) != VBA_OK) return st;
        if (gnc.auto_scale_readapt && gnc.has_auto_scale) {
            double est;
            if (estimate_scale_mad(s, gnc.auto_scale, &est)) {
                const double c = std::max(est, gnc.c);
                state.set_inlier_scale(c);
                inlier_scale = c;
            }
        }
        for (size_t i = 0; i < n; ++i) w[i] = std::isfinite(s[i]) ? state.weight(s[i]) : 1.0;
        if ((st = set_weights(ctx, pr, w.data())) != VBA_OK) return st;
        vba_result r{};
        if ((st = run(ctx, pr, inner, nullptr, 0, &r)) != VBA_OK) return st;
        if ((st = adopt_final_state(ctx, pr)) != VBA_OK) return st;
        if (terminal) {
            converged = true;
            break;
        }
        state.anneal();
    }
    if ((st = squared_residuals(ctx, pr, options, s.data())) != VBA_OK) return st;
    for (size_t i = 0; i < n; ++i)
        w[i] = std::isfinite(s[i]) ? state.weight(s[i]) : std::numeric_limits<double>::quiet_NaN();
    std::vector<double> mask(n);
    for (size_t i = 0; i < n; ++i)
        mask[i] = std::isfinite(w[i]) ? (w[i] >= 0.5 ? 1.0 : 0.0) : std::numeric_limits<double>::quiet_NaN();
    res->initial_cost = initial_cost;
    res->final_cost = weighted_l2(s, w);
    res->inlier_cost = weighted_l2(s, mask);
    res->inlier_scale = inlier_scale;
    res->observation_count = n;
    res->outer_iterations = outer;
    res->converged = converged ? 1 : 0;
    if (weights_out) std::copy(w.begin(), w.end(), weights_out);
    return VBA_OK;
}

}  // namespace vba
