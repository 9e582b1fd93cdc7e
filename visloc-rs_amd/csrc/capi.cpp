// capi.cpp — the extern "C" entry points of include/vba.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/vba.h"
#include "host_common.hpp"
#include "problem.hpp"
#include "capi_types.hpp"
#include "reduced.hpp"
#include "solver.hpp"
#include "gnc.hpp"



namespace {

// Validate and lay out a problem into `prep` (host side of vba_prepare).
int32_t build_prepared(vba_context* ctx, const vba_problem& problem, const vba_options& options,
                       vba_prepared* prep, vba_result* result) {
    std::string reason;
    int32_t st = vba::check_support(problem, options, &reason);
    if (st != VBA_OK) {
        ctx->last_error = "declined: " + reason;
        return st;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->last_error = "hipSetDevice failed";
        return VBA_ERR_DEVICE;
    }
    uint64_t error_id = 0;
    std::string msg;
    try {
        st = vba::build_model(problem, &prep->m, &error_id, &msg);
    } catch (const std::bad_alloc&) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
    if (result) result->error_id = error_id;
    if (st != VBA_OK) {
        ctx->last_error = msg.empty() ? "invalid problem" : msg;
        return st;
    }
    if (!vba::dense_family(options.linear_solver)) {
        st = vba::validate_matrix_free(prep->m, options, &msg);
        if (st != VBA_OK) {
            ctx->last_error = msg;
            return st;
        }
    }
    st = vba::prepare(*ctx, *prep, &msg);
    if (st != VBA_OK) ctx->last_error = msg;
    return st;
}

}  // namespace

extern "C" {

uint32_t vba_abi_version(void) { return VBA_ABI_VERSION; }

vba_context* vba_create(int32_t device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        vba::set_thread_error("vba_create: no HIP device available");
        return nullptr;
    }
    if (device < 0 || device >= count) {
        vba::set_thread_error("vba_create: device %d out of range (%d devices)", device, count);
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        vba::set_thread_error("vba_create: hipSetDevice(%d) failed", device);
        return nullptr;
    }
    std::unique_ptr<vba_context> ctx(new (std::nothrow) vba_context());
    if (!ctx) {
        vba::set_thread_error("vba_create: out of host memory");
        return nullptr;
    }
    ctx->device = device;
    if (hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        ctx->num_cus = 0;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess) {
        vba::set_thread_error("vba_create: stream/event creation failed");
        return nullptr;
    }
    // the pinned status word of the dense solves, and the device address of
    // the no-allocation solve's array (looked up once per device)
    if (hipHostMalloc((void**)&ctx->host_status, sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        ctx->host_status = nullptr;
        vba::set_thread_error("vba_create: pinned status allocation failed");
        (void)hipEventDestroy(ctx->ev_start);
        (void)hipEventDestroy(ctx->ev_stop);
        (void)hipStreamDestroy(ctx->stream);
        return nullptr;
    }
    (void)vba::scratch_setup(device);
    return ctx.release();
}

void vba_destroy(vba_context* ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        vba::release_context(*ctx);
        if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
        if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

const char* vba_last_error(const vba_context* ctx) {
    if (!ctx) return vba::thread_error().c_str();
    return ctx->last_error.c_str();
}

void vba_default_options(vba_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    // BaConfig::default (pipelines/slam/src/bundle/mod.rs:729-752)
    o->max_iterations = 20;
    o->has_initial_lambda = 1;
    o->initial_lambda = 1e-4;
    o->lambda_increase_factor = 10.0;
    o->lambda_decrease_factor = 0.1;
    o->max_lambda = 1e12;
    o->min_lambda = 1e-9;
    o->step_tolerance = 1e-7;
    o->cost_tolerance = 1e-9;
    o->robust_kernel = VBA_KERNEL_NONE;
    o->linear_solver = VBA_SOLVER_DENSE;
    o->precision = VBA_PRECISION_FP64;
    // MatrixFreeBaOptions::default (mod.rs:789-798)
    o->max_pcg_iterations = 128;
    o->pcg_relative_tolerance = 1e-12;
    o->pcg_absolute_tolerance = 1e-12;
    o->pcg_maxit_policy = VBA_PCG_MAXIT_REJECT;
}

int32_t vba_check(const vba_problem* problem, const vba_options* options) {
    if (!problem || !options) {
        vba::set_thread_error("vba_check: null argument");
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::string reason;
    int32_t st = vba::check_support(*problem, *options, &reason);
    if (st != VBA_OK) vba::set_thread_error("declined: %s", reason.c_str());
    return st;
}

int32_t vba_optimize(vba_context* ctx, const vba_problem* problem, const vba_options* options, double* poses_out,
                     double* landmarks_out, vba_iteration_stats* stats, uint32_t stats_capacity,
                     vba_result* result) {
    if (!ctx || !problem || !options || !result) {
        vba::set_thread_error("vba_optimize: null argument");
        if (ctx) ctx->last_error = "vba_optimize: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::memset(result, 0, sizeof(*result));
    ctx->last_error.clear();
    vba_prepared prep;
    int32_t st = build_prepared(ctx, *problem, *options, &prep, result);
    if (st == VBA_OK) {
        try {
            st = vba::run(*ctx, prep, *options, stats, stats_capacity, result);
        } catch (const std::bad_alloc&) {
            ctx->last_error = "out of host memory";
            st = VBA_ERR_OUT_OF_MEMORY;
        }
    }
    if (st == VBA_OK) st = vba::download_state(*ctx, prep, poses_out, landmarks_out);
    vba::release(prep);
    return st;
}

int32_t vba_optimize_intrinsics(vba_context* ctx, const vba_problem* problem, const vba_options* options,
                                double* poses_out, double* landmarks_out, vba_camera* camera_out,
                                vba_iteration_stats* stats, uint32_t stats_capacity, vba_result* result) {
    if (!ctx || !problem || !options || !result || !camera_out) {
        vba::set_thread_error("vba_optimize_intrinsics: null argument");
        if (ctx) ctx->last_error = "vba_optimize_intrinsics: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    *camera_out = problem->camera;
    const vba_camera& c = problem->camera;
    const uint32_t flags = options->refine_intrinsics;
    double fx, fy, cx, cy;
    // BundleAdjustment::optimize (optimize.rs:8-28): the joint solve for a
    // PINHOLE / OPENCV camera with intrinsics and neither general stereo nor
    // rig observations; the pose / structure solve otherwise
    const bool joint = (flags & VBA_REFINE_INTRINSICS) && !(problem->stereo && problem->stereo->num_general) &&
                       problem->num_rig_observations == 0 &&
                       (c.model == VBA_CAMERA_PINHOLE || c.model == VBA_CAMERA_OPENCV) &&
                       vba::camera_intrinsics(c, &fx, &fy, &cx, &cy);
    vba_options plain = *options;
    plain.refine_intrinsics = 0;
    if (!joint) return vba_optimize(ctx, problem, &plain, poses_out, landmarks_out, stats, stats_capacity, result);
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::memset(result, 0, sizeof(*result));
    ctx->last_error.clear();
    std::string reason;
    int32_t st = vba::check_support(*problem, plain, &reason);
    if (st != VBA_OK) {
        ctx->last_error = "declined: " + reason;
        return st;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->last_error = "hipSetDevice failed";
        return VBA_ERR_DEVICE;
    }
    vba_prepared prep;
    uint64_t error_id = 0;
    std::string msg;
    try {
        st = vba::build_model(*problem, &prep.m, &error_id, &msg);
    } catch (const std::bad_alloc&) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
    result->error_id = error_id;
    // the joint solve refines the camera even with every pose and landmark fixed
    if (st != VBA_OK && st != VBA_ERR_ALL_POSES_FIXED) {
        ctx->last_error = msg.empty() ? "invalid problem" : msg;
        return st;
    }
    // the camera block (joint_intrinsics.rs:49-99): (k1, k2) on a monocular
    // problem when asked, (p1, p2) on top (the camera then becomes OPENCV),
    // fx = fy = their mean for a shared focal
    const bool any_stereo = problem->stereo && problem->stereo->num_stereo;
    const bool dist = (flags & VBA_REFINE_DISTORTION) && !any_stereo;
    const bool tan = dist && (flags & VBA_REFINE_TANGENTIAL);
    vba_camera out = c;
    if (dist)
        while (out.num_params < 6) out.params[out.num_params++] = 0.0;
    if (tan) {
        out.model = VBA_CAMERA_OPENCV;
        while (out.num_params < 8) out.params[out.num_params++] = 0.0;
    }
    vba::Lens<double>& L = prep.m.sensors[0].lens;
    vba::lens_from_camera(out, &L);
    if (flags & VBA_REFINE_SHARED_FOCAL) L.fx = L.fy = 0.5 * (L.fx + L.fy);
    // the lens family that the refined parameters can reach
    if (tan) L.kind = vba::LENS_OPENCV;
    else if (dist && L.kind == vba::LENS_PINHOLE) L.kind = vba::LENS_RADIAL;
    const vba::JointSpec spec{tan ? 8 : dist ? 6 : 4, (flags & VBA_REFINE_SHARED_FOCAL) ? 1 : 0};
    plain.linear_solver = VBA_SOLVER_DENSE;
    st = vba::prepare(*ctx, prep, &msg);
    if (st != VBA_OK) ctx->last_error = msg;
    if (st == VBA_OK) {
        try {
            st = vba::run(*ctx, prep, plain, stats, stats_capacity, result, &spec);
        } catch (const std::bad_alloc&) {
            ctx->last_error = "out of host memory";
            st = VBA_ERR_OUT_OF_MEMORY;
        }
    }
    if (st == VBA_OK) st = vba::download_state(*ctx, prep, poses_out, landmarks_out);
    if (st == VBA_OK) {
        const vba::Lens<double>& F = prep.joint_lens;
        const double v[8] = {F.fx, F.fy, F.cx, F.cy, F.k1, F.k2, F.p1, F.p2};
        for (int i = 0; i < spec.k; ++i) out.params[i] = v[i];
        *camera_out = out;
    }
    vba::release(prep);
    return st;
}

int32_t vba_optimize_weighted(vba_context* ctx, const vba_problem* problem, const vba_options* options,
                              const double* weights, uint64_t num_weights, double* poses_out, double* landmarks_out,
                              vba_iteration_stats* stats, uint32_t stats_capacity, vba_result* result) {
    if (!ctx || !problem || !options || !result) {
        vba::set_thread_error("vba_optimize_weighted: null argument");
        if (ctx) ctx->last_error = "vba_optimize_weighted: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::memset(result, 0, sizeof(*result));
    ctx->last_error.clear();
    // validate_observation_weights (optimize.rs:628-650), then the intrinsics
    // refusal (:49-51)
    const uint64_t expected = vba::observation_count(*problem);
    if (num_weights != expected || (num_weights && !weights)) {
        ctx->last_error = "observation weight count " + std::to_string(num_weights) + ", expected " +
                          std::to_string(expected);
        return VBA_ERR_OBSERVATION_WEIGHT_COUNT;
    }
    for (uint64_t i = 0; i < num_weights; ++i)
        if (!std::isfinite(weights[i]) || weights[i] < 0.0) {
            result->error_id = i;
            ctx->last_error = "observation weight " + std::to_string(i) + " is negative or not finite";
            return VBA_ERR_INVALID_OBSERVATION_WEIGHT;
        }
    if (options->refine_intrinsics) {
        ctx->last_error = "observation weights with intrinsics refinement";
        return VBA_ERR_WEIGHTS_WITH_INTRINSICS;
    }
    vba_problem view = *problem;
    view.declined_features &= ~(uint32_t)VBA_FEATURE_OBSERVATION_WEIGHTS;
    vba_prepared prep;
    int32_t st = build_prepared(ctx, view, *options, &prep, result);
    if (st == VBA_OK) st = vba::set_weights(*ctx, prep, weights);
    if (st == VBA_OK) {
        try {
            st = vba::run(*ctx, prep, *options, stats, stats_capacity, result);
        } catch (const std::bad_alloc&) {
            ctx->last_error = "out of host memory";
            st = VBA_ERR_OUT_OF_MEMORY;
        }
    }
    if (st == VBA_OK) st = vba::download_state(*ctx, prep, poses_out, landmarks_out);
    vba::release(prep);
    return st;
}

void vba_default_gnc_options(vba_gnc_options* gnc) {
    if (!gnc) return;
    std::memset(gnc, 0, sizeof(*gnc));
    gnc->kernel = VBA_GNC_GEMAN_MCCLURE;  // GncConfig::default (gnc.rs:103-115)
    gnc->c = 1.0;
    gnc->anneal_factor = 1.4;
    gnc->max_outer = 100;
    gnc->inner_iterations = 5;
}

int32_t vba_optimize_gnc(vba_context* ctx, const vba_problem* problem, const vba_options* options,
                         const vba_gnc_options* gnc, double* poses_out, double* landmarks_out, double* weights_out,
                         vba_gnc_result* result) {
    if (!ctx || !problem || !options || !gnc || !result) {
        vba::set_thread_error("vba_optimize_gnc: null argument");
        if (ctx) ctx->last_error = "vba_optimize_gnc: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::memset(result, 0, sizeof(*result));
    ctx->last_error.clear();
    if (!(gnc->c > 0.0) || !std::isfinite(gnc->c) || !(gnc->anneal_factor > 1.0) ||
        !std::isfinite(gnc->anneal_factor) || gnc->kernel > VBA_GNC_TRUNCATED_LEAST_SQUARES) {
        ctx->last_error = "vba_optimize_gnc: c must be positive and anneal_factor above 1";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    vba_prepared prep;
    vba_result r{};
    int32_t st = build_prepared(ctx, *problem, *options, &prep, &r);
    if (st == VBA_OK) {
        try {
            st = vba::optimize_gnc(*ctx, prep, *options, *gnc, weights_out, result);
        } catch (const std::bad_alloc&) {
            ctx->last_error = "out of host memory";
            st = VBA_ERR_OUT_OF_MEMORY;
        }
    }
    if (st == VBA_OK) st = vba::download_state(*ctx, prep, poses_out, landmarks_out);
    vba::release(prep);
    return st;
}

int32_t vba_prepare(vba_context* ctx, const vba_problem* problem, const vba_options* options, vba_prepared** out,
                    vba_result* result) {
    if (!ctx || !problem || !options || !out) {
        vba::set_thread_error("vba_prepare: null argument");
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = nullptr;
    if (result) std::memset(result, 0, sizeof(*result));
    ctx->last_error.clear();
    std::unique_ptr<vba_prepared> prep(new (std::nothrow) vba_prepared());
    if (!prep) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
    int32_t st = build_prepared(ctx, *problem, *options, prep.get(), result);
    if (st != VBA_OK) {
        vba::release(*prep);
        return st;
    }
    *out = prep.release();
    return VBA_OK;
}

int32_t vba_solve_prepared(vba_context* ctx, vba_prepared* prepared, const vba_options* options,
                           vba_iteration_stats* stats, uint32_t stats_capacity, vba_result* result) {
    if (!ctx || !prepared || !options || !result) {
        vba::set_thread_error("vba_solve_prepared: null argument");
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::memset(result, 0, sizeof(*result));
    ctx->last_error.clear();
    std::string msg;
    if (vba::check_options(*options, &msg) != VBA_OK) {
        ctx->last_error = "declined: " + msg;
        return VBA_DECLINED;
    }
    if (!vba::dense_family(options->linear_solver)) {
        int32_t st = vba::validate_matrix_free(prepared->m, *options, &msg);
        if (st != VBA_OK) {
            ctx->last_error = msg;
            return st;
        }
    }
    try {
        return vba::run(*ctx, *prepared, *options, stats, stats_capacity, result);
    } catch (const std::bad_alloc&) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
}

int32_t vba_prepared_state(vba_context* ctx, const vba_prepared* prepared, double* poses_out,
                           double* landmarks_out) {
    if (!ctx || !prepared) return VBA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    return vba::download_state(*ctx, *prepared, poses_out, landmarks_out);
}

void vba_prepared_free(vba_context* ctx, vba_prepared* prepared) {
    if (!prepared) return;
    if (ctx) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipStreamSynchronize(ctx->stream);
        vba::release(*prepared);
    } else {
        vba::release(*prepared);
    }
    delete prepared;
}

namespace {

// Validate every window (vba_check, then the model build) and lay out the
// valid ones as one batch.
int32_t build_batch(vba_context* ctx, uint32_t count, const vba_problem* problems, const vba_options& options,
                    vba_prepared_batch* pb, vba_result* results, int32_t* statuses) {
    if (options.linear_solver != VBA_SOLVER_DENSE) {
        ctx->last_error = "batched windows use the dense reduced solve (linear_solver = VBA_SOLVER_DENSE)";
        return VBA_ERR_INVALID_CONFIGURATION;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->last_error = "hipSetDevice failed";
        return VBA_ERR_DEVICE;
    }
    std::vector<vba::Model> models;
    std::vector<int32_t> valid(count, 0);
    try {
        models.resize(count);
        for (uint32_t i = 0; i < count; ++i) {
            if (results) std::memset(&results[i], 0, sizeof(vba_result));
            std::string reason;
            int32_t st = vba::check_support(problems[i], options, &reason);
            if (st == VBA_OK) {
                uint64_t error_id = 0;
                st = vba::build_model(problems[i], &models[i], &error_id, &reason);
                if (results) results[i].error_id = error_id;
            }
            statuses[i] = st;
            valid[i] = st == VBA_OK;
        }
        std::string msg;
        int32_t st = vba::prepare_batch(*ctx, models, valid, pb, statuses, &msg);
        if (st != VBA_OK) ctx->last_error = msg;
        return st;
    } catch (const std::bad_alloc&) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
}

}  // namespace

int32_t vba_optimize_batch(vba_context* ctx, uint32_t count, const vba_problem* problems, const vba_options* options,
                           double* const* poses_out, double* const* landmarks_out, vba_iteration_stats* stats,
                           uint32_t stats_capacity, vba_result* results, int32_t* statuses) {
    if (!ctx || !options || !results || !statuses || (count && !problems)) {
        vba::set_thread_error("vba_optimize_batch: null argument");
        if (ctx) ctx->last_error = "vba_optimize_batch: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->last_error.clear();
    vba_prepared_batch pb;
    int32_t st = build_batch(ctx, count, problems, *options, &pb, results, statuses);
    if (st == VBA_OK) {
        try {
            st = vba::run_batch(*ctx, pb, *options, stats, stats_capacity, results, statuses);
        } catch (const std::bad_alloc&) {
            ctx->last_error = "out of host memory";
            st = VBA_ERR_OUT_OF_MEMORY;
        }
    }
    if (st == VBA_OK)
        for (uint32_t i = 0; i < count && st == VBA_OK; ++i)
            if (pb.window_of[i] >= 0)
                st = vba::download_batch_window(*ctx, pb, i, poses_out ? poses_out[i] : nullptr,
                                                landmarks_out ? landmarks_out[i] : nullptr);
    vba::release(pb);
    return st;
}

int32_t vba_prepare_batch(vba_context* ctx, uint32_t count, const vba_problem* problems, const vba_options* options,
                          vba_prepared_batch** out, vba_result* results, int32_t* statuses) {
    if (!ctx || !options || !out || !statuses || (count && !problems)) {
        vba::set_thread_error("vba_prepare_batch: null argument");
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = nullptr;
    ctx->last_error.clear();
    std::unique_ptr<vba_prepared_batch> pb(new (std::nothrow) vba_prepared_batch());
    if (!pb) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
    int32_t st = build_batch(ctx, count, problems, *options, pb.get(), results, statuses);
    if (st != VBA_OK) {
        vba::release(*pb);
        return st;
    }
    *out = pb.release();
    return VBA_OK;
}

int32_t vba_solve_prepared_batch(vba_context* ctx, vba_prepared_batch* batch, const vba_options* options,
                                 vba_iteration_stats* stats, uint32_t stats_capacity, vba_result* results,
                                 int32_t* statuses) {
    if (!ctx || !batch || !options || !results || !statuses) {
        vba::set_thread_error("vba_solve_prepared_batch: null argument");
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->last_error.clear();
    std::string reason;
    if (vba::check_options(*options, &reason) != VBA_OK) {
        ctx->last_error = "declined: " + reason;
        return VBA_DECLINED;
    }
    if (options->linear_solver != VBA_SOLVER_DENSE) {
        ctx->last_error = "batched windows use the dense reduced solve (linear_solver = VBA_SOLVER_DENSE)";
        return VBA_ERR_INVALID_CONFIGURATION;
    }
    try {
        return vba::run_batch(*ctx, *batch, *options, stats, stats_capacity, results, statuses, true);
    } catch (const std::bad_alloc&) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
}

int32_t vba_prepared_batch_state(vba_context* ctx, const vba_prepared_batch* batch, uint32_t window,
                                 double* poses_out, double* landmarks_out) {
    if (!ctx || !batch) return VBA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    return vba::download_batch_window(*ctx, *batch, window, poses_out, landmarks_out);
}

void vba_prepared_batch_free(vba_context* ctx, vba_prepared_batch* batch) {
    if (!batch) return;
    if (ctx) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipStreamSynchronize(ctx->stream);
    }
    vba::release(*batch);
    delete batch;
}

void vba_profile_enable(vba_context* ctx, int32_t enable) {
    if (!ctx) return;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->profile_stride = enable > 0 ? enable : 0;
    ctx->profile_seen.assign(ctx->profile_seen.size(), 0);
}

void vba_profile_reset(vba_context* ctx) {
    if (!ctx) return;
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (vba::KernelProfile& k : ctx->kernels) {
        k.launches = 0;
        k.seconds = 0.0;
        k.bytes = 0.0;
    }
}

int32_t vba_dense_solve(vba_context* ctx, uint32_t n, const double* h, const double* b, double* x) {
    if (!ctx || (n && (!h || !b || !x))) {
        vba::set_thread_error("vba_dense_solve: null argument");
        if (ctx) ctx->last_error = "vba_dense_solve: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->last_error.clear();
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->last_error = "hipSetDevice failed";
        return VBA_ERR_DEVICE;
    }
    try {
        return vba::dense_solve(*ctx, n, h, b, x, &ctx->last_error);
    } catch (const std::bad_alloc&) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
}

int32_t vba_dense_solve_noalloc(vba_context* ctx, uint32_t n, const double* h, const double* b, double* x) {
    if (!ctx || (n && (!h || !b || !x))) {
        vba::set_thread_error("vba_dense_solve_noalloc: null argument");
        if (ctx) ctx->last_error = "vba_dense_solve_noalloc: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->last_error.clear();
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->last_error = "hipSetDevice failed";
        return VBA_ERR_DEVICE;
    }
    return vba::dense_solve_noalloc(*ctx, n, h, b, x, &ctx->last_error);
}

int32_t vba_reduced_solve(vba_context* ctx, const vba_reduced_system* system, double* step, uint32_t* iterations) {
    if (!ctx || !system || !step || !iterations) {
        vba::set_thread_error("vba_reduced_solve: null argument");
        if (ctx) ctx->last_error = "vba_reduced_solve: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->last_error.clear();
    *iterations = 0;
    try {
        return vba::reduced_solve(*ctx, *system, step, iterations, &ctx->last_error);
    } catch (const std::bad_alloc&) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
}

int32_t vba_reduced_eliminate(vba_context* ctx, const vba_point_elimination* elimination, double* run_blocks,
                              double* coarse_sums) {
    if (!ctx || !elimination) {
        vba::set_thread_error("vba_reduced_eliminate: null argument");
        if (ctx) ctx->last_error = "vba_reduced_eliminate: null argument";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->last_error.clear();
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->last_error = "hipSetDevice failed";
        return VBA_ERR_DEVICE;
    }
    try {
        return vba::reduced_eliminate(*ctx, *elimination, run_blocks, coarse_sums, &ctx->last_error);
    } catch (const std::bad_alloc&) {
        ctx->last_error = "out of host memory";
        return VBA_ERR_OUT_OF_MEMORY;
    }
}

int32_t vba_profile_query(vba_context* ctx, const char* kernel, uint64_t* launches, double* seconds,
                          double* bytes) {
    if (!ctx || !kernel) return VBA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (const vba::KernelProfile& k : ctx->kernels)
        if (k.name == kernel) {
            if (launches) *launches = k.launches;
            if (seconds) *seconds = k.seconds;
            if (bytes) *bytes = k.bytes;
            return VBA_OK;
        }
    if (launches) *launches = 0;
    if (seconds) *seconds = 0.0;
    if (bytes) *bytes = 0.0;
    return VBA_OK;
}

int32_t vba_set_tuning(vba_context* ctx, const vba_tuning* tuning) {
    if (!ctx) return VBA_ERR_INVALID_ARGUMENT;
    vba_tuning t{10, 0, 0, 0};
    if (tuning) {
        if (tuning->pcg_chunk < 1 || tuning->pcg_chunk > 1000 || tuning->pcg_lookahead > 2 ||
            tuning->apply_waves_per_cu > 32) {
            std::lock_guard<std::mutex> lock(ctx->mu);
            ctx->last_error =
                "vba_set_tuning: pcg_chunk must be 1..1000, pcg_lookahead 0..2 and apply_waves_per_cu 0..32";
            return VBA_ERR_INVALID_ARGUMENT;
        }
        t = *tuning;
        t.pcg_predict = t.pcg_predict ? 1u : 0u;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tuning = t;
    return VBA_OK;
}

int32_t vba_get_tuning(const vba_context* ctx, vba_tuning* tuning) {
    if (!ctx || !tuning) return VBA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(const_cast<vba_context*>(ctx)->mu);
    *tuning = ctx->tuning;
    return VBA_OK;
}

// The candidate schedules: the reference's chunk of 10 alone (the default)
// and with prediction and lookahead, and shorter and longer chunks.
int32_t vba_autotune(vba_context* ctx, vba_prepared* prepared, const vba_options* options, uint32_t lm_iterations,
                     vba_tuning* best, double* best_seconds, double* default_seconds) {
    if (!ctx || !prepared || !options) return VBA_ERR_INVALID_ARGUMENT;
    {
        std::string reason;
        if (vba::check_options(*options, &reason) != VBA_OK) {
            std::lock_guard<std::mutex> lock(ctx->mu);
            ctx->last_error = "declined: " + reason;
            return VBA_DECLINED;
        }
    }
    if (vba::dense_family(options->linear_solver)) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->last_error = "vba_autotune: the schedule applies to the PCG solvers";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    vba_options o = *options;
    if (lm_iterations) o.max_iterations = lm_iterations;
    // launch schedules, each with the one-wave-per-tile product and with the
    // persistent one at its residency (16 waves per compute unit)
    static const vba_tuning kCandidates[] = {
        {10, 0, 0, 0},  {10, 1, 0, 0},  {10, 1, 1, 0},  {5, 1, 1, 0},   {5, 2, 1, 0},  {20, 1, 1, 0},
        {20, 0, 1, 0},  {4, 2, 1, 0},   {10, 0, 0, 16}, {10, 1, 1, 16}, {5, 1, 1, 16}, {20, 1, 1, 16},
        {20, 0, 1, 16}, {4, 2, 1, 16},
    };
    const uint32_t cap = o.max_iterations ? o.max_iterations : 1;
    std::vector<vba_iteration_stats> ref_stats(cap), stats(cap);
    vba_result ref{}, res{};
    double t_best = 0.0, t_default = 0.0;
    vba_tuning chosen = kCandidates[0];
    const vba_tuning saved = ctx->tuning;
    auto same = [&](const vba_result& a, const std::vector<vba_iteration_stats>& sa, const vba_result& b,
                    const std::vector<vba_iteration_stats>& sb) {
        if (std::memcmp(&a.final_cost, &b.final_cost, sizeof(double)) || a.num_iterations != b.num_iterations ||
            a.total_pcg_iterations != b.total_pcg_iterations)
            return false;
        for (uint32_t i = 0; i < std::min(a.num_iterations, cap); ++i)
            if (sa[i].step_accepted != sb[i].step_accepted || sa[i].pcg_iterations != sb[i].pcg_iterations ||
                std::memcmp(&sa[i].cost_after, &sb[i].cost_after, sizeof(double)))
                return false;
        return true;
    };
    for (size_t c = 0; c < sizeof(kCandidates) / sizeof(kCandidates[0]); ++c) {
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            ctx->tuning = kCandidates[c];
        }
        double t = 0.0;
        for (int rep = 0; rep < 3; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            const int32_t st = vba_solve_prepared(ctx, prepared, &o, stats.data(), cap, &res);
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (st != VBA_OK) {
                std::lock_guard<std::mutex> lock(ctx->mu);
                ctx->tuning = saved;
                return st;
            }
            if (c == 0 && rep == 0) {
                ref = res;
                ref_stats = stats;
            } else if (!same(ref, ref_stats, res, stats)) {
                std::lock_guard<std::mutex> lock(ctx->mu);
                ctx->tuning = saved;
                ctx->last_error = "vba_autotune: a PCG schedule changed the result";
                return VBA_ERR_DEVICE;
            }
            t = rep == 0 ? dt : std::min(t, dt);
        }
        if (c == 0) t_default = t;
        if (c == 0 || t < t_best) {
            t_best = t;
            chosen = kCandidates[c];
        }
    }
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->tuning = chosen;
    }
    if (best) *best = chosen;
    if (best_seconds) *best_seconds = t_best;
    if (default_seconds) *default_seconds = t_default;
    return VBA_OK;
}

}  // extern "C"
