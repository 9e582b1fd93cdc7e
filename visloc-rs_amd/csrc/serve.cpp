// serve.cpp — a request loop over the solver: continuous batching of
// independent bundle-adjustment solves on one GPU.
//
// Requests are copied at submission and queued by lane:
//   - window lane: dense-solver requests with at most 256 variable poses
//     (pose-only refinements, local windows). A lane thread takes the oldest
//     request, gathers every queued request with the same options (up to
//     max_window_batch, waiting at most batch_wait_seconds for more) and
//     solves them as one device batch (prepare_batch / run_batch), where
//     each window keeps its own LM state on the device;
//   - submap lane: everything else (matrix-free solves, large dense ones),
//     one problem per lane thread through the single-problem driver.
// Every lane thread owns a context (its own HIP stream), so batches and
// submaps admitted while others run overlap on the device. Admission waits
// while the estimated device footprint of the running work plus the next
// request would exceed hbm_budget_bytes. A window's result does not depend
// on what else is in its batch (no arithmetic crosses windows), so every
// result equals the one the request gets alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/vba.h"
#include "capi_types.hpp"
#include "host_common.hpp"
#include "problem.hpp"
#include "solver.hpp"

namespace {

using Clock = std::chrono::steady_clock;

// A deep copy of a vba_problem and a view over it.
struct OwnedProblem {
    vba_problem view{};  // over the copies below, or over the caller's arrays when borrowed
    std::vector<uint64_t> pose_ids, lm_ids, obs_pose, obs_lm, rig_pose, rig_lm;
    std::vector<double> poses, lms, obs_xy, rig_xy;
    std::vector<uint8_t> pose_fixed, lm_fixed;
    std::vector<uint32_t> rig_sensor;
    std::vector<vba_sensor> sensors;
    vba_stereo_terms stereo{};
    std::vector<uint64_t> st_pose, st_lm, ge_pose, ge_lm;
    std::vector<double> st_obs, ge_obs;
    std::vector<uint32_t> ge_pair;
    std::vector<vba_sensor> pairs;

    template <typename T>
    static void take(std::vector<T>& dst, const T* src, uint64_t n) {
        if (src && n) dst.assign(src, src + n);
        else dst.clear();
    }
    void copy_from(const vba_problem& p) {
        take(pose_ids, p.pose_ids, p.num_poses);
        take(poses, p.poses, 7 * p.num_poses);
        take(pose_fixed, p.pose_fixed, p.num_poses);
        take(lm_ids, p.landmark_ids, p.num_landmarks);
        take(lms, p.landmarks, 3 * p.num_landmarks);
        take(lm_fixed, p.landmark_fixed, p.num_landmarks);
        take(obs_pose, p.obs_pose_ids, p.num_observations);
        take(obs_lm, p.obs_landmark_ids, p.num_observations);
        take(obs_xy, p.obs_xy, 2 * p.num_observations);
        take(rig_pose, p.rig_pose_ids, p.num_rig_observations);
        take(rig_lm, p.rig_landmark_ids, p.num_rig_observations);
        take(rig_xy, p.rig_xy, 2 * p.num_rig_observations);
        take(rig_sensor, p.rig_sensor, p.num_rig_observations);
        take(sensors, p.sensors, p.num_sensors);
        view = p;
        auto ptr = [](auto& v) { return v.empty() ? nullptr : v.data(); };
        view.pose_ids = ptr(pose_ids);
        view.poses = ptr(poses);
        view.pose_fixed = p.pose_fixed ? ptr(pose_fixed) : nullptr;
        view.landmark_ids = ptr(lm_ids);
        view.landmarks = ptr(lms);
        view.landmark_fixed = p.landmark_fixed ? ptr(lm_fixed) : nullptr;
        view.obs_pose_ids = ptr(obs_pose);
        view.obs_landmark_ids = ptr(obs_lm);
        view.obs_xy = ptr(obs_xy);
        view.rig_pose_ids = ptr(rig_pose);
        view.rig_landmark_ids = ptr(rig_lm);
        view.rig_xy = ptr(rig_xy);
        view.rig_sensor = ptr(rig_sensor);
        view.sensors = ptr(sensors);
        if (p.stereo) {
            const vba_stereo_terms& t = *p.stereo;
            take(st_pose, t.stereo_pose_ids, t.num_stereo);
            take(st_lm, t.stereo_landmark_ids, t.num_stereo);
            take(st_obs, t.stereo_obs, 3 * t.num_stereo);
            take(ge_pose, t.general_pose_ids, t.num_general);
            take(ge_lm, t.general_landmark_ids, t.num_general);
            take(ge_obs, t.general_obs, 4 * t.num_general);
            take(ge_pair, t.general_pair, t.num_general);
            take(pairs, t.pairs, t.num_pairs);
            stereo = t;
            stereo.stereo_pose_ids = ptr(st_pose);
            stereo.stereo_landmark_ids = ptr(st_lm);
            stereo.stereo_obs = ptr(st_obs);
            stereo.general_pose_ids = ptr(ge_pose);
            stereo.general_landmark_ids = ptr(ge_lm);
            stereo.general_obs = ptr(ge_obs);
            stereo.general_pair = ptr(ge_pair);
            stereo.pairs = ptr(pairs);
            view.stereo = &stereo;
        }
    }
};

struct Request {
    uint64_t ticket = 0;
    OwnedProblem prob;
    vba_options opts{};
    Clock::time_point submitted, finished;
    uint64_t est_bytes = 0;
    bool window_lane = false;
    bool done = false;
    int32_t status = VBA_OK;
    vba_result result{};
    std::vector<double> poses_out, lms_out;
};

// Device bytes a request holds while it runs (structure, working buffers
// and, for a window, its dense system), an upper estimate.
uint64_t estimate_bytes(const vba_problem& p, bool window) {
    const uint64_t n = vba::observation_count(p);
    const uint64_t L = p.num_landmarks, P = p.num_poses;
    uint64_t b = n * 40 + n * 196 + L * 260 + P * 1400 + (64u << 10);
    if (window) {
        uint64_t var = 0;
        for (uint64_t i = 0; i < P; ++i) var += p.pose_fixed ? (p.pose_fixed[i] == 0) : 1;
        b += 2 * (6 * var) * (6 * var) * sizeof(double) + n * 24;
    }
    return b;
}

}  // namespace

struct vba_server {
    vba_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> window_q, submap_q;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
    uint64_t next_ticket = 1;
    uint64_t hbm_in_use = 0;
    uint64_t completed = 0, batches = 0, batched_windows = 0;
    bool stopping = false;
    std::string error;
    std::vector<std::thread> threads;
    std::vector<vba_context*> contexts;

    // `bytes` more of resident work stays within the budget.
    bool fits(uint64_t bytes) const { return hbm_in_use + bytes <= cfg.hbm_budget_bytes; }
    // A lane may start a request of `bytes`: within the budget, or alone on an
    // idle device (a request larger than the budget still runs, by itself).
    bool admits(uint64_t bytes) const { return hbm_in_use == 0 || fits(bytes); }

    void finish(const std::shared_ptr<Request>& r) {
        r->finished = Clock::now();
        r->done = true;
        ++completed;
    }

    // Window lane: batches of dense windows with equal options.
    void window_lane(vba_context* ctx) {
        for (;;) {
            std::vector<std::shared_ptr<Request>> batch;
            uint64_t bytes = 0;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv_work.wait(lock, [&] { return stopping || (!window_q.empty() && admits(window_q.front()->est_bytes)); });
                if (stopping && window_q.empty()) return;
                if (window_q.size() < cfg.max_window_batch && cfg.batch_wait_seconds > 0.0) {
                    const auto until =
                        Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double>(cfg.batch_wait_seconds));
                    cv_work.wait_until(lock, until,
                                       [&] { return stopping || window_q.size() >= cfg.max_window_batch; });
                    // the lock was released: another lane may have taken the
                    // queue's head or admitted work meanwhile
                    if (window_q.empty() || !admits(window_q.front()->est_bytes)) continue;
                }
                // the head is admitted as admits() allows; every later window
                // only within the budget, counting the batch gathered so far
                const vba_options key = window_q.front()->opts;
                for (auto it = window_q.begin(); it != window_q.end() && batch.size() < cfg.max_window_batch;) {
                    if (std::memcmp(&(*it)->opts, &key, sizeof(key)) == 0 &&
                        (batch.empty() ? admits((*it)->est_bytes) : fits(bytes + (*it)->est_bytes))) {
                        bytes += (*it)->est_bytes;
                        batch.push_back(*it);
                        it = window_q.erase(it);
                    } else {
                        ++it;
                    }
                }
                if (batch.empty()) continue;
                hbm_in_use += bytes;
            }
            solve_windows(ctx, batch);
            {
                std::lock_guard<std::mutex> lock(mu);
                hbm_in_use -= bytes;
                for (auto& r : batch) finish(r);
                ++batches;
                batched_windows += batch.size();
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    void solve_windows(vba_context* ctx, std::vector<std::shared_ptr<Request>>& batch) {
        const uint32_t n = (uint32_t)batch.size();
        std::vector<vba::Model> models(n);
        std::vector<int32_t> valid(n, 0), statuses(n, VBA_OK);
        std::vector<vba_result> results(n);
        std::vector<vba_iteration_stats> stats;
        vba::PreparedBatch pb;
        try {
            // host layout of the windows, on up to cfg.build_threads threads
            auto build = [&](uint32_t begin, uint32_t step) {
                for (uint32_t i = begin; i < n; i += step) {
                    std::memset(&results[i], 0, sizeof(vba_result));
                    std::string reason;
                    int32_t st = vba::check_support(batch[i]->prob.view, batch[i]->opts, &reason);
                    if (st == VBA_OK) {
                        uint64_t error_id = 0;
                        try {
                            st = vba::build_model(batch[i]->prob.view, &models[i], &error_id, &reason);
                        } catch (const std::bad_alloc&) {
                            st = VBA_ERR_OUT_OF_MEMORY;
                        }
                        results[i].error_id = error_id;
                    }
                    statuses[i] = st;
                    valid[i] = st == VBA_OK;
                }
            };
            const uint32_t nt = std::max<uint32_t>(1, std::min<uint32_t>(cfg.build_threads, n));
            std::vector<std::thread> pool;
            for (uint32_t t = 1; t < nt; ++t) pool.emplace_back(build, t, nt);
            build(0, nt);
            for (std::thread& th : pool) th.join();
            std::lock_guard<std::mutex> lock(ctx->mu);
            (void)hipSetDevice(ctx->device);
            std::string msg;
            int32_t st = vba::prepare_batch(*ctx, models, valid, &pb, statuses.data(), &msg);
            models.clear();
            const uint32_t cap = std::max<uint32_t>(batch[0]->opts.max_iterations, 1);
            stats.resize((size_t)n * cap);
            if (st == VBA_OK)
                st = vba::run_batch(*ctx, pb, batch[0]->opts, stats.data(), cap, results.data(), statuses.data());
            for (uint32_t i = 0; i < n; ++i) {
                Request& r = *batch[i];
                if (st != VBA_OK) {
                    r.status = st;
                    continue;
                }
                r.status = statuses[i];
                r.result = results[i];
                if (r.status == VBA_OK && pb.window_of[i] >= 0) {
                    r.poses_out.resize(7 * r.prob.view.num_poses);
                    r.lms_out.resize(3 * r.prob.view.num_landmarks);
                    r.status = vba::download_batch_window(*ctx, pb, i, r.poses_out.data(), r.lms_out.data());
                }
            }
            vba::release(pb);
        } catch (const std::bad_alloc&) {
            vba::release(pb);
            for (auto& r : batch) r->status = VBA_ERR_OUT_OF_MEMORY;
        }
    }

    // Submap lane: one problem at a time through the single-problem driver.
    void submap_lane(vba_context* ctx) {
        for (;;) {
            std::shared_ptr<Request> r;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv_work.wait(lock, [&] { return stopping || (!submap_q.empty() && admits(submap_q.front()->est_bytes)); });
                if (stopping && submap_q.empty()) return;
                r = submap_q.front();
                submap_q.pop_front();
                hbm_in_use += r->est_bytes;
            }
            r->poses_out.resize(7 * r->prob.view.num_poses);
            r->lms_out.resize(3 * r->prob.view.num_landmarks);
            std::vector<vba_iteration_stats> stats(std::max<uint32_t>(r->opts.max_iterations, 1));
            r->status = vba_optimize(ctx, &r->prob.view, &r->opts, r->poses_out.data(), r->lms_out.data(),
                                     stats.data(), (uint32_t)stats.size(), &r->result);
            {
                std::lock_guard<std::mutex> lock(mu);
                hbm_in_use -= r->est_bytes;
                finish(r);
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

void vba_server_default_config(vba_server_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->device = 0;
    c->max_window_batch = 256;
    c->window_streams = 2;
    c->submap_streams = 6;
    c->hbm_budget_bytes = 200ull << 30;
    c->batch_wait_seconds = 0.002;
    c->build_threads = 8;
}

vba_server* vba_server_create(const vba_server_config* config) {
    vba_server_config c;
    vba_server_default_config(&c);
    if (config) c = *config;
    if (c.max_window_batch == 0 || (c.window_streams == 0 && c.submap_streams == 0) ||
        c.window_streams + c.submap_streams > 16 || !(c.batch_wait_seconds >= 0.0) || c.build_threads > 64) {
        vba::set_thread_error("vba_server_create: invalid configuration");
        return nullptr;
    }
    std::unique_ptr<vba_server> s(new (std::nothrow) vba_server());
    if (!s) {
        vba::set_thread_error("vba_server_create: out of host memory");
        return nullptr;
    }
    s->cfg = c;
    for (uint32_t i = 0; i < c.window_streams + c.submap_streams; ++i) {
        vba_context* ctx = vba_create(c.device);
        if (!ctx) {
            for (vba_context* x : s->contexts) vba_destroy(x);
            return nullptr;
        }
        s->contexts.push_back(ctx);
    }
    vba_server* raw = s.get();
    for (uint32_t i = 0; i < c.window_streams; ++i)
        s->threads.emplace_back([raw, i] { raw->window_lane(raw->contexts[i]); });
    for (uint32_t i = 0; i < c.submap_streams; ++i)
        s->threads.emplace_back([raw, i] { raw->submap_lane(raw->contexts[raw->cfg.window_streams + i]); });
    return s.release();
}

namespace {

// Dense solves of at most 256 variable poses go to the window lane.
bool window_lane(const vba_server& s, const vba_problem& problem, const vba_options& options) {
    uint64_t var = 0;
    for (uint64_t i = 0; i < problem.num_poses; ++i) var += problem.pose_fixed ? (problem.pose_fixed[i] == 0) : 1;
    return options.linear_solver == VBA_SOLVER_DENSE && var <= 256 && s.cfg.window_streams > 0;
}

uint64_t submit(vba_server* s, const vba_problem* problem, const vba_options* options, bool borrow) {
    if (!s || !problem || !options) {
        vba::set_thread_error("vba_server_submit: null argument");
        return 0;
    }
    auto r = std::make_shared<Request>();
    try {
        if (borrow) r->prob.view = *problem;
        else r->prob.copy_from(*problem);
    } catch (const std::bad_alloc&) {
        vba::set_thread_error("vba_server_submit: out of host memory");
        return 0;
    }
    r->opts = *options;
    r->window_lane = window_lane(*s, *problem, *options);
    if (!r->window_lane && s->cfg.submap_streams == 0) {
        vba::set_thread_error("vba_server_submit: no lane serves this request");
        return 0;
    }
    r->est_bytes = estimate_bytes(*problem, r->window_lane);
    {
        std::lock_guard<std::mutex> lock(s->mu);
        if (s->stopping) {
            vba::set_thread_error("vba_server_submit: server is shutting down");
            return 0;
        }
        r->ticket = s->next_ticket++;
        r->submitted = Clock::now();
        s->requests[r->ticket] = r;
        (r->window_lane ? s->window_q : s->submap_q).push_back(r);
    }
    s->cv_work.notify_all();
    return r->ticket;
}

}  // namespace

uint64_t vba_server_submit(vba_server* s, const vba_problem* problem, const vba_options* options) {
    return submit(s, problem, options, false);
}

uint64_t vba_server_submit_borrowed(vba_server* s, const vba_problem* problem, const vba_options* options) {
    return submit(s, problem, options, true);
}

int32_t vba_server_wait(vba_server* s, uint64_t ticket, double timeout_seconds, double* poses_out,
                        double* landmarks_out, vba_result* result, double* latency_seconds) {
    if (!s) return VBA_ERR_INVALID_ARGUMENT;
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lock(s->mu);
        auto it = s->requests.find(ticket);
        if (it == s->requests.end()) {
            vba::set_thread_error("vba_server_wait: unknown ticket");
            return VBA_ERR_INVALID_ARGUMENT;
        }
        r = it->second;
        const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                              std::chrono::duration<double>(std::max(0.0, timeout_seconds)));
        if (!s->cv_done.wait_until(lock, until, [&] { return r->done; })) return VBA_PENDING;
        s->requests.erase(it);
    }
    if (result) *result = r->result;
    if (latency_seconds) *latency_seconds = std::chrono::duration<double>(r->finished - r->submitted).count();
    if (r->status == VBA_OK) {
        if (poses_out) std::memcpy(poses_out, r->poses_out.data(), r->poses_out.size() * sizeof(double));
        if (landmarks_out) std::memcpy(landmarks_out, r->lms_out.data(), r->lms_out.size() * sizeof(double));
    }
    return r->status;
}

uint64_t vba_server_request_bytes(const vba_server* s, const vba_problem* problem, const vba_options* options) {
    if (!s || !problem || !options) return 0;
    return estimate_bytes(*problem, window_lane(*s, *problem, *options));
}

void vba_server_counters(vba_server* s, uint64_t* completed, uint64_t* batches, uint64_t* batched_windows,
                         uint64_t* queued) {
    if (!s) return;
    std::lock_guard<std::mutex> lock(s->mu);
    if (completed) *completed = s->completed;
    if (batches) *batches = s->batches;
    if (batched_windows) *batched_windows = s->batched_windows;
    if (queued) *queued = s->window_q.size() + s->submap_q.size();
}

void vba_server_destroy(vba_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (std::thread& t : s->threads) t.join();
    for (vba_context* c : s->contexts) vba_destroy(c);
    delete s;
}

}  // extern "C"
