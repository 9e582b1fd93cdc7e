// solver.cpp — device-resident Levenberg–Marquardt loop.
//
// Control flow and acceptance rules follow optimize_weighted_backend
// (pipelines/slam/src/bundle/lm.rs:162-1017): strict cost decrease plus a
// non-increasing non-projectable count, lambda x10 (capped) on reject or on a
// failed linear solve, x0.1 (floored) on accept, and the step / cost /
// relative-cost convergence tests. The state stays on the device; the host
// reads back a few scalars per iteration (cost, count, step norms, PCG
// status).
#include "solver.hpp"

#include <algorithm>
#include <optional>
#include <cmath>
#include <cstring>
#include <limits>
#include <unordered_map>

#include "host_common.hpp"
#include "reduced.hpp"

namespace vba {

namespace {

enum BufId {
    B_POSES0, B_LMS0, B_POSES, B_LMS, B_POSES_T, B_LMS_T, B_SENSORS, B_POSE_VAR, B_LM_VAR, B_OBS_UV,
    B_OBS_SENSOR, B_EDGE_LM, B_EDGE_OBS, B_PERM_CL, B_L_POSE, B_LM_STORE, B_LM_EDGE_PTR, B_CAM_TILES,
    B_CAM_TILE_PTR, B_LM_TILES, B_TILE_PERM, B_BIG_TILES, B_POSE_RUN_PTR, B_RUN_SLOT, B_S_LM, B_GL64, B_GL32,
    B_HL64, B_TILE27, B_U, B_G, B_V, B_GL, B_VINV64, B_VINV32, B_VM, B_MINV, B_VB, B_VX, B_VR, B_VZ, B_VD,
    B_VQ, B_PART64, B_PART21, B_BPART, B_CPART, B_PCG, B_DELTA_L, B_TILE_SCALAR,
    B_SCALARS, B_DENSE_BLOCK_PQ, B_DENSE_BLOCK_PTR, B_DENSE_TERMS, B_DENSE_VAR_SLOT, B_DENSE_S,
    B_DENSE_WORK, B_DENSE_STATUS, B_X_BLOCK_POS, B_X_ROW_PTR, B_X_COL, B_X_DIAG_POS, B_X_S64, B_X_S32, B_VD2,
    B_G32_POS, B_OBS_W, B_OBS_S, B_OBS_KIND, B_OBS_EXTRA, B_SENSORS_T, B_J_E, B_J_TILE, B_J_HPK, B_J_KK, B_J_WK,
    B_J_PART, B_J_RHS, B_J_DELTA, B_J_GL, B_T_LM, B_T_HEAD, B_T_RPOSE, B_T_RSLOT, B_COUNT
};

// Profiled kernel families.
enum ProfId { K_LINEARIZE, K_SCHUR_APPLY, K_PCG_POSE, K_PCG_STEP, K_COST, K_EXPLICIT_ASSEMBLE, K_EXPLICIT_APPLY,
              K_DENSE_ASSEMBLE, K_DENSE_SOLVE, K_BATCH_DENSE_ASSEMBLE, K_BATCH_DENSE_SOLVE, K_COUNT };
// dense_solve / batch_dense_solve record the factor's FLOPs, n^3 / 3 per
// system, in the bytes slot (their bound is FP64 MFMA, not HBM)
const char* kProfNames[K_COUNT] = {"linearize",      "schur_apply",          "pcg_pose",         "pcg_step",
                                   "cost",           "explicit_assemble",    "explicit_apply",   "dense_assemble",
                                   "dense_solve",    "batch_dense_assemble", "batch_dense_solve"};

struct Fail {
    int32_t code;
    std::string msg;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw Fail{e == hipErrorOutOfMemory ? VBA_ERR_OUT_OF_MEMORY : VBA_ERR_DEVICE,
                   std::string(what) + ": " + hipGetErrorString(e)};
}

void* ensure(Prepared& pr, int id, size_t bytes) {
    if (pr.bufs.size() < B_COUNT) pr.bufs.resize(B_COUNT);
    DevBuffer& b = pr.bufs[id];
    if (bytes == 0) bytes = 64;
    if (b.bytes < bytes) {
        if (b.ptr) hip_check(hipFree(b.ptr), "hipFree");
        b.ptr = nullptr;
        b.bytes = 0;
        hip_check(hipMalloc(&b.ptr, bytes), "hipMalloc");
        b.bytes = bytes;
        ++pr.alloc_gen;
    }
    return b.ptr;
}
template <typename T>
T* upload(Prepared& pr, int id, const std::vector<T>& v, hipStream_t s) {
    T* p = (T*)ensure(pr, id, v.size() * sizeof(T));
    if (!v.empty())
        hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    return p;
}
template <typename T>
T* buf(Prepared& pr, int id) {
    return (T*)pr.bufs[id].ptr;
}

// Dense reduced system structure: one block per (var p <= var q) sharing a
// landmark plus every diagonal block; each term is (landmark, e_p, e_q).
struct DenseLists {
    std::vector<int32_t> block_pq, block_ptr, terms, var_slot;
};
void build_dense_lists(const Model& m, DenseLists* out) {
    const uint32_t P = m.P;
    out->var_slot.assign(P, 0);
    for (uint32_t s = 0; s < m.num_pose_slots; ++s)
        if (m.pose_var[s] >= 0) out->var_slot[m.pose_var[s]] = (int32_t)s;
    // block ids by (var p, var q): a P x P table while it stays small, else a hash map
    const bool table = (uint64_t)P * P <= (1ull << 24);
    std::vector<int32_t> block_id(table ? (size_t)P * P : 0, -1);
    std::unordered_map<uint64_t, int32_t> block_map;
    auto id_of = [&](int32_t a, int32_t b) -> int32_t& {
        if (table) return block_id[(size_t)a * P + b];
        return block_map.try_emplace(((uint64_t)(uint32_t)a << 32) | (uint32_t)b, -1).first->second;
    };
    std::vector<int64_t> count;
    out->block_pq.clear();
    auto block_of = [&](int32_t a, int32_t b) {
        int32_t& id = id_of(a, b);
        if (id < 0) {
            id = (int32_t)count.size();
            count.push_back(0);
            out->block_pq.push_back(out->var_slot[a]);
            out->block_pq.push_back(out->var_slot[b]);
        }
        return id;
    };
    for (uint32_t v = 0; v < P; ++v) block_of((int32_t)v, (int32_t)v);
    std::vector<std::pair<int32_t, int32_t>> ends;  // (var pose, l-order edge)
    auto collect = [&](uint32_t l) {
        ends.clear();
        for (int32_t e = m.lm_edge_ptr[l]; e < m.lm_edge_ptr[l + 1]; ++e)
            if (m.pose_var[m.l_pose[e]] >= 0) ends.push_back({m.pose_var[m.l_pose[e]], m.lm_store[e]});
    };
    for (uint32_t l = 0; l < m.num_lm_slots; ++l) {
        if (m.lm_var[l] < 0) continue;
        collect(l);
        for (size_t a = 0; a < ends.size(); ++a)
            for (size_t b = a; b < ends.size(); ++b) ++count[block_of(ends[a].first, ends[b].first)];
    }
    const size_t nb = count.size();
    int64_t total = 0;
    for (int64_t c : count) total += c;
    if (total > (int64_t)INT32_MAX / 3) throw Fail{VBA_ERR_INVALID_ARGUMENT, "Schur term lists too large"};
    out->block_ptr.assign(nb + 1, 0);
    for (size_t k = 0; k < nb; ++k) out->block_ptr[k + 1] = out->block_ptr[k] + (int32_t)count[k];
    out->terms.assign(3 * (size_t)total, 0);
    std::vector<int32_t> cursor(out->block_ptr.begin(), out->block_ptr.end() - 1);
    for (uint32_t l = 0; l < m.num_lm_slots; ++l) {
        if (m.lm_var[l] < 0) continue;
        collect(l);
        for (size_t a = 0; a < ends.size(); ++a)
            for (size_t b = a; b < ends.size(); ++b) {
                int32_t k = cursor[id_of(ends[a].first, ends[b].first)]++;
                out->terms[3 * (size_t)k] = (int32_t)l;
                out->terms[3 * (size_t)k + 1] = ends[a].second;
                out->terms[3 * (size_t)k + 2] = ends[b].second;
            }
    }
}
void build_dense(Prepared& pr) {
    DenseLists d;
    build_dense_lists(pr.m, &d);
    pr.dense_block_pq = std::move(d.block_pq);
    pr.dense_block_ptr = std::move(d.block_ptr);
    pr.dense_terms = std::move(d.terms);
    pr.dense_var_slot = std::move(d.var_slot);
    pr.dense_built = true;
}

// Full block rows of the explicit reduced camera matrix over the term lists'
// upper blocks: row p holds S_pq for every block (p, q) and S_qp^T for every
// block (q, p), columns ascending.
void build_explicit(Prepared& pr) {
    const Model& m = pr.m;
    const size_t nb = pr.dense_block_pq.size() / 2, NP = m.num_pose_slots;
    std::vector<int32_t> count(NP + 1, 0);
    for (size_t b = 0; b < nb; ++b) {
        const int32_t p = pr.dense_block_pq[2 * b], q = pr.dense_block_pq[2 * b + 1];
        ++count[p + 1];
        if (p != q) ++count[q + 1];
    }
    for (size_t p = 0; p < NP; ++p) count[p + 1] += count[p];
    pr.x_row_ptr = count;
    const size_t nfull = (size_t)count[NP];
    // (column, block * 2 + transposed) per full position, then each row by column
    std::vector<std::pair<int32_t, int32_t>> ent(nfull);
    std::vector<int32_t> cursor(count.begin(), count.end() - 1);
    for (size_t b = 0; b < nb; ++b) {
        const int32_t p = pr.dense_block_pq[2 * b], q = pr.dense_block_pq[2 * b + 1];
        ent[cursor[p]++] = {q, (int32_t)(2 * b)};
        if (p != q) ent[cursor[q]++] = {p, (int32_t)(2 * b + 1)};
    }
    pr.x_col.assign(nfull, 0);
    pr.x_block_pos.assign(2 * nb, -1);
    pr.x_diag_pos.assign(NP, -1);
    for (size_t p = 0; p < NP; ++p) {
        std::sort(ent.begin() + count[p], ent.begin() + count[p + 1]);
        for (int32_t t = count[p]; t < count[p + 1]; ++t) {
            pr.x_col[t] = ent[t].first;
            pr.x_block_pos[ent[t].second] = t;
            if (ent[t].first == (int32_t)p) pr.x_diag_pos[p] = t;
        }
    }
    pr.explicit_built = true;
}

// ---- profiling with HIP events on the context stream ----
hipEvent_t take_event(Context& ctx) {
    if (!ctx.event_pool.empty()) {
        hipEvent_t e = ctx.event_pool.back();
        ctx.event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}
// Times every profile_stride-th launch of a kernel. Stream mode records an
// event before and after the launch on the stream; bound mode (one kernel
// per scope, launched with events()) has the launch bind the pair to its
// dispatch, so the interval is the kernel's own.
struct ProfScope {
    Context& ctx;
    int kernel;
    double bytes;
    int32_t pcg_iter;
    bool bound;
    hipEvent_t a = nullptr, bb = nullptr;
    ProfScope(Context& c, int k, double b, int32_t it = -1, bool bind = false)
        : ctx(c), kernel(k), bytes(b), pcg_iter(it), bound(bind) {
        if (ctx.profile_stride > 0) {
            if (ctx.profile_seen.size() < K_COUNT) ctx.profile_seen.resize(K_COUNT, 0);
            if (ctx.profile_seen[kernel]++ % (uint64_t)ctx.profile_stride != 0) return;
            a = take_event(ctx);
            if (bound) {
                try {
                    bb = take_event(ctx);
                } catch (...) {
                    ctx.event_pool.push_back(a);
                    throw;
                }
            } else {
                hip_check(hipEventRecord(a, ctx.stream), "hipEventRecord");
            }
        }
    }
    // the events a bound scope's launch takes (none when not sampled)
    LaunchEvents events() const { return bound && a ? LaunchEvents{a, bb} : LaunchEvents{}; }
    ~ProfScope() {
        if (!a) return;
        if (bound) {
            ctx.pending.push_back({a, bb, kernel, bytes, pcg_iter});
            return;
        }
        hipEvent_t b = nullptr;
        if (!ctx.event_pool.empty()) {
            b = ctx.event_pool.back();
            ctx.event_pool.pop_back();
        } else if (hipEventCreate(&b) != hipSuccess) {
            ctx.event_pool.push_back(a);
            return;
        }
        if (hipEventRecord(b, ctx.stream) != hipSuccess) {
            ctx.event_pool.push_back(a);
            ctx.event_pool.push_back(b);
            return;
        }
        ctx.pending.push_back({a, b, kernel, bytes, pcg_iter});
    }
};
// Fold the first n pending event pairs, which must have completed (dropped
// ones may still be in flight: they are only returned to the pool), into the
// per-kernel totals. Returns n.
size_t harvest_upto(Context& ctx, size_t n) {
    if (ctx.kernels.size() < K_COUNT) {
        ctx.kernels.resize(K_COUNT);
        for (int k = 0; k < K_COUNT; ++k) ctx.kernels[k].name = kProfNames[k];
    }
    n = std::min(n, ctx.pending.size());
    for (size_t i = 0; i < n; ++i) {
        const Context::Pending& p = ctx.pending[i];
        if (p.kernel < 0) {  // a launch that exited early (PCG already finished)
            ctx.event_pool.push_back(p.a);
            ctx.event_pool.push_back(p.b);
            continue;
        }
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        KernelProfile& k = ctx.kernels[p.kernel];
        ++k.launches;
        k.seconds += ms * 1e-3;
        k.bytes += p.bytes;
        ctx.event_pool.push_back(p.a);
        ctx.event_pool.push_back(p.b);
    }
    ctx.pending.erase(ctx.pending.begin(), ctx.pending.begin() + (ptrdiff_t)n);
    return n;
}
// All of them (after a sync).
void harvest(Context& ctx) { harvest_upto(ctx, ctx.pending.size()); }

// Drop the pending PCG launches of iterations the device did not run (it
// had stopped; those launches exited at once).
void drop_after(Context& ctx, size_t from, int32_t iterations_run) {
    for (size_t i = from; i < ctx.pending.size(); ++i)
        if (ctx.pending[i].pcg_iter >= iterations_run) ctx.pending[i].kernel = -1;
}

}  // namespace

int32_t validate_matrix_free(const Model& m, const vba_options& o, std::string* msg) {
    // optimize.rs:382-460
    if (o.refine_intrinsics) {
        *msg = "intrinsics/distortion refinement is not supported";
        return VBA_ERR_INELIGIBLE;
    }
    const double il = o.initial_lambda;
    if (o.max_iterations == 0 || !o.has_initial_lambda || !(std::isfinite(il) && il > 0.0) ||
        !std::isfinite(o.lambda_increase_factor) || o.lambda_increase_factor <= 1.0 ||
        !std::isfinite(o.lambda_decrease_factor) || o.lambda_decrease_factor <= 0.0 ||
        o.lambda_decrease_factor >= 1.0 || !std::isfinite(o.max_lambda) || o.max_lambda < il ||
        !std::isfinite(o.min_lambda) || o.min_lambda <= 0.0 || o.min_lambda > il ||
        !std::isfinite(o.step_tolerance) || o.step_tolerance < 0.0 || !std::isfinite(o.cost_tolerance) ||
        o.cost_tolerance < 0.0 ||
        (o.has_relative_cost_tolerance &&
         (!std::isfinite(o.relative_cost_tolerance) || o.relative_cost_tolerance < 0.0))) {
        *msg = "LM damping/tolerance values are not finite or ordered";
        return VBA_ERR_INVALID_CONFIGURATION;
    }
    if (o.max_pcg_iterations == 0 || !std::isfinite(o.pcg_relative_tolerance) || o.pcg_relative_tolerance < 0.0 ||
        !std::isfinite(o.pcg_absolute_tolerance) || o.pcg_absolute_tolerance <= 0.0) {
        *msg = "PCG iteration and tolerance values are invalid";
        return VBA_ERR_INVALID_CONFIGURATION;
    }
    if (o.robust_kernel != VBA_KERNEL_NONE && !(std::isfinite(o.robust_scale) && o.robust_scale > 0.0)) {
        *msg = "robust-kernel scale must be finite and positive";
        return VBA_ERR_INVALID_CONFIGURATION;
    }
    auto finite_lens = [](const Lens<double>& L) {
        const double v[8] = {L.fx, L.fy, L.cx, L.cy, L.k1, L.k2, L.p1, L.p2};
        for (double x : v)
            if (!std::isfinite(x)) return false;
        return L.fx > 0.0 && L.fy > 0.0;
    };
    // every sensor an observation uses (one without intrinsics is refused
    // when the problem is built): lens, and the rig or stereo pair
    // extrinsic (optimize.rs:572-622)
    for (const DevSensor& s : m.sensors) {
        if (s.lens.kind < 0) continue;
        bool ok = finite_lens(s.lens);
        for (double x : s.R) ok = ok && std::isfinite(x);
        for (double x : s.t) ok = ok && std::isfinite(x);
        if (!ok) {
            *msg = "camera calibration is unsupported or non-finite";
            return VBA_ERR_INELIGIBLE;
        }
    }
    if (!m.any_fixed_pose) {
        *msg = "at least one existing pose must be fixed as a gauge anchor";
        return VBA_ERR_INELIGIBLE;
    }
    if (m.P == 0) {
        *msg = "matrix-free backend requires a variable pose";
        return VBA_ERR_INELIGIBLE;
    }
    for (double v : m.poses)
        if (!std::isfinite(v)) {
            *msg = "pose or landmark state is non-finite";
            return VBA_ERR_INELIGIBLE;
        }
    for (double v : m.lms)
        if (!std::isfinite(v)) {
            *msg = "pose or landmark state is non-finite";
            return VBA_ERR_INELIGIBLE;
        }
    for (const std::vector<double>* px : {&m.obs_uv, &m.obs_extra})
        for (double v : *px)
            if (!std::isfinite(v)) {
                *msg = "observation has an invalid pixel";
                return VBA_ERR_INELIGIBLE;
            }
    return VBA_OK;
}

void release(Prepared& pr) {
    for (DevBuffer& b : pr.bufs)
        if (b.ptr) (void)hipFree(b.ptr);
    pr.bufs.clear();
}

void release(PreparedBatch& pb) {
    if (pb.graph) (void)hipGraphExecDestroy(pb.graph);
    pb.graph = nullptr;
    pb.graph_key.clear();
    release(static_cast<Prepared&>(pb));
}

void release_context(Context& ctx) {
    if (ctx.host_scalars) (void)hipHostFree(ctx.host_scalars);
    if (ctx.host_pcg) (void)hipHostFree(ctx.host_pcg);
    if (ctx.host_status) (void)hipHostFree(ctx.host_status);
    if (ctx.host_pcg_ring) (void)hipHostFree(ctx.host_pcg_ring);
    ctx.host_pcg_ring = nullptr;
    for (hipEvent_t& e : ctx.pcg_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    ctx.host_scalars = nullptr;
    ctx.host_pcg = nullptr;
    ctx.host_status = nullptr;
    for (const Context::Pending& p : ctx.pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    ctx.pending.clear();
    for (hipEvent_t e : ctx.event_pool) (void)hipEventDestroy(e);
    ctx.event_pool.clear();
    release_reduced(ctx);
}

int32_t prepare(Context& ctx, Prepared& pr, std::string* msg) {
    hipStream_t s = ctx.stream;
    const Model& m = pr.m;
    try {
        if (!ctx.host_scalars) {
            hip_check(hipHostMalloc((void**)&ctx.host_scalars, sizeof(StepScalars), hipHostMallocDefault), "hipHostMalloc");
            hip_check(hipHostMalloc((void**)&ctx.host_pcg, sizeof(PcgState), hipHostMallocDefault), "hipHostMalloc");
        }
        if (!ctx.host_status)
            hip_check(hipHostMalloc((void**)&ctx.host_status, sizeof(int32_t), hipHostMallocDefault), "hipHostMalloc");
        if (!ctx.host_pcg_ring) {
            hip_check(hipHostMalloc((void**)&ctx.host_pcg_ring, Context::kPcgRing * sizeof(PcgState),
                                    hipHostMallocDefault), "hipHostMalloc");
            for (hipEvent_t& e : ctx.pcg_ev)
                hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        }
        upload(pr, B_POSES0, m.poses, s);
        upload(pr, B_LMS0, m.lms, s);
        ensure(pr, B_POSES, m.poses.size() * sizeof(double));
        // 16 bytes past the last landmark: the operator moves landmark
        // positions into LDS in whole 16-byte pieces from a 16-byte boundary
        ensure(pr, B_LMS, m.lms.size() * sizeof(double) + 16);
        ensure(pr, B_POSES_T, m.poses.size() * sizeof(double));
        ensure(pr, B_LMS_T, m.lms.size() * sizeof(double) + 16);
        upload(pr, B_SENSORS, m.sensors, s);
        upload(pr, B_POSE_VAR, m.pose_var, s);
        upload(pr, B_LM_VAR, m.lm_var, s);
        upload(pr, B_OBS_UV, m.obs_uv, s);
        if (m.any_rig || m.any_stereo) upload(pr, B_OBS_SENSOR, m.obs_sensor, s);
        if (m.any_stereo) {
            upload(pr, B_OBS_KIND, m.obs_kind, s);
            upload(pr, B_OBS_EXTRA, m.obs_extra, s);
        }
        upload(pr, B_EDGE_LM, m.edge_lm, s);
        upload(pr, B_EDGE_OBS, m.edge_obs, s);
        upload(pr, B_PERM_CL, m.perm_cl, s);
        upload(pr, B_L_POSE, m.s_pose, s);
        upload(pr, B_LM_STORE, m.lm_store, s);
        upload(pr, B_LM_EDGE_PTR, m.lm_edge_ptr, s);
        upload(pr, B_CAM_TILES, m.cam_tiles, s);
        upload(pr, B_CAM_TILE_PTR, m.cam_tile_ptr, s);
        upload(pr, B_LM_TILES, m.lm_tiles, s);
        upload(pr, B_TILE_PERM, m.tile_perm, s);
        upload(pr, B_BIG_TILES, m.big_tiles, s);
        upload(pr, B_POSE_RUN_PTR, m.pose_run_ptr, s);
        upload(pr, B_RUN_SLOT, m.run_slot, s);
        upload(pr, B_S_LM, m.s_lm, s);
        upload(pr, B_G32_POS, m.g32_pos, s);
        upload(pr, B_T_LM, m.t_lm, s);
        upload(pr, B_T_HEAD, m.t_head, s);
        upload(pr, B_T_RPOSE, m.t_rpose, s);
        upload(pr, B_T_RSLOT, m.t_rslot, s);
        if (!m.obs_w.empty()) upload(pr, B_OBS_W, m.obs_w, s);
        hip_check(hipStreamSynchronize(s), "sync");
        pr.final_slot = -1;
        return VBA_OK;
    } catch (const Fail& f) {
        *msg = f.msg;
        (void)hipStreamSynchronize(s);
        return f.code;
    }
}

int32_t dense_solve(Context& ctx, uint32_t n, const double* h, const double* b, double* x, std::string* msg) {
    if (n == 0) return VBA_OK;
    hipStream_t s = ctx.stream;
    const size_t nn = (size_t)n * n;
    // row-major copy of the column-major H: the kernels' lower triangle is
    // the caller's lower triangle, which is all Cholesky reads
    std::vector<double> rows(nn);
    for (size_t c = 0; c < n; ++c)
        for (size_t r = 0; r < n; ++r) rows[r * n + c] = h[c * n + r];
    DevBuffer bS, bwork, bb, bx, bst;
    auto alloc = [](DevBuffer& d, size_t bytes) {
        hip_check(hipMalloc(&d.ptr, bytes), "hipMalloc");
        d.bytes = bytes;
    };
    int32_t code = VBA_OK;
    try {
        if (!ctx.host_status)
            hip_check(hipHostMalloc((void**)&ctx.host_status, sizeof(int32_t), hipHostMallocDefault), "hipHostMalloc");
        alloc(bS, nn * sizeof(double));
        alloc(bwork, (nn + n) * sizeof(double));
        alloc(bb, n * sizeof(double));
        alloc(bx, n * sizeof(double));
        alloc(bst, 2 * sizeof(int32_t));
        hip_check(hipMemcpyAsync(bS.ptr, rows.data(), nn * sizeof(double), hipMemcpyHostToDevice, s), "upload");
        hip_check(hipMemcpyAsync(bb.ptr, b, n * sizeof(double), hipMemcpyHostToDevice, s), "upload");
        hip_check(launch_dense_solve_plain((double*)bS.ptr, (double*)bwork.ptr, (const double*)bb.ptr,
                                           (double*)bx.ptr, (int32_t)n, (int32_t*)bst.ptr, s),
                  "dense solve");
        hip_check(hipMemcpyAsync(ctx.host_status, bst.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipMemcpyAsync(x, bx.ptr, n * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "sync");
        if (*ctx.host_status != 0) {
            *msg = n <= (uint32_t)kLuFallbackMax
                       ? "singular system: neither Cholesky nor LU with partial pivoting solves it"
                       : "singular system: Cholesky fails (no LU fallback beyond 3072 unknowns on the device)";
            code = VBA_ERR_SINGULAR_SYSTEM;
        }
    } catch (const Fail& f) {
        *msg = f.msg;
        (void)hipStreamSynchronize(s);
        code = f.code;
    }
    for (DevBuffer* d : {&bS, &bwork, &bb, &bx, &bst})
        if (d->ptr) (void)hipFree(d->ptr);
    return code;
}

int32_t download_state(Context& ctx, const Prepared& pr, double* poses_out, double* landmarks_out) {
    hipStream_t s = ctx.stream;
    const Model& m = pr.m;
    try {
        std::vector<double> poses(m.poses.size()), lms(m.lms.size());
        const int ps = pr.final_slot >= 0 ? pr.final_slot : B_POSES0;
        const int ls = pr.final_slot >= 0 ? pr.final_lm_slot : B_LMS0;
        hip_check(hipMemcpyAsync(poses.data(), pr.bufs[ps].ptr, poses.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                  "D2H");
        hip_check(hipMemcpyAsync(lms.data(), pr.bufs[ls].ptr, lms.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                  "D2H");
        hip_check(hipStreamSynchronize(s), "sync");
        if (poses_out)
            for (size_t i = 0; i < m.pose_slot_of_input.size(); ++i)
                std::memcpy(poses_out + 7 * i, poses.data() + 7 * (size_t)m.pose_slot_of_input[i], 7 * sizeof(double));
        if (landmarks_out)
            for (size_t i = 0; i < m.lm_slot_of_input.size(); ++i)
                std::memcpy(landmarks_out + 3 * i, lms.data() + 3 * (size_t)m.lm_slot_of_input[i], 3 * sizeof(double));
        return VBA_OK;
    } catch (const Fail& f) {
        ctx.last_error = f.msg;
        return f.code;
    }
}

int32_t set_weights(Context& ctx, Prepared& pr, const double* weights) {
    Model& m = pr.m;
    try {
        if (!weights) {
            m.obs_w.clear();
            return VBA_OK;
        }
        m.obs_w.resize(m.num_obs);
        for (uint64_t k = 0; k < m.num_obs; ++k) m.obs_w[k] = weights[m.obs_input[k]];
        upload(pr, B_OBS_W, m.obs_w, ctx.stream);
        return VBA_OK;
    } catch (const Fail& f) {
        ctx.last_error = f.msg;
        return f.code;
    }
}

int32_t adopt_final_state(Context& ctx, Prepared& pr) {
    if (pr.final_slot < 0) return VBA_OK;
    try {
        const Model& m = pr.m;
        hip_check(hipMemcpyAsync(pr.bufs[B_POSES0].ptr, pr.bufs[pr.final_slot].ptr, m.poses.size() * sizeof(double),
                                 hipMemcpyDeviceToDevice, ctx.stream), "D2D");
        hip_check(hipMemcpyAsync(pr.bufs[B_LMS0].ptr, pr.bufs[pr.final_lm_slot].ptr, m.lms.size() * sizeof(double),
                                 hipMemcpyDeviceToDevice, ctx.stream), "D2D");
        hip_check(hipStreamSynchronize(ctx.stream), "sync");
        pr.final_slot = pr.final_lm_slot = -1;
        return VBA_OK;
    } catch (const Fail& f) {
        ctx.last_error = f.msg;
        return f.code;
    }
}

namespace {
DevProblem bind(const Context& ctx, Prepared& pr, const vba_options& o, bool mixed, bool matrix_free,
                hipStream_t s);
}  // namespace

int32_t squared_residuals(Context& ctx, Prepared& pr, const vba_options& o, double* out) {
    hipStream_t s = ctx.stream;
    const Model& m = pr.m;
    try {
        DevProblem P = bind(ctx, pr, o, false, !dense_family(o.linear_solver), s);
        P.obs_s = (double*)ensure(pr, B_OBS_S, std::max<uint64_t>(m.num_obs, 1) * sizeof(double));
        launch_cost(P, false, s);
        std::vector<double> sc(m.num_obs);
        hip_check(hipMemcpyAsync(sc.data(), P.obs_s, m.num_obs * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "sync");
        for (uint64_t k = 0; k < m.num_obs; ++k) out[m.obs_input[k]] = sc[k];
        return VBA_OK;
    } catch (const Fail& f) {
        ctx.last_error = f.msg;
        return f.code;
    }
}

namespace {

// Device pointers of a prepared problem for one solve: structure, the state
// reset to the prepared initial state, and working buffers (allocated on
// first use, kept for later solves).
DevProblem bind(const Context& ctx, Prepared& pr, const vba_options& o, bool mixed, bool matrix_free,
                hipStream_t s) {
    const Model& m = pr.m;
    DevProblem P{};
    P.num_cus = ctx.num_cus;
    P.num_pose_slots = (int32_t)m.num_pose_slots;
    P.num_lm_slots = (int32_t)m.num_lm_slots;
    P.num_edges = (int32_t)m.num_edges;
    P.num_obs = (int32_t)m.num_obs;
    P.num_cam_tiles = (int32_t)m.cam_tiles.size();
    P.num_lm_tiles = (int32_t)m.lm_tiles.size();
    P.any_rig = m.any_rig || m.any_stereo;  // per-observation sensors
    P.kernel_kind = (int32_t)o.robust_kernel;
    P.kernel_scale = o.robust_scale;
    const size_t NP = m.num_pose_slots, NL = m.num_lm_slots, NE = m.num_edges, NT = m.cam_tiles.size();
    const size_t NLT = m.lm_tiles.size(), NR = std::max<uint64_t>(m.num_runs, 1);  // partials by run slot
    const size_t pose_blocks = (NP + kBlock - 1) / kBlock;
    P.poses = buf<double>(pr, B_POSES);
    P.lms = buf<double>(pr, B_LMS);
    P.poses_trial = buf<double>(pr, B_POSES_T);
    P.lms_trial = buf<double>(pr, B_LMS_T);
    hip_check(hipMemcpyAsync(P.poses, pr.bufs[B_POSES0].ptr, m.poses.size() * sizeof(double), hipMemcpyDeviceToDevice,
                             s), "D2D");
    hip_check(hipMemcpyAsync(P.lms, pr.bufs[B_LMS0].ptr, m.lms.size() * sizeof(double), hipMemcpyDeviceToDevice, s),
              "D2D");
    P.sensors = buf<DevSensor>(pr, B_SENSORS);
    P.pose_var = buf<int32_t>(pr, B_POSE_VAR);
    P.lm_var = buf<int32_t>(pr, B_LM_VAR);
    P.obs_uv = buf<double>(pr, B_OBS_UV);
    P.obs_w = m.obs_w.empty() ? nullptr : buf<double>(pr, B_OBS_W);
    P.obs_s = nullptr;
    P.obs_sensor = m.any_rig || m.any_stereo ? buf<uint16_t>(pr, B_OBS_SENSOR) : nullptr;
    P.obs_kind = m.any_stereo ? buf<uint8_t>(pr, B_OBS_KIND) : nullptr;
    P.obs_extra = m.any_stereo ? buf<double>(pr, B_OBS_EXTRA) : nullptr;
    P.edge_lm = buf<int32_t>(pr, B_EDGE_LM);
    P.edge_obs = buf<int32_t>(pr, B_EDGE_OBS);
    P.perm_cl = buf<int32_t>(pr, B_PERM_CL);
    P.s_pose = buf<int32_t>(pr, B_L_POSE);
    P.lm_store = buf<int32_t>(pr, B_LM_STORE);
    P.lm_edge_ptr = buf<int32_t>(pr, B_LM_EDGE_PTR);
    P.cam_tiles = buf<CamTile>(pr, B_CAM_TILES);
    P.cam_tile_ptr = buf<int32_t>(pr, B_CAM_TILE_PTR);
    P.lm_tiles = buf<LmTile>(pr, B_LM_TILES);
    P.tile_perm = buf<uint16_t>(pr, B_TILE_PERM);
    P.big_tiles = buf<int32_t>(pr, B_BIG_TILES);
    P.num_big_tiles = (int32_t)m.big_tiles.size();
    P.pose_run_ptr = buf<int32_t>(pr, B_POSE_RUN_PTR);
    P.run_slot = buf<int32_t>(pr, B_RUN_SLOT);
    P.s_lm = buf<uint8_t>(pr, B_S_LM);
    P.Gl64 = (double*)ensure(pr, B_GL64, 6 * NE * sizeof(double));
    P.hl64 = (double*)ensure(pr, B_HL64, 3 * NE * sizeof(double));
    P.g32_pos = buf<int32_t>(pr, B_G32_POS);
    P.t_lm = buf<uint8_t>(pr, B_T_LM);
    P.t_head = buf<uint64_t>(pr, B_T_HEAD);
    P.t_rpose = buf<int32_t>(pr, B_T_RPOSE);
    P.t_rslot = buf<int32_t>(pr, B_T_RSLOT);
    if (mixed) {
        P.Gl32 = (float*)ensure(pr, B_GL32, (size_t)kG32Chunk * std::max(m.num_g32_chunks, 1) * sizeof(float));
        P.Vm = (double*)ensure(pr, B_VM, 6 * NL * sizeof(double));
        P.VinvM = (double*)ensure(pr, B_VINV32, 6 * NL * sizeof(double));
    }
    P.tile27 = (double*)ensure(pr, B_TILE27, 27 * std::max<size_t>(NT, 1) * sizeof(double));
    P.U = (double*)ensure(pr, B_U, 21 * NP * sizeof(double));
    P.g = (double*)ensure(pr, B_G, 6 * NP * sizeof(double));
    P.V = (double*)ensure(pr, B_V, 6 * NL * sizeof(double));
    P.gl = (double*)ensure(pr, B_GL, 3 * NL * sizeof(double));
    P.Vinv64 = (double*)ensure(pr, B_VINV64, 6 * NL * sizeof(double));
    P.part64 = (double*)ensure(pr, B_PART64, 6 * NR * sizeof(double));
    P.b = (double*)ensure(pr, B_VB, 6 * NP * sizeof(double));
    P.x = (double*)ensure(pr, B_VX, 6 * NP * sizeof(double));
    if (matrix_free) {
        P.Minv = (double*)ensure(pr, B_MINV, 36 * NP * sizeof(double));
        P.part21 = (double*)ensure(pr, B_PART21, 21 * NR * sizeof(double));
        P.r = (double*)ensure(pr, B_VR, 6 * NP * sizeof(double));
        P.z = (double*)ensure(pr, B_VZ, 6 * NP * sizeof(double));
        P.d = (double*)ensure(pr, B_VD, 6 * NP * sizeof(double));
        P.q = (double*)ensure(pr, B_VQ, 6 * NP * sizeof(double));
        P.num_bpart = (int32_t)((NP + (kBlock / 64) - 1) / (kBlock / 64));
        P.num_cpart = (int32_t)((NP + (kBlock / 64) * kPosesPerWave - 1) / ((kBlock / 64) * kPosesPerWave));
        P.bpart = (double*)ensure(pr, B_BPART, 2 * (size_t)std::max(P.num_bpart, 1) * sizeof(double));
        P.cpart = (double*)ensure(pr, B_CPART, 2 * (size_t)std::max(P.num_cpart, 1) * sizeof(double));
        P.pcg = (PcgState*)ensure(pr, B_PCG, sizeof(PcgState));
    }
    P.delta_l = (double*)ensure(pr, B_DELTA_L, 3 * NL * sizeof(double));
    const size_t nscalar = 4 * (NT + NLT + pose_blocks + 1);
    P.tile_scalar = (double*)ensure(pr, B_TILE_SCALAR, nscalar * sizeof(double));
    P.scalars = (StepScalars*)ensure(pr, B_SCALARS, sizeof(StepScalars));
    hip_check(hipMemsetAsync(P.x, 0, 6 * NP * sizeof(double), s), "memset");
    hip_check(hipMemsetAsync(P.delta_l, 0, 3 * NL * sizeof(double), s), "memset");
    hip_check(hipMemsetAsync(P.tile_scalar, 0, nscalar * sizeof(double), s), "memset");
    return P;
}

}  // namespace

int32_t run(Context& ctx, Prepared& pr, const vba_options& o, vba_iteration_stats* stats, uint32_t cap,
            vba_result* res, const JointSpec* joint) {
    hipStream_t s = ctx.stream;
    const Model& m = pr.m;
    const bool matrix_free = !dense_family(o.linear_solver);  // a PCG solver
    const bool xplicit = o.linear_solver == VBA_SOLVER_PCG_EXPLICIT;
    const bool mixed = matrix_free && o.precision == VBA_PRECISION_MIXED;
    const bool accept_iterate = o.pcg_maxit_policy == VBA_PCG_MAXIT_ACCEPT_ITERATE;
    try {
        if (!matrix_free && m.P > kDenseMaxPoses)
            throw Fail{VBA_ERR_INVALID_ARGUMENT, "dense reduced solve limited to 8192 variable poses"};
        DevProblem P = bind(ctx, pr, o, mixed, matrix_free, s);
        const size_t NP = m.num_pose_slots, NL = m.num_lm_slots, NE = m.num_edges, NT = m.cam_tiles.size();
        const size_t NR = m.num_runs;
        DensePairs D{};
        double* dS = nullptr;
        double* dwork = nullptr;
        int32_t* dstatus = nullptr;
        const int32_t JK = joint ? joint->k : 0;
        if (JK) {
            if (matrix_free) throw Fail{VBA_ERR_INVALID_ARGUMENT, "the joint camera solve is dense"};
            const size_t NA = 6 * (size_t)JK + JK * (JK + 1) / 2 + JK;
            P.joint_k = JK;
            P.joint_shared = joint->shared;
            P.sensors_trial = (DevSensor*)ensure(pr, B_SENSORS_T, m.sensors.size() * sizeof(DevSensor));
            hip_check(hipMemcpyAsync(P.sensors_trial, P.sensors, m.sensors.size() * sizeof(DevSensor),
                                     hipMemcpyDeviceToDevice, s), "D2D");
            P.joint_E = (double*)ensure(pr, B_J_E, 3 * JK * std::max<size_t>(NE, 1) * sizeof(double));
            P.joint_tile = (double*)ensure(pr, B_J_TILE, NA * std::max<size_t>(NT, 1) * sizeof(double));
            P.joint_hpk = (double*)ensure(pr, B_J_HPK, 6 * JK * std::max<size_t>(NP, 1) * sizeof(double));
            P.joint_kk = (double*)ensure(pr, B_J_KK, NA * sizeof(double));
            P.joint_Wk = (double*)ensure(pr, B_J_WK, 3 * JK * std::max<size_t>(NL, 1) * sizeof(double));
            P.joint_part = (double*)ensure(pr, B_J_PART, joint_part_size(P) * sizeof(double));
            P.joint_rhs = (double*)ensure(pr, B_J_RHS, JK * sizeof(double));
            P.joint_delta = (double*)ensure(pr, B_J_DELTA, JK * sizeof(double));
            P.joint_gl = (double*)ensure(pr, B_J_GL, 3 * std::max<size_t>(NL, 1) * sizeof(double));
        }
        const int32_t ndense = (int32_t)(6 * m.P) + JK;
        if (!matrix_free || xplicit) {
            if (!pr.dense_built) {
                build_dense(pr);
                upload(pr, B_DENSE_BLOCK_PQ, pr.dense_block_pq, s);
                upload(pr, B_DENSE_BLOCK_PTR, pr.dense_block_ptr, s);
                upload(pr, B_DENSE_TERMS, pr.dense_terms, s);
                upload(pr, B_DENSE_VAR_SLOT, pr.dense_var_slot, s);
            }
            D.num_blocks = (int32_t)(pr.dense_block_ptr.size() - 1);
            D.block_pq = buf<int32_t>(pr, B_DENSE_BLOCK_PQ);
            D.block_ptr = buf<int32_t>(pr, B_DENSE_BLOCK_PTR);
            D.terms = buf<int32_t>(pr, B_DENSE_TERMS);
            D.var_slot = buf<int32_t>(pr, B_DENSE_VAR_SLOT);
        }
        ExplicitDev X{};
        double* dbuf[2] = {P.d, nullptr};  // PCG direction by iteration parity (explicit)
        if (xplicit) {
            if (!pr.explicit_built) {
                build_explicit(pr);
                upload(pr, B_X_BLOCK_POS, pr.x_block_pos, s);
                upload(pr, B_X_ROW_PTR, pr.x_row_ptr, s);
                upload(pr, B_X_COL, pr.x_col, s);
                upload(pr, B_X_DIAG_POS, pr.x_diag_pos, s);
            }
            X.num_blocks = D.num_blocks;
            X.num_full = (int32_t)pr.x_col.size();
            X.block_pos = buf<int32_t>(pr, B_X_BLOCK_POS);
            X.row_ptr = buf<int32_t>(pr, B_X_ROW_PTR);
            X.col = buf<int32_t>(pr, B_X_COL);
            X.diag_pos = buf<int32_t>(pr, B_X_DIAG_POS);
            X.S64 = (double*)ensure(pr, B_X_S64, 36 * (size_t)X.num_full * sizeof(double));
            X.S32 = mixed ? (float*)ensure(pr, B_X_S32, 36 * (size_t)X.num_full * sizeof(float)) : nullptr;
            dbuf[1] = (double*)ensure(pr, B_VD2, 6 * NP * sizeof(double));
            hip_check(hipMemsetAsync(dbuf[1], 0, 6 * NP * sizeof(double), s), "memset");
        }
        if (!matrix_free) {
            dS = (double*)ensure(pr, B_DENSE_S, (size_t)ndense * ndense * sizeof(double));
            dwork = (double*)ensure(pr, B_DENSE_WORK, ((size_t)ndense * ndense + ndense) * sizeof(double));
            dstatus = (int32_t*)ensure(pr, B_DENSE_STATUS, 2 * sizeof(int32_t));
        }
        // Algorithmic bytes per launch of the profiled kernels: what the data
        // layout must move at least once (include/vba.h, vba_profile_query).
        // Pose vectors gathered per edge are counted once per pose.
        const double real = mixed ? 4.0 : 8.0;
        // tile pass: per edge G (6 reals) and its local landmark (1); per
        // landmark X (24) and V^-1 (48); per run the FP64 partial (48); per
        // pose z and d (96); per tile its bounds (20), run-start mask (32)
        // and the pose and slot tables of its 64 runs (512)
        const double bytes_apply = NE * (6.0 * real + 1.0) + NL * (24.0 + 48.0) + NR * 48.0 + NP * 96.0 +
                                   m.lm_tiles.size() * (20.0 + 32.0 + 512.0);
        const double bytes_pcg_pose = NR * 48.0 + NP * (8.0 + 96.0 + 168.0 + 96.0);
        const double bytes_pcg_step = NP * (4.0 * 48.0 + 288.0 + 3.0 * 48.0);
        const double bytes_linearize = m.num_obs * 16.0 + NE * (4.0 + 4.0 + 4.0 + 48.0 + 24.0 + (mixed ? 24.0 : 0.0)) +
                                       NL * 24.0 + NT * (16.0 + 27.0 * 8.0) + NP * 56.0;
        const double bytes_cost = m.num_obs * 16.0 + NE * 8.0 + NL * 24.0 + NT * (16.0 + 24.0) + NP * 56.0;
        // explicit matrix: per term its indices (12); per edge G and per
        // landmark X and V^-1 once; per stored block 36 values (FP64, and the
        // FP32 copy); per pose U, the diagonal read and Minv. Product: per
        // stored block 36 reals and its column; per pose z, d_old, d_new, q.
        const double nterms = xplicit ? (double)pr.dense_terms.size() / 3.0 : 0.0;
        const double bytes_xassemble = nterms * 12.0 + NE * 48.0 + NL * 72.0 +
                                       X.num_full * 36.0 * (8.0 + (mixed ? 4.0 : 0.0)) + NP * (168.0 + 288.0 + 288.0);
        const double bytes_xapply = X.num_full * (36.0 * real + 4.0) + NP * (4.0 + 4.0 * 48.0);
        // dense assembly: per term its indices (12); per edge G, per landmark X
        // and V^-1, per pose U once; per block its pair and term pointer (12);
        // the whole n x n S written. Dense factor: n^3 / 3 FLOPs.
        const double dterms = !matrix_free ? (double)pr.dense_terms.size() / 3.0 : 0.0;
        const double bytes_dassemble = dterms * 12.0 + NE * 48.0 + NL * 72.0 + NP * 168.0 + D.num_blocks * 12.0 +
                                       8.0 * (double)ndense * ndense;
        const double flops_dfactor = (double)ndense * ndense * ndense / 3.0;

        hip_check(hipEventRecord(ctx.ev_start, s), "event");

        auto read_scalars = [&]() {
            launch_scalars_finalize(P, s);
            hip_check(hipGetLastError(), "kernel launch");
            hip_check(hipMemcpyAsync(ctx.host_scalars, P.scalars, sizeof(StepScalars), hipMemcpyDeviceToHost, s), "D2H");
            hip_check(hipStreamSynchronize(s), "sync");
            harvest(ctx);
            return *ctx.host_scalars;
        };
        auto read_pcg = [&]() {
            hip_check(hipGetLastError(), "kernel launch");
            hip_check(hipMemcpyAsync(ctx.host_pcg, P.pcg, sizeof(PcgState), hipMemcpyDeviceToHost, s), "D2H");
            hip_check(hipStreamSynchronize(s), "sync");
            return *ctx.host_pcg;
        };

        // ---- initial cost (lm.rs:323-327) ----
        hip_check(hipMemsetAsync(P.scalars, 0, sizeof(StepScalars), s), "memset");
        {
            ProfScope ps(ctx, K_COST, bytes_cost);
            launch_cost(P, false, s);
        }
        StepScalars sc = read_scalars();
        const double initial_cost = sc.cost_sum + sc.cost_comp;
        uint64_t current_nonproj = sc.nonprojectable;
        if (matrix_free && !std::isfinite(initial_cost))
            throw Fail{VBA_ERR_INELIGIBLE, "initial visual cost is non-finite"};

        double current_cost = initial_cost;
        double lambda = o.has_initial_lambda ? o.initial_lambda : 0.0;
        bool converged = false, need_lin = true;
        uint32_t written = 0, total_pcg = 0, linearizations = 0;
        const int32_t max_pcg = (int32_t)o.max_pcg_iterations;
        auto push = [&](const vba_iteration_stats& st) {
            if (stats && written < cap) stats[written] = st;
            ++written;
        };

        for (uint32_t iteration = 0; iteration < o.max_iterations; ++iteration) {
            if (need_lin) {
                {
                    ProfScope ps(ctx, K_LINEARIZE, bytes_linearize);
                    launch_linearize(P, mixed, s);
                }
                launch_camera_finalize(P, s);
                launch_landmark_reduce(P, s);
                if (JK) launch_joint_linearize(P, s);
                need_lin = false;
                ++linearizations;
            }
            hip_check(hipMemsetAsync(P.scalars, 0, sizeof(StepScalars), s), "memset");
            launch_landmark_prepare(P, lambda, s);
            launch_rhs(P, s);

            vba_iteration_stats st;
            std::memset(&st, 0, sizeof(st));
            st.iteration = iteration;
            st.pcg_iterations = -1;
            bool solved = true;
            if (!matrix_free) {
                hip_check(hipMemsetAsync(dS, 0, (size_t)ndense * ndense * sizeof(double), s), "memset");
                {
                    ProfScope a(ctx, K_DENSE_ASSEMBLE, bytes_dassemble);
                    launch_dense_assemble(P, D, dS, ndense, lambda, s);
                }
                if (JK) launch_joint_border(P, dS, ndense, lambda, s);
                {
                    ProfScope f(ctx, K_DENSE_SOLVE, flops_dfactor);
                    hip_check(launch_dense_solve(P, D, dS, dwork, ndense, dstatus, s,
                                                 o.linear_solver != VBA_SOLVER_SPARSE),
                              "dense solve");
                }
                hip_check(hipMemcpyAsync(ctx.host_status, dstatus, sizeof(int32_t), hipMemcpyDeviceToHost, s), "D2H");
                hip_check(hipStreamSynchronize(s), "sync");
                solved = *ctx.host_status == 0;
            } else {
                if (xplicit) {
                    ProfScope a(ctx, K_EXPLICIT_ASSEMBLE, bytes_xassemble);
                    launch_explicit_assemble(P, D, X, lambda, s);
                } else {
                    launch_precondition(P, lambda, s);
                }
                launch_pcg_init(P, o.pcg_relative_tolerance, o.pcg_absolute_tolerance, max_pcg, lambda, s);
                int32_t k = 0;
                int32_t fresh = 1;
                PcgState ps{};
                // one PCG iteration's launches; fresh_kk: the first of a
                // solve or restart (d starts from z)
                auto launch_iteration = [&](int32_t kk, int32_t fresh_kk) {
                    if (xplicit) {
                        // iteration kk reads d from dbuf[kk & 1] and writes the next
                        double* d_new = dbuf[(kk + 1) & 1];
                        {
                            ProfScope a(ctx, K_EXPLICIT_APPLY, bytes_xapply, kk);
                            launch_pcg_explicit(P, X, mixed, dbuf[kk & 1], d_new, kk, fresh_kk, s);
                        }
                        DevProblem Ps = P;
                        Ps.d = d_new;
                        ProfScope c(ctx, K_PCG_STEP, bytes_pcg_step, kk);
                        launch_pcg_step(Ps, kk, s);
                        return;
                    }
                    {
                        ProfScope a(ctx, K_SCHUR_APPLY, bytes_apply, kk, true);
                        launch_schur_apply(P, mixed, P.z, P.d, kk, fresh_kk, s,
                                           (int32_t)ctx.tuning.apply_waves_per_cu * ctx.num_cus, a.events());
                    }
                    {
                        ProfScope b(ctx, K_PCG_POSE, bytes_pcg_pose, kk, true);
                        launch_pcg_pose(P, s, b.events());
                    }
                    {
                        ProfScope c(ctx, K_PCG_STEP, bytes_pcg_step, kk, true);
                        launch_pcg_step(P, kk, s, c.events());
                    }
                };
                while (true) {
                    // Iterations k .. in chunks (the context's schedule): each
                    // chunk ends with a status kernel and a copy of the PCG
                    // state into a pinned ring slot; the host waits for one
                    // chunk's slot while `lookahead` later chunks are already
                    // queued. The device stops itself, so launches past the
                    // stop do nothing and no schedule changes a result.
                    {
                        struct Ticket {
                            int32_t end;
                            int slot;
                            size_t pending_end;
                        };
                        Ticket inflight[Context::kPcgRing];
                        int n_inflight = 0;
                        const vba_tuning& T = ctx.tuning;
                        const int32_t k0 = k, fresh0 = fresh;
                        int32_t next = k0;
                        auto enqueue = [&](int32_t end) {
                            for (int32_t kk = next; kk < end; ++kk) launch_iteration(kk, fresh0 && kk == k0);
                            launch_pcg_status(P, end, fresh0 && end == k0, s);
                            const int slot = (int)(ctx.pcg_ring_next++ % Context::kPcgRing);
                            hip_check(hipMemcpyAsync(ctx.host_pcg_ring + slot, P.pcg, sizeof(PcgState),
                                                     hipMemcpyDeviceToHost, s), "D2H");
                            hip_check(hipEventRecord(ctx.pcg_ev[slot], s), "hipEventRecord");
                            inflight[n_inflight++] = Ticket{end, slot, ctx.pending.size()};
                            next = end;
                        };
                        int32_t first = k0 + (int32_t)T.pcg_chunk;
                        if (T.pcg_predict && k0 == 0 && pr.last_pcg > first) first = pr.last_pcg;
                        enqueue(std::min(first, max_pcg));
                        for (uint32_t a = 0; a < T.pcg_lookahead && next < max_pcg; ++a)
                            enqueue(std::min(next + (int32_t)T.pcg_chunk, max_pcg));
                        int32_t seen = k0;
                        while (true) {
                            if (n_inflight == 0) throw Fail{VBA_ERR_DEVICE, "PCG did not stop at its cap"};
                            const Ticket t = inflight[0];
                            for (int i = 1; i < n_inflight; ++i) inflight[i - 1] = inflight[i];
                            --n_inflight;
                            hip_check(hipGetLastError(), "kernel launch");
                            hip_check(hipEventSynchronize(ctx.pcg_ev[t.slot]), "hipEventSynchronize");
                            ps = ctx.host_pcg_ring[t.slot];
                            if (ps.stop) {
                                // every later launch found the stop flag; their
                                // profile records are dropped (the records still
                                // pending are this segment's and non-PCG ones)
                                drop_after(ctx, 0, ps.iter);
                                harvest_upto(ctx, ctx.pending.size());
                                break;
                            }
                            const size_t done = harvest_upto(ctx, t.pending_end);
                            for (int i = 0; i < n_inflight; ++i) inflight[i].pending_end -= done;
                            if (ps.iter <= seen) throw Fail{VBA_ERR_DEVICE, "PCG made no progress"};
                            seen = ps.iter;
                            if (next < max_pcg) enqueue(std::min(next + (int32_t)T.pcg_chunk, max_pcg));
                        }
                        k = ps.iter;
                        fresh = 0;
                    }
                    if (ps.status != PCG_CONVERGED && ps.status != PCG_MAXIT) {
                        solved = false;
                        break;
                    }
                    // true residual re-check (implicit_schur.rs:613-690, 750-761)
                    if (xplicit) launch_explicit_true_residual(P, X, 0, s);
                    else launch_pcg_true_residual(P, s);
                    ps = read_pcg();
                    if (std::isfinite(ps.true_norm) && ps.true_norm <= ps.target) break;
                    if (!std::isfinite(ps.true_norm)) {
                        solved = false;
                        break;
                    }
                    if (mixed && ps.iter < ps.max_iter) {
                        // FP64 refinement: continue from the true residual
                        if (xplicit) {
                            // the next iteration (ps.iter) starts fresh from d = 0
                            DevProblem Pr = P;
                            Pr.d = dbuf[ps.iter & 1];
                            launch_explicit_true_residual(Pr, X, 1, s);
                        } else {
                            launch_pcg_restart(P, s);
                        }
                        ps = read_pcg();
                        if (ps.stop) {
                            solved = ps.status == PCG_MAXIT && accept_iterate;
                            break;
                        }
                        k = ps.iter;
                        fresh = 1;
                        continue;
                    }
                    solved = accept_iterate;
                    break;
                }
                st.pcg_iterations = ps.iter;
                st.pcg_residual_norm = ps.true_norm;
                st.pcg_target = ps.target;
                total_pcg += (uint32_t)ps.iter;
                pr.last_pcg = ps.iter;
            }
            const double cost_before = current_cost;
            if (!solved) {  // lm.rs:610-662: a failed solve is a rejected step
                lambda = std::min(lambda * o.lambda_increase_factor, o.max_lambda);
                st.cost_before = cost_before;
                st.cost_after = cost_before;
                st.lambda = lambda;
                st.nonprojectable_after = (uint32_t)current_nonproj;
                push(st);
                if (lambda >= o.max_lambda) break;
                continue;
            }
            if (JK) {
                // the trial camera, and back-substitution with g_l + W_lk delta_k
                launch_joint_step(P, s);
                DevProblem Pb = P;
                Pb.gl = P.joint_gl;
                launch_backsub(Pb, s);
            } else {
                launch_backsub(P, s);
            }
            launch_pose_update(P, s);
            {
                ProfScope ps(ctx, K_COST, bytes_cost);
                launch_cost(P, true, s);
            }
            sc = read_scalars();
            const double cost_after = sc.cost_sum + sc.cost_comp;
            const uint64_t nonproj_after = sc.nonprojectable;
            const bool cost_accepted = o.has_initial_lambda ? (cost_after < cost_before) : true;
            const bool accepted = cost_accepted && nonproj_after <= current_nonproj;
            st.cost_before = cost_before;
            st.cost_after = cost_after;
            st.max_pose_step = sc.max_pose_step;
            st.max_landmark_step = sc.max_landmark_step;
            st.nonprojectable_after = (uint32_t)nonproj_after;
            if (!accepted) {  // lm.rs:890-947
                lambda = std::min(lambda * o.lambda_increase_factor, o.max_lambda);
                st.lambda = lambda;
                st.step_accepted = 0;
                push(st);
                if (!o.has_initial_lambda) break;
                if (lambda >= o.max_lambda) break;
                continue;
            }
            st.lambda = lambda;  // lm.rs:949-991
            st.step_accepted = 1;
            push(st);
            std::swap(P.poses, P.poses_trial);
            std::swap(P.lms, P.lms_trial);
            if (JK) {
                DevSensor* cur = const_cast<DevSensor*>(P.sensors);
                P.sensors = P.sensors_trial;
                P.sensors_trial = cur;
            }
            need_lin = true;
            current_cost = cost_after;
            current_nonproj = nonproj_after;
            if (o.has_initial_lambda) lambda = std::max(lambda * o.lambda_decrease_factor, o.min_lambda);
            if (sc.max_pose_step < o.step_tolerance && sc.max_landmark_step < o.step_tolerance) {
                converged = true;
                break;
            }
            if (std::fabs(cost_before - cost_after) < o.cost_tolerance) {
                converged = true;
                break;
            }
            if (o.has_relative_cost_tolerance && std::isfinite(o.relative_cost_tolerance) &&
                o.relative_cost_tolerance >= 0.0 &&
                (cost_before - cost_after) / std::max(std::fabs(cost_before), std::numeric_limits<double>::epsilon()) <
                    o.relative_cost_tolerance) {
                converged = true;
                break;
            }
        }
        hip_check(hipEventRecord(ctx.ev_stop, s), "event");
        hip_check(hipStreamSynchronize(s), "sync");
        harvest(ctx);
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, ctx.ev_start, ctx.ev_stop), "event");
        if (JK) {
            DevSensor cam;
            hip_check(hipMemcpy(&cam, P.sensors, sizeof(DevSensor), hipMemcpyDeviceToHost), "D2H");
            pr.joint_lens = cam.lens;
        }
        pr.final_slot = P.poses == buf<double>(pr, B_POSES) ? B_POSES : B_POSES_T;
        pr.final_lm_slot = P.lms == buf<double>(pr, B_LMS) ? B_LMS : B_LMS_T;
        res->initial_cost = initial_cost;
        res->final_cost = current_cost;
        res->num_iterations = std::min(written, cap);
        res->converged = converged;
        res->num_observations_used = m.num_obs;
        res->total_pcg_iterations = total_pcg;
        res->linearizations = linearizations;
        res->device_seconds = ms * 1e-3;
        return VBA_OK;
    } catch (const Fail& f) {
        ctx.last_error = f.msg;
        (void)hipStreamSynchronize(s);
        for (const Context::Pending& p : ctx.pending) {
            ctx.event_pool.push_back(p.a);
            ctx.event_pool.push_back(p.b);
        }
        ctx.pending.clear();
        return f.code;
    }
}

// ===========================================================================
// Batches of independent windows
// ===========================================================================
namespace {

enum BatchBufId {
    BB_LAYOUT = B_COUNT, BB_STATE, BB_STATS, BB_POSE_WIN, BB_LM_WIN, BB_CAM_TILE_WIN, BB_LM_TILE_WIN,
    BB_S, BB_WORK, BB_WORK_OFFSET, BB_COUNT
};

// Concatenate per-window models into one problem with disjoint slots.
void merge_models(const std::vector<const Model*>& ms, Model* out, std::vector<WindowLayout>* layout,
                  std::vector<int32_t>* pose_win, std::vector<int32_t>* lm_win, std::vector<int32_t>* cam_tile_win,
                  std::vector<int32_t>* lm_tile_win, const std::vector<DenseLists>& dense, DenseLists* dense_out) {
    Model& M = *out;
    M = Model();
    M.any_rig = true;
    for (const Model* mw : ms)
        if (mw->any_stereo) M.any_stereo = true;
    layout->clear();
    pose_win->clear();
    lm_win->clear();
    cam_tile_win->clear();
    lm_tile_win->clear();
    *dense_out = DenseLists();
    int64_t po = 0, lo = 0, eo = 0, oo = 0, to = 0, vo = 0, vlo = 0, so = 0, soff = 0, term_off = 0, ro = 0;
    for (size_t w = 0; w < ms.size(); ++w) {
        const Model& m = *ms[w];
        WindowLayout L{};
        L.tile_begin = (int32_t)to;
        L.tile_end = (int32_t)(to + m.cam_tiles.size());
        L.pose_begin = (int32_t)po;
        L.pose_end = (int32_t)(po + m.num_pose_slots);
        L.lm_begin = (int32_t)lo;
        L.lm_end = (int32_t)(lo + m.num_lm_slots);
        L.var_begin = (int32_t)vo;
        L.var_end = (int32_t)(vo + m.P);
        L.s_offset = soff;
        layout->push_back(L);
        soff += (int64_t)36 * m.P * m.P;
        for (uint32_t s2 = 0; s2 < m.num_pose_slots; ++s2) {
            M.pose_var.push_back(m.pose_var[s2] >= 0 ? (int32_t)(m.pose_var[s2] + vo) : -1);
            pose_win->push_back((int32_t)w);
        }
        for (uint32_t l = 0; l < m.num_lm_slots; ++l) {
            M.lm_var.push_back(m.lm_var[l] >= 0 ? (int32_t)(m.lm_var[l] + vlo) : -1);
            lm_win->push_back((int32_t)w);
        }
        M.poses.insert(M.poses.end(), m.poses.begin(), m.poses.end());
        M.lms.insert(M.lms.end(), m.lms.begin(), m.lms.end());
        for (DevSensor d : m.sensors) {
            d.left += (int32_t)so;
            M.sensors.push_back(d);
        }
        M.obs_uv.insert(M.obs_uv.end(), m.obs_uv.begin(), m.obs_uv.end());
        const bool per_sensor = m.any_rig || m.any_stereo;
        for (uint64_t o = 0; o < m.num_obs; ++o)
            M.obs_sensor.push_back((uint16_t)(so + (per_sensor ? m.obs_sensor[o] : 0)));
        if (M.any_stereo) {
            if (m.any_stereo) {
                M.obs_kind.insert(M.obs_kind.end(), m.obs_kind.begin(), m.obs_kind.end());
                M.obs_extra.insert(M.obs_extra.end(), m.obs_extra.begin(), m.obs_extra.end());
            } else {
                M.obs_kind.insert(M.obs_kind.end(), m.num_obs, OBS_PIXEL);
                M.obs_extra.insert(M.obs_extra.end(), 2 * m.num_obs, 0.0);
            }
        }
        for (uint64_t e = 0; e < m.num_edges; ++e) {
            M.edge_pose.push_back((int32_t)(m.edge_pose[e] + po));
            M.edge_lm.push_back((int32_t)(m.edge_lm[e] + lo));
            M.edge_obs.push_back((int32_t)(m.edge_obs[e] + oo));
            M.l_pose.push_back((int32_t)(m.l_pose[e] + po));
            M.s_pose.push_back((int32_t)(m.s_pose[e] + po));
            M.lm_store.push_back((int32_t)(m.lm_store[e] + eo));
            M.perm_cl.push_back((int32_t)(m.perm_cl[e] + eo));
        }
        for (uint32_t s2 = 0; s2 < m.num_pose_slots; ++s2) {
            M.cam_edge_ptr.push_back((int32_t)(m.cam_edge_ptr[s2] + eo));
            M.cam_tile_ptr.push_back((int32_t)(m.cam_tile_ptr[s2] + to));
        }
        for (uint32_t l = 0; l < m.num_lm_slots; ++l) M.lm_edge_ptr.push_back((int32_t)(m.lm_edge_ptr[l] + eo));
        for (const CamTile& t : m.cam_tiles) {
            M.cam_tiles.push_back({(int32_t)(t.pose + po), (int32_t)(t.begin + eo), (int32_t)(t.end + eo), 0});
            cam_tile_win->push_back((int32_t)w);
        }
        for (int32_t t : m.big_tiles) M.big_tiles.push_back((int32_t)(t + M.lm_tiles.size()));
        for (const LmTile& t : m.lm_tiles) {
            M.lm_tiles.push_back({(int32_t)(t.lm_begin + lo), (int32_t)(t.lm_end + lo), (int32_t)(t.e_begin + eo),
                                  (int32_t)(t.e_end + eo), t.g_chunk + M.num_g32_chunks});
            lm_tile_win->push_back((int32_t)w);
        }
        // camera runs: tile-local permutations are unchanged; run ends shift
        // with the edges
        M.tile_perm.insert(M.tile_perm.end(), m.tile_perm.begin(), m.tile_perm.end());
        M.s_lm.insert(M.s_lm.end(), m.s_lm.begin(), m.s_lm.end());
        for (uint32_t s2 = 0; s2 < m.num_pose_slots; ++s2)
            M.pose_run_ptr.push_back((int32_t)(m.pose_run_ptr[s2] + ro));
        for (int32_t r : m.run_slot) M.run_slot.push_back(r < 0 ? -1 : (int32_t)(r + ro));
        for (int32_t g : m.g32_pos) M.g32_pos.push_back(g + kTileEdges * M.num_g32_chunks);
        M.t_lm.insert(M.t_lm.end(), m.t_lm.begin(), m.t_lm.end());
        M.t_head.insert(M.t_head.end(), m.t_head.begin(), m.t_head.end());
        for (int32_t p : m.t_rpose) M.t_rpose.push_back((int32_t)(p + po));
        for (int32_t r : m.t_rslot) M.t_rslot.push_back((int32_t)(r + ro));
        M.num_g32_chunks += m.num_g32_chunks;
        ro += (int64_t)m.num_runs;
        M.max_lm_tile_edges = std::max(M.max_lm_tile_edges, m.max_lm_tile_edges);
        if (m.any_fixed_pose) M.any_fixed_pose = true;
        const DenseLists& d = dense[w];
        for (size_t k = 0; k < d.block_pq.size(); ++k) dense_out->block_pq.push_back((int32_t)(d.block_pq[k] + po));
        for (size_t k = 0; k + 1 < d.block_ptr.size(); ++k)
            dense_out->block_ptr.push_back((int32_t)(d.block_ptr[k] + term_off));
        for (size_t k = 0; k < d.terms.size(); k += 3) {
            dense_out->terms.push_back((int32_t)(d.terms[k] + lo));
            dense_out->terms.push_back((int32_t)(d.terms[k + 1] + eo));
            dense_out->terms.push_back((int32_t)(d.terms[k + 2] + eo));
        }
        for (int32_t v : d.var_slot) dense_out->var_slot.push_back((int32_t)(v + po));
        term_off += d.block_ptr.empty() ? 0 : d.block_ptr.back();
        po += m.num_pose_slots;
        lo += m.num_lm_slots;
        eo += (int64_t)m.num_edges;
        oo += (int64_t)m.num_obs;
        to += (int64_t)m.cam_tiles.size();
        vo += m.P;
        vlo += m.L;
        so += (int64_t)m.sensors.size();
    }
    M.num_pose_slots = (uint32_t)po;
    M.num_lm_slots = (uint32_t)lo;
    M.P = (uint32_t)vo;
    M.L = (uint32_t)vlo;
    M.num_obs = (uint64_t)oo;
    M.num_edges = (uint64_t)eo;
    M.edge_obs.push_back((int32_t)oo);
    M.cam_edge_ptr.push_back((int32_t)eo);
    M.cam_tile_ptr.push_back((int32_t)to);
    M.lm_edge_ptr.push_back((int32_t)eo);
    M.pose_run_ptr.push_back((int32_t)ro);
    M.num_runs = (uint64_t)ro;
    dense_out->block_ptr.push_back((int32_t)term_off);
}

}  // namespace

int32_t prepare_batch(Context& ctx, std::vector<Model>& models, const std::vector<int32_t>& valid,
                      PreparedBatch* pbp, int32_t* statuses, std::string* message) {
    PreparedBatch& pb = *pbp;
    hipStream_t s = ctx.stream;
    try {
        std::vector<const Model*> ms;
        std::vector<DenseLists> dense;
        pb.count = (uint32_t)models.size();
        pb.index.clear();
        pb.window_of.assign(models.size(), -1);
        for (size_t i = 0; i < models.size(); ++i) {
            if (!valid[i]) continue;
            if (models[i].P > 256) {
                statuses[i] = VBA_ERR_INVALID_ARGUMENT;
                continue;
            }
            pb.window_of[i] = (int32_t)ms.size();
            ms.push_back(&models[i]);
            pb.index.push_back((int32_t)i);
            dense.emplace_back();
            build_dense_lists(models[i], &dense.back());
        }
        const int32_t W = (int32_t)ms.size();
        if (W == 0) return VBA_OK;
        size_t num_sensors = 0;
        for (const Model* mw : ms) num_sensors += mw->sensors.size();
        if (num_sensors > 65535) throw Fail{VBA_ERR_INVALID_ARGUMENT, "batch holds more than 65535 sensors"};
        std::vector<int32_t> pose_win, lm_win, cam_tile_win, lm_tile_win;
        DenseLists dl;
        merge_models(ms, &pb.m, &pb.layout, &pose_win, &lm_win, &cam_tile_win, &lm_tile_win, dense, &dl);
        pb.pose_slot_of_input.clear();
        pb.lm_slot_of_input.clear();
        pb.num_obs.clear();
        for (const Model* mw : ms) {
            pb.pose_slot_of_input.push_back(mw->pose_slot_of_input);
            pb.lm_slot_of_input.push_back(mw->lm_slot_of_input);
            pb.num_obs.push_back(mw->num_obs);
        }
        std::string msg;
        int32_t st = prepare(ctx, pb, &msg);
        if (st != VBA_OK) throw Fail{st, msg};
        pb.bufs.resize(BB_COUNT);
        upload(pb, BB_POSE_WIN, pose_win, s);
        upload(pb, BB_LM_WIN, lm_win, s);
        upload(pb, BB_CAM_TILE_WIN, cam_tile_win, s);
        upload(pb, BB_LM_TILE_WIN, lm_tile_win, s);
        upload(pb, BB_LAYOUT, pb.layout, s);
        pb.work_offset.assign(W, 0);
        pb.s_total = pb.work_total = 0;
        for (int32_t w = 0; w < W; ++w) {
            const int64_t n = 6 * (int64_t)(pb.layout[w].var_end - pb.layout[w].var_begin);
            pb.s_total += n * n;
            pb.work_offset[w] = pb.work_total;
            pb.work_total += n * n + n;
        }
        upload(pb, BB_WORK_OFFSET, pb.work_offset, s);
        pb.num_dense_blocks = (int32_t)(dl.block_ptr.size() - 1);
        upload(pb, B_DENSE_BLOCK_PQ, dl.block_pq, s);
        upload(pb, B_DENSE_BLOCK_PTR, dl.block_ptr, s);
        upload(pb, B_DENSE_TERMS, dl.terms, s);
        upload(pb, B_DENSE_VAR_SLOT, dl.var_slot, s);
        ensure(pb, BB_S, std::max<int64_t>(pb.s_total, 1) * sizeof(double));
        ensure(pb, BB_WORK, std::max<int64_t>(pb.work_total, 1) * sizeof(double));
        ensure(pb, BB_STATE, W * sizeof(WindowState));
        hip_check(hipStreamSynchronize(s), "sync");
        // the structure now lives on the device; keep only what outputs need
        pb.m.obs_uv = std::vector<double>();
        pb.m.obs_sensor = std::vector<uint16_t>();
        pb.m.obs_kind = std::vector<uint8_t>();
        pb.m.obs_extra = std::vector<double>();
        pb.m.obs_input = std::vector<uint32_t>();
        pb.solved = false;
        return VBA_OK;
    } catch (const Fail& f) {
        *message = f.msg;
        (void)hipStreamSynchronize(s);
        return f.code;
    }
}

int32_t run_batch(Context& ctx, PreparedBatch& pb, const vba_options& o, vba_iteration_stats* stats, uint32_t cap,
                  vba_result* results, int32_t* statuses, bool use_graph) {
    hipStream_t s = ctx.stream;
    const int32_t W = (int32_t)pb.index.size();
    if (W == 0) return VBA_OK;
    try {
        if (!ctx.host_scalars) {
            hip_check(hipHostMalloc((void**)&ctx.host_scalars, sizeof(StepScalars), hipHostMallocDefault), "hipHostMalloc");
            hip_check(hipHostMalloc((void**)&ctx.host_pcg, sizeof(PcgState), hipHostMallocDefault), "hipHostMalloc");
        }
        if (!ctx.host_status)
            hip_check(hipHostMalloc((void**)&ctx.host_status, sizeof(int32_t), hipHostMallocDefault), "hipHostMalloc");
        DevProblem P = bind(ctx, pb, o, false, false, s);
        P.any_rig = 1;
        P.obs_sensor = buf<uint16_t>(pb, B_OBS_SENSOR);
        P.pose_win = buf<int32_t>(pb, BB_POSE_WIN);
        P.lm_win = buf<int32_t>(pb, BB_LM_WIN);
        P.cam_tile_win = buf<int32_t>(pb, BB_CAM_TILE_WIN);
        P.lm_tile_win = buf<int32_t>(pb, BB_LM_TILE_WIN);
        P.num_windows = W;
        std::vector<WindowState> init(W);
        std::memset(init.data(), 0, W * sizeof(WindowState));
        for (int32_t w = 0; w < W; ++w) init[w].active = 1, init[w].relin = 1;
        WindowState* dstate = upload(pb, BB_STATE, init, s);
        P.wstate = dstate;
        BatchDev B{};
        B.num_windows = W;
        B.layout = buf<WindowLayout>(pb, BB_LAYOUT);
        B.state = dstate;
        B.stats_capacity = (int32_t)std::max<uint32_t>(cap, 1);
        B.stats = (vba_iteration_stats*)ensure(pb, BB_STATS, (size_t)W * B.stats_capacity * sizeof(vba_iteration_stats));
        B.S = buf<double>(pb, BB_S);
        B.work = buf<double>(pb, BB_WORK);
        B.work_offset = buf<int64_t>(pb, BB_WORK_OFFSET);
        DensePairs D{};
        D.num_blocks = pb.num_dense_blocks;
        D.block_pq = buf<int32_t>(pb, B_DENSE_BLOCK_PQ);
        D.block_ptr = buf<int32_t>(pb, B_DENSE_BLOCK_PTR);
        D.terms = buf<int32_t>(pb, B_DENSE_TERMS);
        D.var_slot = buf<int32_t>(pb, B_DENSE_VAR_SLOT);

        // ---- initial cost and state, then max_iterations LM steps with no
        // host synchronisation: every window decides on the device. Kernel
        // launches only (no memsets or copies), so the sequence can be
        // recorded as a graph ----
        // algorithmic bytes of the batch's dense assembly and FLOPs of its
        // factors, summed over the windows (as for the single problem)
        const Model& bm = pb.m;
        double flops_bfactor = 0.0;
        for (const WindowLayout& L : pb.layout) {
            const double n = 6.0 * (L.var_end - L.var_begin);
            flops_bfactor += n * n * n / 3.0;
        }
        const double bytes_bassemble = (double)pb.dense_terms.size() / 3.0 * 12.0 + bm.num_edges * 48.0 +
                                       bm.num_lm_slots * 72.0 + bm.num_pose_slots * 168.0 +
                                       pb.num_dense_blocks * 12.0 + 8.0 * (double)pb.s_total;
        bool capturing = false;
        auto launch_all = [&]() {
            launch_cost(P, false, s);
            launch_batch_scalars(P, B, s);
            launch_batch_init(P, B, o, s);
            for (uint32_t it = 0; it < o.max_iterations; ++it) {
                launch_linearize(P, false, s);
                launch_camera_finalize(P, s);
                launch_landmark_reduce(P, s);
                launch_landmark_prepare(P, 0.0, s);
                launch_rhs(P, s);
                launch_batch_dense_zero(P, B, s);
                {
                    // events only when the launches are not being recorded
                    // into the batch's graph
                    std::optional<ProfScope> a;
                    if (!capturing) a.emplace(ctx, K_BATCH_DENSE_ASSEMBLE, bytes_bassemble);
                    launch_batch_dense_assemble(P, D, B, s);
                }
                {
                    std::optional<ProfScope> f;
                    if (!capturing) f.emplace(ctx, K_BATCH_DENSE_SOLVE, flops_bfactor);
                    launch_batch_dense_solve(P, D, B, s);
                }
                launch_backsub(P, s);
                launch_pose_update(P, s);
                launch_cost(P, true, s);
                launch_batch_scalars(P, B, s);
                launch_batch_control(P, B, o, (int32_t)it, s);
                launch_batch_commit(P, B, s);
            }
        };
        const bool graph = use_graph && !(o.flags & VBA_FLAG_NO_GRAPH);
        if (graph) {
            // what the recorded kernels' arguments depend on: the options, the
            // stats capacity and the device buffers (their generation)
            std::vector<uint8_t> key(sizeof(vba_options) + sizeof(uint32_t) + sizeof(uint64_t));
            std::memcpy(key.data(), &o, sizeof(vba_options));
            const uint32_t capk = (uint32_t)B.stats_capacity;
            std::memcpy(key.data() + sizeof(vba_options), &capk, sizeof(capk));
            std::memcpy(key.data() + sizeof(vba_options) + sizeof(capk), &pb.alloc_gen, sizeof(uint64_t));
            if (!pb.graph || pb.graph_key != key) {
                if (pb.graph) (void)hipGraphExecDestroy(pb.graph);
                pb.graph = nullptr;
                hipGraph_t g = nullptr;
                hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                capturing = true;
                launch_all();
                capturing = false;
                const hipError_t launched = hipGetLastError();
                hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
                hip_check(launched, "kernel launch (capture)");
                const hipError_t inst = hipGraphInstantiate(&pb.graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                hip_check(inst, "hipGraphInstantiate");
                pb.graph_key = key;
            }
        }
        hip_check(hipEventRecord(ctx.ev_start, s), "event");
        if (graph) hip_check(hipGraphLaunch(pb.graph, s), "hipGraphLaunch");
        else launch_all();
        hip_check(hipGetLastError(), "kernel launch");
        hip_check(hipEventRecord(ctx.ev_stop, s), "event");
        std::vector<WindowState> fin(W);
        std::vector<vba_iteration_stats> st_host((size_t)W * B.stats_capacity);
        hip_check(hipMemcpyAsync(fin.data(), dstate, W * sizeof(WindowState), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipMemcpyAsync(st_host.data(), B.stats, st_host.size() * sizeof(vba_iteration_stats),
                                 hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "sync");
        harvest(ctx);
        float ms_elapsed = 0.f;
        hip_check(hipEventElapsedTime(&ms_elapsed, ctx.ev_start, ctx.ev_stop), "event");
        pb.final_slot = B_POSES;
        pb.final_lm_slot = B_LMS;
        pb.solved = true;
        for (int32_t w = 0; w < W; ++w) {
            const int32_t i = pb.index[w];
            const uint32_t n = (uint32_t)std::min<int32_t>(fin[w].iterations, B.stats_capacity);
            if (stats)
                std::memcpy(stats + (size_t)i * cap, st_host.data() + (size_t)w * B.stats_capacity,
                            std::min(n, cap) * sizeof(vba_iteration_stats));
            vba_result& r = results[i];
            std::memset(&r, 0, sizeof(r));
            r.initial_cost = fin[w].initial_cost;
            r.final_cost = fin[w].current_cost;
            r.num_iterations = std::min(n, cap);
            r.converged = fin[w].converged;
            r.num_observations_used = pb.num_obs[w];
            r.device_seconds = ms_elapsed * 1e-3;
            statuses[i] = VBA_OK;
        }
        return VBA_OK;
    } catch (const Fail& f) {
        ctx.last_error = f.msg;
        (void)hipStreamSynchronize(s);
        return f.code;
    }
}

int32_t download_batch_window(Context& ctx, const PreparedBatch& pb, uint32_t i, double* poses_out,
                              double* landmarks_out) {
    if (i >= pb.window_of.size() || pb.window_of[i] < 0) {
        ctx.last_error = "window is not in the batch";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    const int32_t w = pb.window_of[i];
    const WindowLayout& L = pb.layout[w];
    hipStream_t s = ctx.stream;
    try {
        const int ps = pb.final_slot >= 0 ? pb.final_slot : B_POSES0;
        const int ls = pb.final_slot >= 0 ? pb.final_lm_slot : B_LMS0;
        std::vector<double> poses(7 * (size_t)(L.pose_end - L.pose_begin)), lms(3 * (size_t)(L.lm_end - L.lm_begin));
        hip_check(hipMemcpyAsync(poses.data(), (const double*)pb.bufs[ps].ptr + 7 * (size_t)L.pose_begin,
                                 poses.size() * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipMemcpyAsync(lms.data(), (const double*)pb.bufs[ls].ptr + 3 * (size_t)L.lm_begin,
                                 lms.size() * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "sync");
        const auto& pso = pb.pose_slot_of_input[w];
        const auto& lso = pb.lm_slot_of_input[w];
        if (poses_out)
            for (size_t k = 0; k < pso.size(); ++k)
                std::memcpy(poses_out + 7 * k, poses.data() + 7 * (size_t)pso[k], 7 * sizeof(double));
        if (landmarks_out)
            for (size_t k = 0; k < lso.size(); ++k)
                std::memcpy(landmarks_out + 3 * k, lms.data() + 3 * (size_t)lso[k], 3 * sizeof(double));
        return VBA_OK;
    } catch (const Fail& f) {
        ctx.last_error = f.msg;
        return f.code;
    }
}

}  // namespace vba
