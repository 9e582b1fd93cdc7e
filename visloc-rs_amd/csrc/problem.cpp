// problem.cpp — validation and structure-of-arrays layout (problem.hpp).
#include "problem.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace vba {

namespace {

bool supported_model(int32_t m) { return m >= VBA_CAMERA_PINHOLE && m <= VBA_CAMERA_OPENCV; }

// Camera::intrinsics (crates/core/src/types/camera.rs:241-271).

void quat_to_matrix(const double* q, double R[9]) {
    Quat u{q[0], q[1], q[2], q[3]};
    qmatrix<double>(u, R);
}

// Stable argsort of ids; false on a duplicate id.
bool slot_order(uint64_t n, const uint64_t* ids, std::vector<uint32_t>* slot_of_input,
                std::vector<uint64_t>* sorted) {
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ids[a] < ids[b]; });
    slot_of_input->assign(n, 0);
    sorted->resize(n);
    for (uint64_t s = 0; s < n; ++s) {
        (*slot_of_input)[order[s]] = (uint32_t)s;
        (*sorted)[s] = ids[order[s]];
        if (s > 0 && (*sorted)[s] == (*sorted)[s - 1]) return false;
    }
    return true;
}
// Id -> slot over the ascending ids: arithmetic when the ids are
// consecutive (the common case), binary search otherwise.
struct IdIndex {
    const std::vector<uint64_t>& sorted;
    bool dense;
    explicit IdIndex(const std::vector<uint64_t>& s)
        : sorted(s), dense(!s.empty() && s.back() - s.front() == s.size() - 1) {}
    bool operator()(uint64_t id, uint32_t* slot) const {
        if (dense) {
            if (id < sorted.front() || id > sorted.back()) return false;
            *slot = (uint32_t)(id - sorted.front());
            return true;
        }
        auto it = std::lower_bound(sorted.begin(), sorted.end(), id);
        if (it == sorted.end() || *it != id) return false;
        *slot = (uint32_t)(it - sorted.begin());
        return true;
    }
};

// Camera runs of every landmark tile (Model::tile_perm and the run lists).
// Within a landmark, l-order edges are already in pose order and each pose
// appears once, so a tile of one landmark needs no permutation.
void build_runs(Model* mp) {
    Model& m = *mp;
    m.tile_perm.assign(m.num_edges, 0);
    m.s_lm.assign(m.num_edges, 0);
    m.s_pose.assign(m.num_edges, 0);
    m.lm_store.assign(m.num_edges, 0);
    m.g32_pos.assign(m.num_edges, 0);
    m.num_g32_chunks = 0;
    m.big_tiles.clear();
    std::vector<int32_t> run_pose, run_end, count;
    std::vector<uint64_t> keys;
    std::vector<uint8_t> lm_of_k;
    for (size_t t = 0; t < m.lm_tiles.size(); ++t) {
        const int32_t e0 = m.lm_edge_ptr[m.lm_tiles[t].lm_begin], e1 = m.lm_edge_ptr[m.lm_tiles[t].lm_end];
        const int32_t n = e1 - e0;
        if (n > kTileEdges) {
            m.big_tiles.push_back((int32_t)t);  // identity order; tile_perm is not read
        } else {
            // stable sort by pose: counting sort over the tile's pose range when
            // it is narrow (landmarks ordered by first camera), else a key sort
            int32_t lo = m.l_pose[e0], hi = lo;
            for (int32_t k = 1; k < n; ++k) {
                lo = std::min(lo, m.l_pose[e0 + k]);
                hi = std::max(hi, m.l_pose[e0 + k]);
            }
            const int32_t range = hi - lo + 1;
            if (range <= 8 * n + 64) {
                count.assign(range + 1, 0);
                for (int32_t k = 0; k < n; ++k) ++count[m.l_pose[e0 + k] - lo + 1];
                for (int32_t r = 0; r < range; ++r) count[r + 1] += count[r];
                for (int32_t k = 0; k < n; ++k) m.tile_perm[e0 + count[m.l_pose[e0 + k] - lo]++] = (uint16_t)k;
            } else {
                keys.resize(n);
                for (int32_t k = 0; k < n; ++k) keys[k] = ((uint64_t)(uint32_t)m.l_pose[e0 + k] << 16) | (uint64_t)k;
                std::sort(keys.begin(), keys.end());  // by pose, then l-order index
                for (int32_t s = 0; s < n; ++s) m.tile_perm[e0 + s] = (uint16_t)(keys[s] & 0xffff);
            }
        }
        // tile-local landmark of each l-order index (a tile holds at most
        // kTileLandmarks landmarks; a tile of one landmark has local index 0)
        lm_of_k.assign(n, 0);
        for (int32_t l = m.lm_tiles[t].lm_begin; l < m.lm_tiles[t].lm_end; ++l)
            for (int32_t k = m.lm_edge_ptr[l] - e0; k < m.lm_edge_ptr[l + 1] - e0; ++k)
                lm_of_k[k] = (uint8_t)(l - m.lm_tiles[t].lm_begin);
        m.lm_tiles[t].g_chunk = m.num_g32_chunks;
        for (int32_t s = 0; s < n; ++s) {
            const int32_t k = n > kTileEdges ? s : m.tile_perm[e0 + s];
            m.s_pose[e0 + s] = m.l_pose[e0 + k];
            m.s_lm[e0 + s] = lm_of_k[k];
            m.lm_store[e0 + k] = e0 + s;
            m.g32_pos[e0 + s] = m.num_g32_chunks * kTileEdges + s;
        }
        m.num_g32_chunks += std::max<int32_t>(1, (n + kTileEdges - 1) / kTileEdges);
        for (int32_t s = 0; s < n; ++s)
            if (s == n - 1 || m.s_pose[e0 + s + 1] != m.s_pose[e0 + s]) {
                run_pose.push_back(m.s_pose[e0 + s]);
                run_end.push_back(e0 + s);
            }
    }
    for (int32_t& pos : m.perm_cl) pos = m.lm_store[pos];
    m.num_runs = run_pose.size();
    m.pose_run_ptr.assign(m.num_pose_slots + 1, 0);
    for (int32_t p : run_pose) ++m.pose_run_ptr[p + 1];
    for (uint32_t s = 0; s < m.num_pose_slots; ++s) m.pose_run_ptr[s + 1] += m.pose_run_ptr[s];
    // slots by pose, each pose's runs in ascending storage order
    m.run_slot.assign(m.num_edges, -1);
    std::vector<int32_t> cursor(m.pose_run_ptr.begin(), m.pose_run_ptr.end() - 1);
    for (size_t r = 0; r < run_pose.size(); ++r) m.run_slot[run_end[r]] = cursor[run_pose[r]]++;
    const size_t tiles = m.lm_tiles.size();
    m.t_lm.assign(tiles * kTileEdges, 0);
    m.t_head.assign(tiles * (kTileEdges / 64), 0);
    m.t_rpose.assign(tiles * kTileRuns, 0);
    m.t_rslot.assign(tiles * kTileRuns, 0);
    for (size_t t = 0; t < tiles; ++t) {
        const LmTile& lt = m.lm_tiles[t];
        const int32_t e0 = lt.e_begin, n = lt.e_end - e0;
        if (n > kTileEdges || (lt.lm_end - lt.lm_begin == 1 && n > kTileRuns)) continue;
        int32_t r = -1;
        for (int32_t s = 0; s < n; ++s) {
            m.t_lm[t * kTileEdges + s] = m.s_lm[e0 + s];
            if (s == 0 || m.s_pose[e0 + s] != m.s_pose[e0 + s - 1]) {
                m.t_head[t * (kTileEdges / 64) + s / 64] |= 1ull << (s % 64);
                m.t_rpose[t * kTileRuns + ++r] = m.s_pose[e0 + s];
            }
            if (m.run_slot[e0 + s] >= 0) m.t_rslot[t * kTileRuns + r] = m.run_slot[e0 + s];
        }
    }
}

uint32_t next_pow2(uint64_t v) {
    uint32_t p = 1;
    while (p < v && p < (1u << 30)) p <<= 1;
    return p;
}

}  // namespace

bool camera_intrinsics(const vba_camera& c, double* fx, double* fy, double* cx, double* cy) {
    switch (c.model) {
        case VBA_CAMERA_PINHOLE:
        case VBA_CAMERA_OPENCV:
            if (c.num_params < 4) return false;
            *fx = c.params[0]; *fy = c.params[1]; *cx = c.params[2]; *cy = c.params[3];
            return true;
        case VBA_CAMERA_SIMPLE_PINHOLE:
        case VBA_CAMERA_SIMPLE_RADIAL:
        case VBA_CAMERA_RADIAL:
            if (c.num_params < 3) return false;
            *fx = *fy = c.params[0]; *cx = c.params[1]; *cy = c.params[2];
            return true;
        default:
            return false;
    }
}

bool lens_from_camera(const vba_camera& c, Lens<double>* L) {
    double fx, fy, cx, cy;
    if (!supported_model(c.model) || !camera_intrinsics(c, &fx, &fy, &cx, &cy)) return false;
    *L = Lens<double>{LENS_PINHOLE, 0, fx, fy, cx, cy, 0.0, 0.0, 0.0, 0.0};
    // Camera::radial_distortion / tangential_distortion (camera.rs:213-242).
    double k1 = 0.0, k2 = 0.0, p1 = 0.0, p2 = 0.0;
    switch (c.model) {
        case VBA_CAMERA_PINHOLE:
        case VBA_CAMERA_OPENCV:
            if (c.num_params > 4) k1 = c.params[4];
            if (c.num_params > 5) k2 = c.params[5];
            if (c.model == VBA_CAMERA_OPENCV) {
                if (c.num_params > 6) p1 = c.params[6];
                if (c.num_params > 7) p2 = c.params[7];
            }
            break;
        case VBA_CAMERA_SIMPLE_RADIAL:
            if (c.num_params > 3) k1 = c.params[3];
            break;
        case VBA_CAMERA_RADIAL:
            if (c.num_params > 3) k1 = c.params[3];
            if (c.num_params > 4) k2 = c.params[4];
            break;
        default:
            break;
    }
    L->k1 = k1;
    L->k2 = k2;
    L->p1 = p1;
    L->p2 = p2;
    if (p1 != 0.0 || p2 != 0.0) L->kind = LENS_OPENCV;
    else if (k1 != 0.0 || k2 != 0.0) L->kind = LENS_RADIAL;
    return true;
}

int32_t check_options(const vba_options& o, std::string* reason) {
    // The option half of GpuBundleAdjuster::check (crates/ba-gpu/src/solver.rs:212-251),
    // also applied to the options of every solve of a prepared problem or batch.
    if (o.refine_intrinsics) {
        *reason = "intrinsics refinement";
        return VBA_DECLINED;
    }
    if (o.robust_kernel > VBA_KERNEL_CAUCHY) {
        *reason = "robust kernel";
        return VBA_DECLINED;
    }
    if (o.linear_solver > VBA_SOLVER_PCG_EXPLICIT) {
        *reason = "linear solver";
        return VBA_DECLINED;
    }
    if (o.precision > VBA_PRECISION_MIXED) {
        *reason = "precision";
        return VBA_DECLINED;
    }
    if (o.pcg_maxit_policy > VBA_PCG_MAXIT_ACCEPT_ITERATE) {
        *reason = "PCG iteration-cap policy";
        return VBA_DECLINED;
    }
    return VBA_OK;
}

int32_t check_support(const vba_problem& pb, const vba_options& o, std::string* reason) {
    // GpuBundleAdjuster::check (crates/ba-gpu/src/solver.rs:212-251) widened to
    // the lens models, rig observations and kernels implemented here.
    const int32_t st = check_options(o, reason);
    if (st != VBA_OK) return st;
    if (pb.declined_features & (VBA_FEATURE_STEREO_OBSERVATIONS | VBA_FEATURE_GENERAL_STEREO)) {
        *reason = "stereo observations not passed in vba_problem.stereo";
        return VBA_DECLINED;
    }
    if (pb.declined_features & (VBA_FEATURE_PRIORS | VBA_FEATURE_INERTIAL)) {
        *reason = "prior/inertial terms";
        return VBA_DECLINED;
    }
    if (pb.declined_features & VBA_FEATURE_FIXED_ROTATIONS) {
        *reason = "rotation-fixed poses";
        return VBA_DECLINED;
    }
    if (pb.declined_features & VBA_FEATURE_OBSERVATION_WEIGHTS) {
        *reason = "observation weights";
        return VBA_DECLINED;
    }
    if (pb.declined_features) {
        *reason = "unknown declined feature";
        return VBA_DECLINED;
    }
    if (!supported_model(pb.camera.model)) {
        *reason = "camera model";
        return VBA_DECLINED;
    }
    for (uint64_t s = 0; s < pb.num_sensors; ++s)
        if (!supported_model(pb.sensors[s].camera.model)) {
            *reason = "rig camera model";
            return VBA_DECLINED;
        }
    const uint64_t ns = pb.stereo ? pb.stereo->num_stereo : 0, ng = pb.stereo ? pb.stereo->num_general : 0;
    const uint64_t npairs = ng ? pb.stereo->num_pairs : 0;
    if (pb.num_sensors + npairs > 65534) {
        *reason = "more than 65534 rig sensors and stereo pairs";
        return VBA_DECLINED;
    }
    if (pb.num_observations + ns + ng + pb.num_rig_observations > (uint64_t)INT32_MAX) {
        *reason = "more than 2^31 observations";
        return VBA_DECLINED;
    }
    return VBA_OK;
}

int32_t build_model(const vba_problem& pb, Model* mp, uint64_t* error_id, std::string* message) {
    Model& m = *mp;
    *error_id = 0;
    double fx, fy, cx, cy;
    // lm.rs:168-229: intrinsics, poses, observations, landmarks, ids.
    if (!supported_model(pb.camera.model) || !camera_intrinsics(pb.camera, &fx, &fy, &cx, &cy)) {
        *message = "camera model has no intrinsics";
        return VBA_ERR_UNSUPPORTED_CAMERA_MODEL;
    }
    const vba_stereo_terms* st = pb.stereo;
    const uint64_t ns = st ? st->num_stereo : 0, ng = st ? st->num_general : 0;
    // check_stereo_camera_model (problem.rs:733-746)
    if ((ns || ng) && pb.camera.model != VBA_CAMERA_PINHOLE && pb.camera.model != VBA_CAMERA_SIMPLE_PINHOLE) {
        *message = "stereo observations need a PINHOLE or SIMPLE_PINHOLE camera";
        return VBA_ERR_UNSUPPORTED_CAMERA_MODEL;
    }
    if (pb.num_poses == 0) return VBA_ERR_NO_POSES;
    if (pb.num_observations == 0 && ns == 0 && ng == 0 && pb.num_rig_observations == 0)
        return VBA_ERR_NO_OBSERVATIONS;
    if (pb.num_landmarks == 0) return VBA_ERR_NO_LANDMARKS;
    if (!pb.pose_ids || !pb.poses || !pb.landmark_ids || !pb.landmarks ||
        (pb.num_observations && (!pb.obs_pose_ids || !pb.obs_landmark_ids || !pb.obs_xy)) ||
        (pb.num_rig_observations &&
         (!pb.rig_pose_ids || !pb.rig_landmark_ids || !pb.rig_xy || !pb.rig_sensor || !pb.sensors)) ||
        (ns && (!st->stereo_pose_ids || !st->stereo_landmark_ids || !st->stereo_obs)) ||
        (ng && (!st->general_pose_ids || !st->general_landmark_ids || !st->general_obs || !st->general_pair ||
                !st->pairs))) {
        *message = "null array in problem";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    std::vector<uint64_t> pose_ids, lm_ids;
    if (!slot_order(pb.num_poses, pb.pose_ids, &m.pose_slot_of_input, &pose_ids)) {
        *message = "duplicate pose id";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    if (!slot_order(pb.num_landmarks, pb.landmark_ids, &m.lm_slot_of_input, &lm_ids)) {
        *message = "duplicate landmark id";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    m.num_pose_slots = (uint32_t)pb.num_poses;
    m.num_lm_slots = (uint32_t)pb.num_landmarks;
    m.poses.assign(7 * (size_t)pb.num_poses, 0.0);
    std::vector<uint8_t> pose_fixed(pb.num_poses, 0), lm_fixed(pb.num_landmarks, 0);
    for (uint64_t i = 0; i < pb.num_poses; ++i) {
        uint32_t s = m.pose_slot_of_input[i];
        std::memcpy(&m.poses[7 * (size_t)s], pb.poses + 7 * i, 7 * sizeof(double));
        pose_fixed[s] = pb.pose_fixed ? (pb.pose_fixed[i] != 0) : 0;
    }
    m.lms.assign(3 * (size_t)pb.num_landmarks, 0.0);
    for (uint64_t i = 0; i < pb.num_landmarks; ++i) {
        uint32_t s = m.lm_slot_of_input[i];
        std::memcpy(&m.lms[3 * (size_t)s], pb.landmarks + 3 * i, 3 * sizeof(double));
        lm_fixed[s] = pb.landmark_fixed ? (pb.landmark_fixed[i] != 0) : 0;
    }

    // Sensors: 0 = shared camera, 1 + s = rig sensor s.
    m.sensors.clear();
    DevSensor shared{};
    lens_from_camera(pb.camera, &shared.lens);
    shared.R[0] = shared.R[4] = shared.R[8] = 1.0;
    shared.identity = 1;
    m.sensors.push_back(shared);
    for (uint64_t s = 0; s < pb.num_sensors; ++s) {
        DevSensor d{};
        const vba_sensor& vs = pb.sensors[s];
        if (!lens_from_camera(vs.camera, &d.lens)) d.lens.kind = -1;  // checked per observation
        quat_to_matrix(vs.sensor_from_rig, d.R);
        d.t[0] = vs.sensor_from_rig[4];
        d.t[1] = vs.sensor_from_rig[5];
        d.t[2] = vs.sensor_from_rig[6];
        d.identity = 0;
        m.sensors.push_back(d);
    }
    // general stereo pairs: the right camera's pinhole intrinsics
    // (sqrt_qr.rs:471-476) and T_right<-left; checked per observation
    if (ng && pb.num_sensors + st->num_pairs > 65534) {
        *message = "more than 65534 rig sensors and stereo pairs";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    if (ng)
        for (uint64_t q = 0; q < pb.stereo->num_pairs; ++q) {
            DevSensor d{};
            const vba_sensor& vs = pb.stereo->pairs[q];
            d.lens.kind = LENS_PINHOLE;
            if (!camera_intrinsics(vs.camera, &d.lens.fx, &d.lens.fy, &d.lens.cx, &d.lens.cy)) d.lens.kind = -1;
            quat_to_matrix(vs.sensor_from_rig, d.R);
            d.t[0] = vs.sensor_from_rig[4];
            d.t[1] = vs.sensor_from_rig[5];
            d.t[2] = vs.sensor_from_rig[6];
            d.identity = 0;
            d.left = 0;
            m.sensors.push_back(d);
        }

    // Observations in the reference's flattened order (mono, rectified
    // stereo, general stereo, rig), ids resolved in that order with the
    // checks of lm.rs:193-229.
    const uint64_t nm = pb.num_observations, nr = pb.num_rig_observations;
    const uint64_t so = nm, go = nm + ns, ro = nm + ns + ng, n = ro + nr;
    std::vector<uint32_t> opose(n), olm(n);
    const IdIndex pose_slot(pose_ids), lm_slot(lm_ids);
    std::vector<uint16_t> osensor(n, 0);
    auto resolve = [&](const uint64_t* pids, const uint64_t* lids, uint64_t i, uint64_t k) -> int32_t {
        if (!pose_slot(pids[i], &opose[k])) {
            *error_id = pids[i];
            return VBA_ERR_MISSING_POSE;
        }
        if (!lm_slot(lids[i], &olm[k])) {
            *error_id = lids[i];
            return VBA_ERR_MISSING_LANDMARK;
        }
        return VBA_OK;
    };
    for (uint64_t i = 0; i < nm; ++i)
        if (int32_t e = resolve(pb.obs_pose_ids, pb.obs_landmark_ids, i, i)) return e;
    double baseline = 0.0;
    if (ns) {
        if (!(std::isfinite(st->baseline) && st->baseline > 0.0)) return VBA_ERR_MISSING_STEREO_BASELINE;
        baseline = st->baseline;
        for (uint64_t i = 0; i < ns; ++i)
            if (int32_t e = resolve(st->stereo_pose_ids, st->stereo_landmark_ids, i, so + i)) return e;
    }
    // a general pair is a sensor after the rig sensors: its right camera's
    // pinhole intrinsics and T_right<-left
    const uint32_t pair0 = 1 + (uint32_t)pb.num_sensors;
    if (ng) {
        for (uint64_t i = 0; i < ng; ++i) {
            if (int32_t e = resolve(st->general_pose_ids, st->general_landmark_ids, i, go + i)) return e;
            const uint32_t q = st->general_pair[i];
            if (q >= st->num_pairs) {
                *message = "stereo pair index out of range";
                return VBA_ERR_INVALID_ARGUMENT;
            }
            if (m.sensors[pair0 + q].lens.kind < 0) {
                *message = "right camera has no intrinsics";
                return VBA_ERR_UNSUPPORTED_CAMERA_MODEL;
            }
            osensor[go + i] = (uint16_t)(pair0 + q);
        }
    }
    for (uint64_t i = 0; i < nr; ++i) {
        uint64_t k = ro + i;
        if (!pose_slot(pb.rig_pose_ids[i], &opose[k])) {
            *error_id = pb.rig_pose_ids[i];
            return VBA_ERR_MISSING_POSE;
        }
        if (!lm_slot(pb.rig_landmark_ids[i], &olm[k])) {
            *error_id = pb.rig_landmark_ids[i];
            return VBA_ERR_MISSING_LANDMARK;
        }
        uint32_t s = pb.rig_sensor[i];
        if (s >= pb.num_sensors) {
            *message = "rig sensor index out of range";
            return VBA_ERR_INVALID_ARGUMENT;
        }
        if (m.sensors[s + 1].lens.kind < 0) {
            *message = "rig camera model has no intrinsics";
            return VBA_ERR_UNSUPPORTED_CAMERA_MODEL;
        }
        osensor[k] = (uint16_t)(s + 1);
    }
    m.any_rig = nr > 0;
    m.any_stereo = ns + ng > 0;
    // Landmark slots in order of their first, then last, observing pose slot
    // (ties in id order): the landmarks of a tile then span few cameras, so
    // its edges form few, long camera runs. Two stable counting sorts, by the
    // last pose and then by the first. Outputs map back through
    // lm_slot_of_input.
    {
        std::vector<uint32_t> first(m.num_lm_slots, m.num_pose_slots), last(m.num_lm_slots, 0);
        for (uint64_t i = 0; i < n; ++i) {
            first[olm[i]] = std::min(first[olm[i]], opose[i]);
            last[olm[i]] = std::max(last[olm[i]], opose[i]);
        }
        for (uint32_t l = 0; l < m.num_lm_slots; ++l)
            if (first[l] == m.num_pose_slots) last[l] = m.num_pose_slots;  // unobserved: at the end
        std::vector<uint32_t> start(m.num_pose_slots + 2, 0), by_last(m.num_lm_slots);
        for (uint32_t l = 0; l < m.num_lm_slots; ++l) ++start[last[l] + 1];
        for (uint32_t p = 0; p <= m.num_pose_slots; ++p) start[p + 1] += start[p];
        for (uint32_t l = 0; l < m.num_lm_slots; ++l) by_last[start[last[l]]++] = l;
        start.assign(m.num_pose_slots + 2, 0);
        for (uint32_t l = 0; l < m.num_lm_slots; ++l) ++start[first[l] + 1];
        for (uint32_t p = 0; p <= m.num_pose_slots; ++p) start[p + 1] += start[p];
        std::vector<uint32_t> new_slot(m.num_lm_slots);
        for (uint32_t l : by_last) new_slot[l] = start[first[l]]++;
        for (uint64_t i = 0; i < n; ++i) olm[i] = new_slot[olm[i]];
        for (uint32_t& s : m.lm_slot_of_input) s = new_slot[s];
        std::vector<double> lms(m.lms.size());
        std::vector<uint8_t> fixed(lm_fixed.size());
        for (uint32_t s = 0; s < m.num_lm_slots; ++s) {
            std::memcpy(&lms[3 * (size_t)new_slot[s]], &m.lms[3 * (size_t)s], 3 * sizeof(double));
            fixed[new_slot[s]] = lm_fixed[s];
        }
        m.lms.swap(lms);
        lm_fixed.swap(fixed);
    }
    m.pose_var.assign(m.num_pose_slots, -1);
    m.lm_var.assign(m.num_lm_slots, -1);
    m.P = m.L = 0;
    m.any_fixed_pose = false;
    for (uint32_t s = 0; s < m.num_pose_slots; ++s) {
        if (pose_fixed[s]) m.any_fixed_pose = true;
        else m.pose_var[s] = (int32_t)m.P++;
    }
    for (uint32_t s = 0; s < m.num_lm_slots; ++s)
        if (!lm_fixed[s]) m.lm_var[s] = (int32_t)m.L++;
    // reported after the whole layout is built: the joint camera solve
    // still runs with every pose and landmark fixed
    const bool all_fixed = m.P == 0 && m.L == 0;

    // c-order: (pose, landmark, input index) by two stable counting sorts,
    // by landmark and then by pose.
    m.num_obs = n;
    std::vector<uint32_t> cord(n);
    {
        std::vector<uint32_t> by_lm(n), cursor(std::max(m.num_lm_slots, m.num_pose_slots) + 1, 0);
        for (uint64_t i = 0; i < n; ++i) ++cursor[olm[i] + 1];
        for (uint32_t l = 0; l < m.num_lm_slots; ++l) cursor[l + 1] += cursor[l];
        for (uint64_t i = 0; i < n; ++i) by_lm[cursor[olm[i]]++] = (uint32_t)i;
        std::fill(cursor.begin(), cursor.end(), 0u);
        for (uint64_t i = 0; i < n; ++i) ++cursor[opose[i] + 1];
        for (uint32_t p = 0; p < m.num_pose_slots; ++p) cursor[p + 1] += cursor[p];
        for (uint64_t k = 0; k < n; ++k) {
            const uint32_t i = by_lm[k];
            cord[cursor[opose[i]]++] = i;
        }
    }
    m.obs_uv.resize(2 * n);
    m.obs_input = cord;
    const bool per_sensor = m.any_rig || m.any_stereo;
    if (per_sensor) m.obs_sensor.resize(n);
    else m.obs_sensor.clear();
    m.obs_kind.assign(m.any_stereo ? n : 0, OBS_PIXEL);
    m.obs_extra.assign(m.any_stereo ? 2 * n : 0, 0.0);
    for (uint64_t k = 0; k < n; ++k) {
        uint32_t i = cord[k];
        const double* xy;
        if (i < so) {
            xy = pb.obs_xy + 2 * i;
        } else if (i < go) {
            xy = st->stereo_obs + 3 * (i - so);
            m.obs_kind[k] = OBS_RECTIFIED;
            m.obs_extra[2 * k] = xy[2];
            m.obs_extra[2 * k + 1] = baseline;
        } else if (i < ro) {
            xy = st->general_obs + 4 * (i - go);
            m.obs_kind[k] = OBS_GENERAL;
            m.obs_extra[2 * k] = xy[2];
            m.obs_extra[2 * k + 1] = xy[3];
        } else {
            xy = pb.rig_xy + 2 * (i - ro);
        }
        m.obs_uv[2 * k] = xy[0];
        m.obs_uv[2 * k + 1] = xy[1];
        if (per_sensor) m.obs_sensor[k] = osensor[i];
    }
    // Edges: runs of equal (pose, landmark).
    m.edge_pose.clear();
    m.edge_lm.clear();
    m.edge_obs.clear();
    for (uint64_t k = 0; k < n; ++k) {
        uint32_t i = cord[k];
        if (k == 0 || opose[i] != opose[cord[k - 1]] || olm[i] != olm[cord[k - 1]]) {
            m.edge_pose.push_back((int32_t)opose[i]);
            m.edge_lm.push_back((int32_t)olm[i]);
            m.edge_obs.push_back((int32_t)k);
        }
    }
    m.num_edges = m.edge_pose.size();
    m.edge_obs.push_back((int32_t)n);
    m.cam_edge_ptr.assign(m.num_pose_slots + 1, 0);
    for (uint64_t e = 0; e < m.num_edges; ++e) ++m.cam_edge_ptr[m.edge_pose[e] + 1];
    for (uint32_t s = 0; s < m.num_pose_slots; ++s) m.cam_edge_ptr[s + 1] += m.cam_edge_ptr[s];

    // l-order: counting sort of c-order edges by landmark (stable: pose-sorted).
    m.lm_edge_ptr.assign(m.num_lm_slots + 1, 0);
    for (uint64_t e = 0; e < m.num_edges; ++e) ++m.lm_edge_ptr[m.edge_lm[e] + 1];
    for (uint32_t s = 0; s < m.num_lm_slots; ++s) m.lm_edge_ptr[s + 1] += m.lm_edge_ptr[s];
    m.perm_cl.resize(m.num_edges);
    m.l_pose.resize(m.num_edges);
    {
        std::vector<int32_t> cursor(m.lm_edge_ptr.begin(), m.lm_edge_ptr.end() - 1);
        for (uint64_t e = 0; e < m.num_edges; ++e) {
            int32_t pos = cursor[m.edge_lm[e]]++;
            m.perm_cl[e] = pos;
            m.l_pose[pos] = m.edge_pose[e];
        }
    }

    // Camera tiles: about 1,500 work units over the whole problem, each a
    // slice of one pose's edges (256..4096 edges).
    const uint32_t tc = std::min<uint32_t>(4096, std::max<uint32_t>(256, next_pow2(m.num_edges / 1536 + 1)));
    m.cam_tiles.clear();
    m.cam_tile_ptr.assign(m.num_pose_slots + 1, 0);
    for (uint32_t s = 0; s < m.num_pose_slots; ++s) {
        m.cam_tile_ptr[s] = (int32_t)m.cam_tiles.size();
        for (int32_t b = m.cam_edge_ptr[s]; b < m.cam_edge_ptr[s + 1]; b += (int32_t)tc)
            m.cam_tiles.push_back({(int32_t)s, b, std::min<int32_t>(b + (int32_t)tc, m.cam_edge_ptr[s + 1]), 0});
    }
    m.cam_tile_ptr[m.num_pose_slots] = (int32_t)m.cam_tiles.size();
    // Landmark tiles: runs of whole landmarks with at most kTileEdges edges,
    // kTileLandmarks landmarks and kTileRuns distinct poses; a landmark with
    // more than kTileRuns edges (each a distinct pose) forms its own tile.
    m.lm_tiles.clear();
    m.max_lm_tile_edges = 0;
    {
        std::vector<int32_t> mark(m.num_pose_slots, -1);  // tile that last counted the pose
        auto edges_of = [&](uint32_t l) { return m.lm_edge_ptr[l + 1] - m.lm_edge_ptr[l]; };
        for (uint32_t b = 0; b < m.num_lm_slots;) {
            const int32_t tile = (int32_t)m.lm_tiles.size();
            uint32_t e = b + 1;
            if (edges_of(b) <= kTileRuns) {
                int32_t poses = 0;
                for (int32_t k = m.lm_edge_ptr[b]; k < m.lm_edge_ptr[b + 1]; ++k)
                    if (mark[m.l_pose[k]] != tile) {
                        mark[m.l_pose[k]] = tile;
                        ++poses;
                    }
                while (e < m.num_lm_slots && e - b < (uint32_t)kTileLandmarks && edges_of(e) <= kTileRuns &&
                       m.lm_edge_ptr[e + 1] - m.lm_edge_ptr[b] <= kTileEdges) {
                    int32_t fresh = 0;
                    for (int32_t k = m.lm_edge_ptr[e]; k < m.lm_edge_ptr[e + 1]; ++k)
                        fresh += mark[m.l_pose[k]] != tile;
                    if (poses + fresh > kTileRuns) break;
                    for (int32_t k = m.lm_edge_ptr[e]; k < m.lm_edge_ptr[e + 1]; ++k) mark[m.l_pose[k]] = tile;
                    poses += fresh;
                    ++e;
                }
            }
            m.lm_tiles.push_back({(int32_t)b, (int32_t)e, m.lm_edge_ptr[b], m.lm_edge_ptr[e], 0});
            m.max_lm_tile_edges = std::max(m.max_lm_tile_edges, m.lm_edge_ptr[e] - m.lm_edge_ptr[b]);
            b = e;
        }
    }
    build_runs(&m);
    return all_fixed ? VBA_ERR_ALL_POSES_FIXED : VBA_OK;
}

}  // namespace vba
