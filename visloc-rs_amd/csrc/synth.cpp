// synth.cpp — seeded synthetic bundle-adjustment workloads (vba_synth.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vba_synth.h"
#include "geom.hpp"
#include "host_common.hpp"

namespace {

using vba::Lens;
using vba::Pose;
using vba::Quat;
using vba::V3;

// SplitMix64 (Steele, Lea, Flood 2014) with 53-bit uniforms and Box–Muller
// normals.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    uint32_t below(uint32_t n) { return (uint32_t)(uniform() * n) % n; }
    double normal() {
        double u1 = 1.0 - uniform();  // (0, 1]
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    // Number of failures before the first success, mean m (p = 1/(m+1)).
    uint32_t geometric(double mean) {
        if (mean <= 0.0) return 0;
        double p = 1.0 / (mean + 1.0);
        double u = 1.0 - uniform();
        return (uint32_t)std::floor(std::log(u) / std::log(1.0 - p));
    }
};

Quat quat_from_matrix(const double R[9]) {  // row-major, proper rotation
    double tr = R[0] + R[4] + R[8];
    Quat q;
    if (tr > 0.0) {
        double s = std::sqrt(tr + 1.0) * 2.0;
        q = {0.25 * s, (R[7] - R[5]) / s, (R[2] - R[6]) / s, (R[3] - R[1]) / s};
    } else if (R[0] > R[4] && R[0] > R[8]) {
        double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2.0;
        q = {(R[7] - R[5]) / s, 0.25 * s, (R[1] + R[3]) / s, (R[2] + R[6]) / s};
    } else if (R[4] > R[8]) {
        double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2.0;
        q = {(R[2] - R[6]) / s, (R[1] + R[3]) / s, 0.25 * s, (R[5] + R[7]) / s};
    } else {
        double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2.0;
        q = {(R[3] - R[1]) / s, (R[2] + R[6]) / s, (R[5] + R[7]) / s, 0.25 * s};
    }
    double n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    if (q.w < 0.0) n = -n;
    return {q.w / n, q.x / n, q.y / n, q.z / n};
}
V3 normalize(const V3& v) {
    double n = std::sqrt(vba::dot3(v, v));
    return {v.x / n, v.y / n, v.z / n};
}
// World z is up; the camera looks along `forward` with image y pointing down.
Pose look_along(const V3& centre, const V3& forward) {
    V3 z = normalize(forward);
    V3 down{0.0, 0.0, -1.0};
    double dz = vba::dot3(down, z);
    V3 y = normalize(V3{down.x - dz * z.x, down.y - dz * z.y, down.z - dz * z.z});
    V3 x = vba::cross3(y, z);
    // world->camera rotation rows are the camera axes in world coordinates
    double R[9] = {x.x, x.y, x.z, y.x, y.y, y.z, z.x, z.y, z.z};
    Quat q = quat_from_matrix(R);
    V3 rc = vba::qrot(q, centre);
    return {q, {-rc.x, -rc.y, -rc.z}};
}
V3 camera_centre(const Pose& T) { return vba::pose_inverse(T).t; }

Lens<double> lens_of(const vba_camera& c) {
    Lens<double> L{};
    switch (c.model) {
        case VBA_CAMERA_PINHOLE:
            L = {vba::LENS_PINHOLE, 0, c.params[0], c.params[1], c.params[2], c.params[3], 0, 0, 0, 0};
            if (c.num_params > 4) {
                L.k1 = c.params[4];
                L.k2 = c.num_params > 5 ? c.params[5] : 0.0;
            }
            break;
        case VBA_CAMERA_SIMPLE_RADIAL:
            L = {vba::LENS_PINHOLE, 0, c.params[0], c.params[0], c.params[1], c.params[2], c.params[3], 0, 0, 0};
            break;
        case VBA_CAMERA_OPENCV:
            L = {vba::LENS_OPENCV, 0, c.params[0], c.params[1], c.params[2], c.params[3],
                 c.params[4], c.params[5], c.params[6], c.params[7]};
            break;
        default:
            L = {vba::LENS_PINHOLE, 0, c.params[0], c.params[0], c.params[1], c.params[2], 0, 0, 0, 0};
    }
    if (L.kind == vba::LENS_PINHOLE && (L.k1 != 0.0 || L.k2 != 0.0)) L.kind = vba::LENS_RADIAL;
    return L;
}
// In-frame projection through a (pose, sensor extrinsic, lens).
bool in_frame(const Pose& T, const Pose* sensor, const vba_camera& cam, const V3& X, double uv[2]) {
    V3 p = vba::pose_apply(T, X);
    if (sensor) p = vba::pose_apply(*sensor, p);
    Lens<double> L = lens_of(cam);
    if (!(p.z > 0.05)) return false;
    if (!vba::project(L, p.x, p.y, p.z, uv, (double*)nullptr)) return false;
    return uv[0] >= 0.0 && uv[0] < (double)cam.width && uv[1] >= 0.0 && uv[1] < (double)cam.height;
}

vba_camera make_camera(int model, uint32_t w, uint32_t h, std::initializer_list<double> p) {
    vba_camera c;
    std::memset(&c, 0, sizeof(c));
    c.model = model;
    c.width = w;
    c.height = h;
    c.num_params = (uint32_t)p.size();
    int i = 0;
    for (double v : p) c.params[i++] = v;
    return c;
}

}  // namespace

struct vba_synth_problem {
    vba_synth_params params;
    vba_camera camera;
    std::vector<uint64_t> pose_ids, lm_ids;
    std::vector<double> poses, truth_poses;      // 7 per pose
    std::vector<uint8_t> pose_fixed;
    std::vector<double> lms, truth_lms;          // 3 per landmark
    std::vector<uint8_t> lm_fixed;
    std::vector<uint64_t> obs_pose, obs_lm;      // monocular
    std::vector<double> obs_xy;
    std::vector<uint64_t> rig_pose, rig_lm;      // rig
    std::vector<double> rig_xy;
    std::vector<uint32_t> rig_sensor;
    std::vector<vba_sensor> sensors;
    double extent = 0.0;
};

namespace {

void put_pose(std::vector<double>& v, const Pose& T) {
    v.insert(v.end(), {T.q.w, T.q.x, T.q.y, T.q.z, T.t.x, T.t.y, T.t.z});
}
Pose get_pose(const std::vector<double>& v, size_t i) {
    const double* p = v.data() + 7 * i;
    return {{p[0], p[1], p[2], p[3]}, {p[4], p[5], p[6]}};
}

// Perturb pose i: rotation error Exp(dw) applied on the camera side and a
// translation error of the camera centre; then landmarks.
void perturb(vba_synth_problem& S, Rng& rng) {
    const vba_synth_params& p = S.params;
    const size_t np = S.pose_ids.size();
    S.poses.clear();
    for (size_t i = 0; i < np; ++i) {
        Pose T = get_pose(S.truth_poses, i);
        if (S.pose_fixed[i] || (p.rotation_noise == 0.0 && p.translation_noise == 0.0)) {
            put_pose(S.poses, T);
            continue;
        }
        double xi[6] = {0, 0, 0, rng.normal() * p.rotation_noise, rng.normal() * p.rotation_noise,
                        rng.normal() * p.rotation_noise};
        Pose dR = vba::se3_exp(xi);
        V3 C = camera_centre(T);
        C = {C.x + rng.normal() * p.translation_noise, C.y + rng.normal() * p.translation_noise,
             C.z + rng.normal() * p.translation_noise};
        Quat q = vba::qmul(dR.q, T.q);
        V3 rc = vba::qrot(q, C);
        put_pose(S.poses, Pose{q, {-rc.x, -rc.y, -rc.z}});
    }
    S.lms = S.truth_lms;
    for (size_t j = 0; j < S.lm_ids.size(); ++j) {
        if (S.lm_fixed[j]) continue;
        for (int k = 0; k < 3; ++k) S.lms[3 * j + k] += rng.normal() * p.landmark_noise;
    }
}
void finish_extent(vba_synth_problem& S) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t j = 0; j < S.lm_ids.size(); ++j)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], S.truth_lms[3 * j + k]);
            hi[k] = std::max(hi[k], S.truth_lms[3 * j + k]);
        }
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) d2 += (hi[k] - lo[k]) * (hi[k] - lo[k]);
    S.extent = std::sqrt(d2);
}
double noisy(Rng& rng, double v, double sigma) { return v + rng.normal() * sigma; }

// Observation record before sorting into camera-major order.
struct RawObs {
    uint32_t cam, lm, sensor;
    double u, v;
};
void emit_camera_major(vba_synth_problem& S, std::vector<RawObs>& raw, bool rig) {
    std::stable_sort(raw.begin(), raw.end(), [](const RawObs& a, const RawObs& b) { return a.cam < b.cam; });
    for (const RawObs& o : raw) {
        if (rig) {
            S.rig_pose.push_back(S.pose_ids[o.cam]);
            S.rig_lm.push_back(S.lm_ids[o.lm]);
            S.rig_xy.push_back(o.u);
            S.rig_xy.push_back(o.v);
            S.rig_sensor.push_back(o.sensor);
        } else {
            S.obs_pose.push_back(S.pose_ids[o.cam]);
            S.obs_lm.push_back(S.lm_ids[o.lm]);
            S.obs_xy.push_back(o.u);
            S.obs_xy.push_back(o.v);
        }
    }
}
void apply_outliers(Rng& rng, const vba_synth_params& p, const vba_camera& cam, RawObs& o) {
    if (p.outlier_fraction > 0.0 && rng.uniform() < p.outlier_fraction) {
        o.u = rng.uniform(0.0, (double)cam.width);
        o.v = rng.uniform(0.0, (double)cam.height);
    }
}

void init_ids(vba_synth_problem& S, uint32_t ncam, uint64_t nlm) {
    S.pose_ids.resize(ncam);
    S.pose_fixed.assign(ncam, 0);
    for (uint32_t i = 0; i < ncam; ++i) {
        S.pose_ids[i] = 10ull * (i + 1);
        S.pose_fixed[i] = i < S.params.fixed_cameras;
    }
    S.lm_ids.resize(nlm);
    S.lm_fixed.assign(nlm, 0);
    for (uint64_t j = 0; j < nlm; ++j) S.lm_ids[j] = j + 1;
}

// Cameras on a circle facing its centre; every in-frame landmark observed.
bool gen_circle(vba_synth_problem& S) {
    const vba_synth_params& p = S.params;
    Rng rng(p.seed);
    const uint32_t n = p.num_cameras;
    S.camera = make_camera(VBA_CAMERA_SIMPLE_RADIAL, p.width, p.height,
                           {p.focal, p.width / 2.0, p.height / 2.0, p.distortion});
    init_ids(S, n, p.num_landmarks);
    const double r = p.scene_scale;
    for (uint32_t i = 0; i < n; ++i) {
        double th = 2.0 * M_PI * i / n;
        V3 C{r * std::cos(th), r * std::sin(th), 0.0};
        put_pose(S.truth_poses, look_along(C, V3{-C.x, -C.y, -C.z}));
    }
    const double half = 0.2 * r;  // 4 m cube at radius 10 m
    for (uint64_t j = 0; j < p.num_landmarks; ++j)
        for (int k = 0; k < 3; ++k) S.truth_lms.push_back(rng.uniform(-half, half));
    std::vector<RawObs> raw;
    for (uint32_t i = 0; i < n; ++i) {
        Pose T = get_pose(S.truth_poses, i);
        for (uint64_t j = 0; j < p.num_landmarks; ++j) {
            V3 X{S.truth_lms[3 * j], S.truth_lms[3 * j + 1], S.truth_lms[3 * j + 2]};
            double uv[2];
            if (!in_frame(T, nullptr, S.camera, X, uv)) continue;
            RawObs o{i, (uint32_t)j, 0, noisy(rng, uv[0], p.pixel_noise), noisy(rng, uv[1], p.pixel_noise)};
            apply_outliers(rng, p, S.camera, o);
            raw.push_back(o);
        }
    }
    emit_camera_major(S, raw, false);
    perturb(S, rng);
    finish_extent(S);
    return true;
}

// Forward-facing cameras on a closed loop; each landmark seen by a run of
// consecutive cameras, placed in view of the run's last camera inside a
// corridor around the loop.
bool gen_loop(vba_synth_problem& S) {
    const vba_synth_params& p = S.params;
    Rng rng(p.seed);
    const uint32_t n = p.num_cameras;
    S.camera = make_camera(VBA_CAMERA_PINHOLE, p.width, p.height,
                           {p.focal, p.focal, p.width / 2.0, p.height / 2.0});
    init_ids(S, n, p.num_landmarks);
    const double R = p.scene_scale / (2.0 * M_PI);
    const double half_corridor = 10.0;
    std::vector<Pose> T(n);
    for (uint32_t i = 0; i < n; ++i) {
        double ph = 2.0 * M_PI * i / n;
        V3 C{R * std::cos(ph), R * std::sin(ph), 0.0};
        V3 f{-std::sin(ph), std::cos(ph), 0.0};
        T[i] = look_along(C, f);
        put_pose(S.truth_poses, T[i]);
    }
    const double tx = 0.45 * p.width / p.focal, ty = 0.45 * p.height / p.focal;
    std::vector<RawObs> raw;
    raw.reserve((size_t)(p.num_landmarks * (p.mean_track_length + 1.0)));
    std::vector<std::pair<double, double>> uvs;
    for (uint64_t j = 0; j < p.num_landmarks; ++j) {
        uint32_t k = p.min_track_length + rng.geometric(p.mean_track_length - p.min_track_length);
        k = std::min(std::max(k, p.min_track_length), std::min(p.max_track_length, n));
        uint32_t s = rng.below(n);
        V3 X{0, 0, 0};
        bool ok = false;
        for (int attempt = 0; attempt < 64 && !ok; ++attempt) {
            uint32_t e = (s + k - 1) % n;
            Pose Ti = vba::pose_inverse(T[e]);
            double d = rng.uniform(4.0, 30.0);
            V3 pc{rng.uniform(-tx, tx) * d, rng.uniform(-ty, ty) * d, d};
            X = vba::pose_apply(Ti, pc);
            double rad = std::sqrt(X.x * X.x + X.y * X.y);
            if (std::fabs(rad - R) > half_corridor) continue;
            uvs.clear();
            ok = true;
            for (uint32_t t = 0; t < k && ok; ++t) {
                double uv[2];
                ok = in_frame(T[(s + t) % n], nullptr, S.camera, X, uv);
                uvs.push_back({uv[0], uv[1]});
            }
            if (!ok && k > p.min_track_length && attempt > 32) --k;
        }
        if (!ok) {  // fall back to a point straight ahead of the last camera
            uint32_t e = (s + k - 1) % n;
            X = vba::pose_apply(vba::pose_inverse(T[e]), V3{0.0, 0.0, 15.0});
            uvs.clear();
            for (uint32_t t = 0; t < k; ++t) {
                double uv[2];
                if (in_frame(T[(s + t) % n], nullptr, S.camera, X, uv)) uvs.push_back({uv[0], uv[1]});
                else uvs.push_back({-1.0, -1.0});
            }
        }
        S.truth_lms.insert(S.truth_lms.end(), {X.x, X.y, X.z});
        for (uint32_t t = 0; t < k; ++t) {
            if (uvs[t].first < 0.0) continue;
            RawObs o{(s + t) % n, (uint32_t)j, 0, noisy(rng, uvs[t].first, p.pixel_noise),
                     noisy(rng, uvs[t].second, p.pixel_noise)};
            apply_outliers(rng, p, S.camera, o);
            raw.push_back(o);
        }
    }
    emit_camera_major(S, raw, false);
    perturb(S, rng);
    finish_extent(S);
    return true;
}

// Stereo-rig window: keyframes moving forward, each landmark seen by a run of
// 2..min(10, K) keyframes in both sensors.
bool gen_stereo_window(vba_synth_problem& S) {
    const vba_synth_params& p = S.params;
    Rng rng(p.seed);
    const uint32_t n = p.num_cameras;
    S.camera = make_camera(VBA_CAMERA_PINHOLE, p.width, p.height,
                           {p.focal, p.focal, p.width / 2.0, p.height / 2.0});
    init_ids(S, n, p.num_landmarks);
    vba_sensor left, right;
    std::memset(&left, 0, sizeof(left));
    left.camera = S.camera;
    left.sensor_from_rig[0] = 1.0;
    right = left;
    right.sensor_from_rig[4] = -p.scene_scale;  // baseline: right camera at +b x
    S.sensors = {left, right};
    Pose sens[2] = {{{1, 0, 0, 0}, {0, 0, 0}}, {{1, 0, 0, 0}, {-p.scene_scale, 0, 0}}};
    std::vector<Pose> T(n);
    double heading = 0.0;
    V3 C{0, 0, 1.5};
    for (uint32_t i = 0; i < n; ++i) {
        heading += rng.normal() * 0.02;
        V3 f{std::cos(heading), std::sin(heading), 0.0};
        T[i] = look_along(C, f);
        put_pose(S.truth_poses, T[i]);
        C = {C.x + 0.3 * f.x, C.y + 0.3 * f.y, C.z};
    }
    const double tx = 0.45 * p.width / p.focal, ty = 0.45 * p.height / p.focal;
    std::vector<RawObs> raw;
    for (uint64_t j = 0; j < p.num_landmarks; ++j) {
        uint32_t kmax = std::min(p.max_track_length, n);
        uint32_t k = p.min_track_length + rng.below(kmax - p.min_track_length + 1);
        uint32_t s = rng.below(n - k + 1);
        V3 X{0, 0, 0};
        bool ok = false;
        double uv[20][2][2];
        for (int attempt = 0; attempt < 64 && !ok; ++attempt) {
            Pose Ti = vba::pose_inverse(T[s + k - 1]);
            double d = rng.uniform(2.0, 20.0);
            X = vba::pose_apply(Ti, V3{rng.uniform(-tx, tx) * d, rng.uniform(-ty, ty) * d, d});
            ok = true;
            for (uint32_t t = 0; t < k && ok; ++t)
                for (int c = 0; c < 2 && ok; ++c) ok = in_frame(T[s + t], &sens[c], S.camera, X, uv[t][c]);
            if (!ok && attempt > 32 && k > p.min_track_length) --k;
        }
        if (!ok) {
            k = 1;
            X = vba::pose_apply(vba::pose_inverse(T[s]), V3{0.0, 0.0, 8.0});
            for (int c = 0; c < 2; ++c) in_frame(T[s], &sens[c], S.camera, X, uv[0][c]);
        }
        S.truth_lms.insert(S.truth_lms.end(), {X.x, X.y, X.z});
        for (uint32_t t = 0; t < k; ++t)
            for (uint32_t c = 0; c < 2; ++c) {
                RawObs o{s + t, (uint32_t)j, c, noisy(rng, uv[t][c][0], p.pixel_noise),
                         noisy(rng, uv[t][c][1], p.pixel_noise)};
                apply_outliers(rng, p, S.camera, o);
                raw.push_back(o);
            }
    }
    emit_camera_major(S, raw, true);
    perturb(S, rng);
    finish_extent(S);
    return true;
}

// One variable camera looking at fixed points.
bool gen_pose_only(vba_synth_problem& S) {
    const vba_synth_params& p = S.params;
    Rng rng(p.seed);
    S.camera = make_camera(VBA_CAMERA_PINHOLE, p.width, p.height,
                           {p.focal, p.focal, p.width / 2.0, p.height / 2.0});
    init_ids(S, 1, p.num_landmarks);
    S.pose_fixed[0] = 0;
    Pose T = look_along(V3{0, 0, 0}, V3{1, 0, 0});
    put_pose(S.truth_poses, T);
    Pose Ti = vba::pose_inverse(T);
    const double tx = 0.45 * p.width / p.focal, ty = 0.45 * p.height / p.focal;
    std::vector<RawObs> raw;
    for (uint64_t j = 0; j < p.num_landmarks; ++j) {
        double d = rng.uniform(3.0, 30.0);
        V3 X = vba::pose_apply(Ti, V3{rng.uniform(-tx, tx) * d, rng.uniform(-ty, ty) * d, d});
        S.truth_lms.insert(S.truth_lms.end(), {X.x, X.y, X.z});
        S.lm_fixed[j] = 1;
        double uv[2];
        in_frame(T, nullptr, S.camera, X, uv);
        RawObs o{0, (uint32_t)j, 0, noisy(rng, uv[0], p.pixel_noise), noisy(rng, uv[1], p.pixel_noise)};
        apply_outliers(rng, p, S.camera, o);
        raw.push_back(o);
    }
    emit_camera_major(S, raw, false);
    perturb(S, rng);
    finish_extent(S);
    return true;
}

// City block: a square block of side `scene_scale` centred at the origin,
// cameras on a square orbit 10 m outside its facades at 1.5 m height facing
// the block; landmarks on the facades (height 0..20 m) and the street.
bool gen_city(vba_synth_problem& S) {
    const vba_synth_params& p = S.params;
    Rng rng(p.seed);
    const uint32_t n = p.num_cameras;
    S.camera = make_camera(VBA_CAMERA_PINHOLE, p.width, p.height,
                           {p.focal, p.focal, p.width / 2.0, p.height / 2.0});
    init_ids(S, n, p.num_landmarks);
    const double h = 0.5 * p.scene_scale, orbit = h + 10.0;
    std::vector<Pose> T(n);
    S.sensors.resize(n);
    std::vector<vba_camera> cams(n);
    for (uint32_t i = 0; i < n; ++i) {
        double s = 8.0 * orbit * i / n;  // arc length along the square orbit
        int side = (int)(s / (2.0 * orbit));
        double u = s - side * 2.0 * orbit - orbit;
        V3 C, f;
        switch (side) {
            case 0: C = {u, -orbit, 1.5}; f = {0, 1, 0}; break;
            case 1: C = {orbit, u, 1.5}; f = {-1, 0, 0}; break;
            case 2: C = {-u, orbit, 1.5}; f = {0, -1, 0}; break;
            default: C = {-orbit, -u, 1.5}; f = {1, 0, 0}; break;
        }
        T[i] = look_along(C, f);
        put_pose(S.truth_poses, T[i]);
        double fi = p.focal * (1.0 + rng.uniform(-0.05, 0.05));
        cams[i] = make_camera(VBA_CAMERA_OPENCV, p.width, p.height,
                              {fi, fi, p.width / 2.0, p.height / 2.0, rng.normal() * p.distortion,
                               rng.normal() * p.distortion, rng.normal() * p.distortion * 0.1,
                               rng.normal() * p.distortion * 0.1});
        std::memset(&S.sensors[i], 0, sizeof(vba_sensor));
        S.sensors[i].camera = cams[i];
        S.sensors[i].sensor_from_rig[0] = 1.0;
    }
    std::vector<RawObs> raw;
    raw.reserve((size_t)(p.num_landmarks * (p.mean_track_length + 1.0)));
    Pose ident{{1, 0, 0, 0}, {0, 0, 0}};
    double uvbuf[64][2];
    for (uint64_t j = 0; j < p.num_landmarks; ++j) {
        uint32_t k = p.min_track_length + rng.geometric(p.mean_track_length - p.min_track_length);
        k = std::min(std::max(k, p.min_track_length), std::min(p.max_track_length, 64u));
        uint32_t s0 = rng.below(n);
        uint32_t mid = (s0 + k / 2) % n;
        V3 X{0, 0, 0};
        bool ok = false;
        for (int attempt = 0; attempt < 64 && !ok; ++attempt) {
            // a ray through a random pixel of the middle camera, hit against
            // the facade plane it faces or the ground
            Pose Ti = vba::pose_inverse(T[mid]);
            V3 C = Ti.t;
            V3 dir_c{rng.uniform(-0.45, 0.45) * p.width / p.focal, rng.uniform(-0.45, 0.45) * p.height / p.focal, 1.0};
            V3 dir = vba::qrot(Ti.q, dir_c);
            double tbest = 1e300;
            if (dir.z < -1e-6) tbest = std::min(tbest, -C.z / dir.z);  // ground
            const double planes[4][2] = {{0, -h}, {0, h}, {1, -h}, {1, h}};
            for (auto& pl : planes) {
                double c0 = pl[0] == 0 ? C.y : C.x, d0 = pl[0] == 0 ? dir.y : dir.x;
                if (std::fabs(d0) < 1e-9) continue;
                double t = (pl[1] - c0) / d0;
                if (t <= 0.5) continue;
                V3 Y{C.x + t * dir.x, C.y + t * dir.y, C.z + t * dir.z};
                double other = pl[0] == 0 ? Y.x : Y.y;
                if (std::fabs(other) <= h && Y.z >= 0.0 && Y.z <= 20.0) tbest = std::min(tbest, t);
            }
            if (tbest > 60.0) continue;
            X = {C.x + tbest * dir.x, C.y + tbest * dir.y, C.z + tbest * dir.z};
            ok = true;
            for (uint32_t t = 0; t < k && ok; ++t) ok = in_frame(T[(s0 + t) % n], &ident, cams[(s0 + t) % n], X, uvbuf[t]);
            if (!ok && attempt > 32 && k > p.min_track_length) --k;
        }
        if (!ok) {
            k = 1;
            X = vba::pose_apply(vba::pose_inverse(T[s0]), V3{0.0, 0.0, 10.0});
            if (!in_frame(T[s0], &ident, cams[s0], X, uvbuf[0])) uvbuf[0][0] = uvbuf[0][1] = 100.0;
        }
        S.truth_lms.insert(S.truth_lms.end(), {X.x, X.y, X.z});
        for (uint32_t t = 0; t < k; ++t) {
            uint32_t c = (s0 + t) % n;
            RawObs o{c, (uint32_t)j, c, noisy(rng, uvbuf[t][0], p.pixel_noise), noisy(rng, uvbuf[t][1], p.pixel_noise)};
            apply_outliers(rng, p, cams[c], o);
            raw.push_back(o);
        }
    }
    emit_camera_major(S, raw, true);
    perturb(S, rng);
    finish_extent(S);
    return true;
}

}  // namespace

extern "C" {

void vba_synth_default(uint32_t kind, vba_synth_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->seed = kind;
    p->rotation_noise = 0.01;
    p->translation_noise = 0.05;
    p->landmark_noise = 0.05;
    p->fixed_cameras = 1;
    p->min_track_length = 2;
    p->max_track_length = 64;
    switch (kind) {
        case VBA_SYNTH_CIRCLE:
            p->num_cameras = 50;
            p->num_landmarks = 5000;
            p->pixel_noise = 0.5;
            p->scene_scale = 10.0;
            p->focal = 500.0;
            p->distortion = 0.01;
            p->width = 640;
            p->height = 480;
            break;
        case VBA_SYNTH_LOOP:
            p->num_cameras = 1200;
            p->num_landmarks = 400000;
            p->pixel_noise = 0.7;
            p->mean_track_length = 8.75;
            p->scene_scale = 200.0;
            p->focal = 800.0;
            p->width = 1024;
            p->height = 768;
            break;
        case VBA_SYNTH_STEREO_WINDOW:
            p->num_cameras = 20;
            p->num_landmarks = 4000;
            p->pixel_noise = 0.5;
            p->outlier_fraction = 0.02;
            p->min_track_length = 2;
            p->max_track_length = 10;
            p->scene_scale = 0.11;
            p->focal = 460.0;
            p->width = 752;
            p->height = 480;
            break;
        case VBA_SYNTH_POSE_ONLY:
            p->num_cameras = 1;
            p->num_landmarks = 1000;
            p->pixel_noise = 1.0;
            p->outlier_fraction = 0.05;
            p->landmark_noise = 0.0;
            p->focal = 500.0;
            p->width = 640;
            p->height = 480;
            p->fixed_cameras = 0;
            break;
        case VBA_SYNTH_CITY:
            p->num_cameras = 5500;
            p->num_landmarks = 2500000;
            p->pixel_noise = 0.8;
            p->outlier_fraction = 0.10;
            p->mean_track_length = 8.0;
            p->scene_scale = 100.0;
            p->focal = 1100.0;
            p->distortion = 0.02;
            p->width = 1920;
            p->height = 1080;
            break;
        default:
            break;
    }
}

vba_synth_problem* vba_synth_generate(const vba_synth_params* params) {
    if (!params || params->num_landmarks == 0 || params->num_cameras == 0) {
        vba::set_thread_error("vba_synth_generate: invalid parameters");
        return nullptr;
    }
    auto S = std::make_unique<vba_synth_problem>();
    S->params = *params;
    bool ok = false;
    switch (params->kind) {
        case VBA_SYNTH_CIRCLE: ok = gen_circle(*S); break;
        case VBA_SYNTH_LOOP: ok = gen_loop(*S); break;
        case VBA_SYNTH_STEREO_WINDOW:
            ok = params->num_cameras >= params->min_track_length && params->min_track_length >= 1 &&
                 params->max_track_length <= 20 && gen_stereo_window(*S);
            break;
        case VBA_SYNTH_POSE_ONLY: ok = gen_pose_only(*S); break;
        case VBA_SYNTH_CITY: ok = gen_city(*S); break;
        default: break;
    }
    if (!ok) {
        vba::set_thread_error("vba_synth_generate: unknown kind or inconsistent sizes");
        return nullptr;
    }
    return S.release();
}

void vba_synth_free(vba_synth_problem* problem) { delete problem; }

void vba_synth_view(const vba_synth_problem* S, vba_problem* v) {
    std::memset(v, 0, sizeof(*v));
    v->camera = S->camera;
    v->num_poses = S->pose_ids.size();
    v->pose_ids = S->pose_ids.data();
    v->poses = S->poses.data();
    v->pose_fixed = S->pose_fixed.data();
    v->num_landmarks = S->lm_ids.size();
    v->landmark_ids = S->lm_ids.data();
    v->landmarks = S->lms.data();
    v->landmark_fixed = S->lm_fixed.data();
    v->num_observations = S->obs_pose.size();
    v->obs_pose_ids = S->obs_pose.data();
    v->obs_landmark_ids = S->obs_lm.data();
    v->obs_xy = S->obs_xy.data();
    v->num_rig_observations = S->rig_pose.size();
    v->rig_pose_ids = S->rig_pose.data();
    v->rig_landmark_ids = S->rig_lm.data();
    v->rig_xy = S->rig_xy.data();
    v->rig_sensor = S->rig_sensor.data();
    v->num_sensors = S->sensors.size();
    v->sensors = S->sensors.data();
}

void vba_synth_truth(const vba_synth_problem* S, const double** poses, const double** landmarks) {
    if (poses) *poses = S->truth_poses.data();
    if (landmarks) *landmarks = S->truth_lms.data();
}

double vba_synth_scene_extent(const vba_synth_problem* S) { return S->extent; }

}  // extern "C"
