// ba_oracle.cpp — CPU FP64 restatement of the reference bundle adjuster.
//
// TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline): see
// ba_oracle.h. It restates, function by function, the Levenberg–Marquardt
// Schur-complement path of visloc-rs:
//   pipelines/slam/src/bundle/lm.rs:162-1017              (LM loop)
//   pipelines/slam/src/bundle/normal_equations.rs:352-651 (assembly)
//   pipelines/slam/src/bundle/normal_equations.rs:1314-1592 (dense Schur)
//   pipelines/slam/src/bundle/implicit_schur.rs:133-824   (matrix-free PCG)
//   pipelines/slam/src/bundle/matrix_free.rs:1106-1240
//   pipelines/slam/src/bundle/problem.rs:252-303, 497-711 (cost, feasibility)
//   pipelines/slam/src/bundle/jacobians.rs:29-124         (projection)
//   pipelines/slam/src/bundle/sqrt_qr.rs:445-463, 960-1005 (rig residual)
//   pipelines/slam/src/pose_graph.rs:87-140, 2259-2270    (kernels, solve)
//   crates/core/src/geometry/se3.rs, pose.rs
//   crates/core/src/types/camera.rs:213-433, 844-941
// Arithmetic follows the evaluation order of those sources and of the
// nalgebra 0.33.3 routines they call (quaternion product and rotation,
// 3x3 cofactor inverse, column Cholesky, partial-pivot LU, the 8-way
// unrolled dynamic dot product), so that a change of operation order shows up
// in the tests. Bit equality with the Rust binary itself cannot be checked
// here (no Rust toolchain); the known-answer tests pin what they can.
//
// Extension (not in the reference): the matrix-free path accepts every lens
// model of the dense path (the reference's matrix-free validation admits
// only distortion-free PINHOLE / SIMPLE_PINHOLE, optimize.rs:462-485). For
// those two models the behaviour is the reference's.

#include "ba_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <map>
#include <vector>

namespace {

constexpr double kEps = 1e-12;  // camera.rs:5

// ---------------------------------------------------------------------------
// Fixed-size row-major matrices; products sum over k left to right, as
// nalgebra's gemv/axpy column sweep does for statically sized operands.
// ---------------------------------------------------------------------------
template <int R, int C>
struct Mat {
    double a[R * C];
    double& operator()(int r, int c) { return a[r * C + c]; }
    double operator()(int r, int c) const { return a[r * C + c]; }
    static Mat zeros() {
        Mat m;
        for (int i = 0; i < R * C; ++i) m.a[i] = 0.0;
        return m;
    }
};
using M3 = Mat<3, 3>;
using M23 = Mat<2, 3>;
using M26 = Mat<2, 6>;
using M36 = Mat<3, 6>;
using M46 = Mat<4, 6>;
using M43 = Mat<4, 3>;
using M63 = Mat<6, 3>;
using M66 = Mat<6, 6>;

// Runs fn(begin, end) over [0, n) in `threads` contiguous ranges, one thread
// each. Every index is handled by exactly one thread with the serial code's
// operations in the serial order, so a result whose entries do not depend
// on one another has the serial bits (the parallel == serial property of
// parallel_ba_tests.rs:113-160).
template <class F>
void parallel_ranges(int threads, size_t n, F&& fn) {
    if (threads <= 1 || n < 2) {
        fn((size_t)0, n);
        return;
    }
    const size_t t = std::min<size_t>((size_t)threads, n);
    std::vector<std::thread> pool;
    pool.reserve(t);
    for (size_t k = 0; k < t; ++k) {
        const size_t b = n * k / t, e = n * (k + 1) / t;
        pool.emplace_back([&fn, b, e]() { fn(b, e); });
    }
    for (auto& th : pool) th.join();
}

template <int R, int K, int C>
Mat<R, C> mul(const Mat<R, K>& A, const Mat<K, C>& B) {
    Mat<R, C> out;
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            double acc = A(r, 0) * B(0, c);
            for (int k = 1; k < K; ++k) acc = A(r, k) * B(k, c) + acc;
            out(r, c) = acc;
        }
    return out;
}
template <int R, int C>
Mat<C, R> tr(const Mat<R, C>& A) {
    Mat<C, R> out;
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) out(c, r) = A(r, c);
    return out;
}
template <int R, int C>
Mat<R, C> scale(double s, const Mat<R, C>& A) {
    Mat<R, C> out;
    for (int i = 0; i < R * C; ++i) out.a[i] = s * A.a[i];
    return out;
}
template <int R, int C>
Mat<R, C> neg(const Mat<R, C>& A) {
    Mat<R, C> out;
    for (int i = 0; i < R * C; ++i) out.a[i] = -A.a[i];
    return out;
}
template <int R, int C>
Mat<R, C> add(const Mat<R, C>& A, const Mat<R, C>& B) {
    Mat<R, C> out;
    for (int i = 0; i < R * C; ++i) out.a[i] = A.a[i] + B.a[i];
    return out;
}
template <int R, int C>
void add_assign(Mat<R, C>& A, const Mat<R, C>& B) {
    for (int i = 0; i < R * C; ++i) A.a[i] = A.a[i] + B.a[i];
}
template <int R, int C>
void sub_assign(Mat<R, C>& A, const Mat<R, C>& B) {
    for (int i = 0; i < R * C; ++i) A.a[i] = A.a[i] - B.a[i];
}
template <int R, int C>
bool all_finite(const Mat<R, C>& A) {
    for (int i = 0; i < R * C; ++i)
        if (!std::isfinite(A.a[i])) return false;
    return true;
}
// Matrix (R x K) times vector (K).
template <int R, int K>
void mulv(const Mat<R, K>& A, const double* x, double* y) {
    for (int r = 0; r < R; ++r) {
        double acc = A(r, 0) * x[0];
        for (int k = 1; k < K; ++k) acc = A(r, k) * x[k] + acc;
        y[r] = acc;
    }
}

struct V3 {
    double x, y, z;
};
inline V3 cross(const V3& a, const V3& b) {  // nalgebra Vector3::cross
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm_sq3(const V3& v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline M3 skew(const V3& v) {  // se3.rs:102-104, sqrt_qr.rs:1003-1005
    M3 m;
    m(0, 0) = 0.0; m(0, 1) = -v.z; m(0, 2) = v.y;
    m(1, 0) = v.z; m(1, 1) = 0.0; m(1, 2) = -v.x;
    m(2, 0) = -v.y; m(2, 1) = v.x; m(2, 2) = 0.0;
    return m;
}
inline M3 identity3() {
    M3 m = M3::zeros();
    m(0, 0) = m(1, 1) = m(2, 2) = 1.0;
    return m;
}

// nalgebra DVector::dot: 8-accumulator unrolled sum (generic dotx path).
double dyn_dot(const double* a, const double* b, size_t n) {
    double res = 0.0;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    size_t i = 0;
    while (n - i >= 8) {
        for (int k = 0; k < 8; ++k) acc[k] += a[i + k] * b[i + k];
        i += 8;
    }
    res += acc[0] + acc[4];
    res += acc[1] + acc[5];
    res += acc[2] + acc[6];
    res += acc[3] + acc[7];
    for (size_t k = i; k < n; ++k) res += a[k] * b[k];
    return res;
}
double dyn_norm(const std::vector<double>& v) { return std::sqrt(dyn_dot(v.data(), v.data(), v.size())); }
bool all_finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

// ---------------------------------------------------------------------------
// Unit quaternion and SE3 (crates/core/src/geometry/se3.rs).
// ---------------------------------------------------------------------------
struct Quat {
    double w, i, j, k;
};
inline Quat qmul(const Quat& a, const Quat& b) {  // nalgebra Quaternion Mul
    return {a.w * b.w - a.i * b.i - a.j * b.j - a.k * b.k,
            a.w * b.i + a.i * b.w + a.j * b.k - a.k * b.j,
            a.w * b.j - a.i * b.k + a.j * b.w + a.k * b.i,
            a.w * b.k + a.i * b.j - a.j * b.i + a.k * b.w};
}
inline Quat qconj(const Quat& q) { return {q.w, -q.i, -q.j, -q.k}; }
// UnitQuaternion × Vector3: t = 2 (v × p); v × t; (t w + v×t) + p.
inline V3 qrot(const Quat& q, const V3& p) {
    V3 v{q.i, q.j, q.k};
    V3 c = cross(v, p);
    V3 t{c.x * 2.0, c.y * 2.0, c.z * 2.0};
    V3 c2 = cross(v, t);
    return {t.x * q.w + c2.x + p.x, t.y * q.w + c2.y + p.y, t.z * q.w + c2.z + p.z};
}
inline M3 qmatrix(const Quat& q) {  // UnitQuaternion::to_rotation_matrix
    double i = q.i, j = q.j, k = q.k, w = q.w;
    double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
    double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0;
    double ik = i * k * 2.0, jk = j * k * 2.0, wi = w * i * 2.0;
    M3 m;
    m(0, 0) = ww + ii - jj - kk; m(0, 1) = ij - wk; m(0, 2) = wj + ik;
    m(1, 0) = ij + wk; m(1, 1) = ww - ii + jj - kk; m(1, 2) = jk - wi;
    m(2, 0) = ik - wj; m(2, 1) = wi + jk; m(2, 2) = ww - ii - jj + kk;
    return m;
}
// UnitQuaternion::from_scaled_axis = Quaternion::from_imag(ω/2).exp()
// (exp_eps with eps = f64::EPSILON).
inline Quat from_scaled_axis(const V3& w) {
    V3 v{w.x / 2.0, w.y / 2.0, w.z / 2.0};
    double nn = norm_sq3(v);
    const double eps = std::numeric_limits<double>::epsilon();
    if (nn <= eps * eps) return {1.0, 0.0, 0.0, 0.0};
    double w_exp = std::exp(0.0);
    double n = std::sqrt(nn);
    double s = w_exp * std::sin(n) / n;
    return {w_exp * std::cos(n), v.x * s, v.y * s, v.z * s};
}
// UnitQuaternion::scaled_axis = axis() * angle(): the axis is the imaginary
// part, negated when w < 0, normalized (None when its norm is 0); the angle
// is 2 atan2(|v|, |w|).
inline V3 scaled_axis(const Quat& q) {
    V3 v = q.w >= 0.0 ? V3{q.i, q.j, q.k} : V3{-q.i, -q.j, -q.k};
    double n = std::sqrt(norm_sq3(v));
    if (n <= 0.0) return {0.0, 0.0, 0.0};
    V3 axis{v.x / n, v.y / n, v.z / n};
    double angle = std::atan2(std::sqrt(norm_sq3(V3{q.i, q.j, q.k})), std::fabs(q.w)) * 2.0;
    return {axis.x * angle, axis.y * angle, axis.z * angle};
}

struct SE3 {
    Quat q;
    V3 t;
};
inline V3 se3_apply(const SE3& T, const V3& p) {
    V3 r = qrot(T.q, p);
    return {r.x + T.t.x, r.y + T.t.y, r.z + T.t.z};
}
inline SE3 se3_compose(const SE3& a, const SE3& b) {  // se3.rs:32-37
    V3 r = qrot(a.q, b.t);
    return {qmul(a.q, b.q), {r.x + a.t.x, r.y + a.t.y, r.z + a.t.z}};
}
inline SE3 se3_inverse(const SE3& a) {  // se3.rs:39-43
    Quat qi = qconj(a.q);
    V3 r = qrot(qi, a.t);
    return {qi, {-r.x, -r.y, -r.z}};
}
M3 so3_left_jacobian(const V3& w) {  // se3.rs:107-120
    double theta_sq = norm_sq3(w);
    M3 S = skew(w);
    if (theta_sq < 1e-10) {
        return add(add(identity3(), scale(0.5, S)), mul(scale(1.0 / 6.0, S), S));
    }
    double theta = std::sqrt(theta_sq);
    double a = (1.0 - std::cos(theta)) / theta_sq;
    double b = (theta - std::sin(theta)) / (theta * theta_sq);
    return add(add(identity3(), scale(a, S)), mul(scale(b, S), S));
}
M3 so3_left_jacobian_inverse(const V3& w) {  // se3.rs:122-134
    double theta_sq = norm_sq3(w);
    M3 S = skew(w);
    if (theta_sq < 1e-10) {
        M3 m = identity3();
        sub_assign(m, scale(0.5, S));
        return add(m, mul(scale(1.0 / 12.0, S), S));
    }
    double theta = std::sqrt(theta_sq);
    double half = 0.5 * theta;
    double c = (1.0 - half * std::cos(half) / std::sin(half)) / theta_sq;
    M3 m = identity3();
    sub_assign(m, scale(0.5, S));
    return add(m, mul(scale(c, S), S));
}
SE3 se3_exp(const double xi[6]) {  // se3.rs:72-82
    V3 rho{xi[0], xi[1], xi[2]};
    V3 omega{xi[3], xi[4], xi[5]};
    Quat q = from_scaled_axis(omega);
    M3 V = so3_left_jacobian(omega);
    double r[3] = {rho.x, rho.y, rho.z}, t[3];
    mulv(V, r, t);
    return {q, {t[0], t[1], t[2]}};
}
void se3_log(const SE3& T, double out[6]) {  // se3.rs:57-69
    V3 omega = scaled_axis(T.q);
    M3 Vi = so3_left_jacobian_inverse(omega);
    double t[3] = {T.t.x, T.t.y, T.t.z}, rho[3];
    mulv(Vi, t, rho);
    out[0] = rho[0]; out[1] = rho[1]; out[2] = rho[2];
    out[3] = omega.x; out[4] = omega.y; out[5] = omega.z;
}

// ---------------------------------------------------------------------------
// Cameras (crates/core/src/types/camera.rs).
// ---------------------------------------------------------------------------
struct Cam {
    int model;
    int np;
    double p[12];
};
Cam to_cam(const vba_camera& c) {
    Cam k;
    k.model = c.model;
    k.np = (int)std::min<uint32_t>(c.num_params, 12);
    for (int i = 0; i < 12; ++i) k.p[i] = c.params[i];
    return k;
}
bool known_model(int m) { return m >= VBA_CAMERA_PINHOLE && m <= VBA_CAMERA_OPENCV; }
struct Intr {
    double fx, fy, cx, cy;
};
// Camera::intrinsics (camera.rs:241-271).
bool intrinsics(const Cam& c, Intr* out) {
    switch (c.model) {
        case VBA_CAMERA_PINHOLE:
        case VBA_CAMERA_OPENCV:
            if (c.np < 4) return false;
            *out = {c.p[0], c.p[1], c.p[2], c.p[3]};
            return true;
        case VBA_CAMERA_SIMPLE_PINHOLE:
        case VBA_CAMERA_SIMPLE_RADIAL:
        case VBA_CAMERA_RADIAL:
            if (c.np < 3) return false;
            *out = {c.p[0], c.p[0], c.p[1], c.p[2]};
            return true;
        default:
            return false;
    }
}
// Camera::radial_distortion (camera.rs:213-227).
bool radial_distortion(const Cam& c, double* k1, double* k2) {
    switch (c.model) {
        case VBA_CAMERA_PINHOLE:
        case VBA_CAMERA_OPENCV:
            if (c.np <= 4) return false;
            *k1 = c.p[4];
            *k2 = c.np > 5 ? c.p[5] : 0.0;
            return true;
        case VBA_CAMERA_SIMPLE_RADIAL:
            if (c.np <= 3) return false;
            *k1 = c.p[3];
            *k2 = 0.0;
            return true;
        case VBA_CAMERA_RADIAL:
            if (c.np <= 3) return false;
            *k1 = c.p[3];
            *k2 = c.np > 4 ? c.p[4] : 0.0;
            return true;
        default:
            return false;
    }
}
// Camera::tangential_distortion (camera.rs:235-242).
bool tangential_distortion(const Cam& c) {
    if (c.model != VBA_CAMERA_OPENCV) return false;
    double p1 = c.np > 6 ? c.p[6] : 0.0;
    double p2 = c.np > 7 ? c.p[7] : 0.0;
    return p1 != 0.0 || p2 != 0.0;
}
void full_opencv_coeffs(const Cam& c, double k[8]) {  // camera.rs:844-850
    for (int i = 0; i < 8; ++i) k[i] = (4 + i < c.np) ? c.p[4 + i] : 0.0;
}
// project_full_opencv (camera.rs:852-877).
bool project_full_opencv(double fx, double fy, double cx, double cy, const double k[8], const V3& P,
                         double uv[2]) {
    if (P.z <= kEps) return false;
    double x = P.x / P.z, y = P.y / P.z;
    double r2 = x * x + y * y;
    double r2_2 = r2 * r2;
    double r2_3 = r2_2 * r2;
    double num = 1.0 + k[0] * r2 + k[1] * r2_2 + k[4] * r2_3;
    double den = 1.0 + k[5] * r2 + k[6] * r2_2 + k[7] * r2_3;
    if (std::fabs(den) <= kEps) return false;
    double radial = num / den;
    double xd = x * radial + 2.0 * k[2] * x * y + k[3] * (r2 + 2.0 * x * x);
    double yd = y * radial + k[2] * (r2 + 2.0 * y * y) + 2.0 * k[3] * x * y;
    uv[0] = fx * xd + cx;
    uv[1] = fy * yd + cy;
    return std::isfinite(uv[0]) && std::isfinite(uv[1]);
}
// full_opencv_point_jacobian + chain_normalized_jacobian (camera.rs:882-941).
bool full_opencv_point_jacobian(double fx, double fy, const double k[8], const V3& P, M23* J) {
    if (P.z <= 0.0) return false;
    double x = P.x / P.z, y = P.y / P.z;
    double r2 = x * x + y * y;
    double r4 = r2 * r2;
    double num = 1.0 + k[0] * r2 + k[1] * r4 + k[4] * r4 * r2;
    double den = 1.0 + k[5] * r2 + k[6] * r4 + k[7] * r4 * r2;
    if (std::fabs(den) <= kEps) return false;
    double radial = num / den;
    double dn = k[0] + 2.0 * k[1] * r2 + 3.0 * k[4] * r4;
    double dd = k[5] + 2.0 * k[6] * r2 + 3.0 * k[7] * r4;
    double g = (dn * den - num * dd) / (den * den);
    double p1 = k[2], p2 = k[3];
    double crs = 2.0 * x * y * g + 2.0 * p1 * x + 2.0 * p2 * y;
    double a = radial + 2.0 * x * x * g + 2.0 * p1 * y + 6.0 * p2 * x;
    double b = crs, c = crs;
    double d = radial + 2.0 * y * y * g + 6.0 * p1 * y + 2.0 * p2 * x;
    double z_inv = 1.0 / P.z;
    (*J)(0, 0) = fx * a * z_inv;
    (*J)(0, 1) = fx * b * z_inv;
    (*J)(0, 2) = -fx * (a * x + b * y) * z_inv;
    (*J)(1, 0) = fy * c * z_inv;
    (*J)(1, 1) = fy * d * z_inv;
    (*J)(1, 2) = -fy * (c * x + d * y) * z_inv;
    return true;
}
// Camera::project (camera.rs:351-433), pinhole-family branches.
bool camera_project(const Cam& c, const V3& P, double uv[2]) {
    if (!(std::isfinite(P.x) && std::isfinite(P.y) && std::isfinite(P.z))) return false;
    if (!known_model(c.model)) return false;
    if (c.model == VBA_CAMERA_OPENCV && tangential_distortion(c)) {
        if (c.np < 4) return false;
        double k[8];
        full_opencv_coeffs(c, k);
        return project_full_opencv(c.p[0], c.p[1], c.p[2], c.p[3], k, P, uv);
    }
    if (P.z <= 0.0) return false;
    Intr in;
    if (!intrinsics(c, &in)) return false;
    double x = P.x / P.z;
    double y = P.y / P.z;
    double k1, k2;
    if (radial_distortion(c, &k1, &k2)) {
        if (k1 != 0.0 || k2 != 0.0) {
            double r2 = x * x + y * y;
            double d = 1.0 + k1 * r2 + k2 * r2 * r2;
            x *= d;
            y *= d;
        }
    }
    uv[0] = in.fx * x + in.cx;
    uv[1] = in.fy * y + in.cy;
    return true;
}

// MonoProjection (jacobians.rs:18-58).
enum ProjKind { PROJ_PINHOLE, PROJ_RADIAL, PROJ_LENS };
struct Proj {
    ProjKind kind;
    double k1, k2;
};
Proj mono_projection_for(const Cam& c) {
    bool radial_family = (c.model == VBA_CAMERA_PINHOLE || c.model == VBA_CAMERA_SIMPLE_PINHOLE ||
                          c.model == VBA_CAMERA_SIMPLE_RADIAL || c.model == VBA_CAMERA_RADIAL) ||
                         (c.model == VBA_CAMERA_OPENCV && !tangential_distortion(c));
    if (!radial_family) return {PROJ_LENS, 0.0, 0.0};
    double k1, k2;
    if (radial_distortion(c, &k1, &k2) && (k1 != 0.0 || k2 != 0.0)) return {PROJ_RADIAL, k1, k2};
    return {PROJ_PINHOLE, 0.0, 0.0};
}
// project_pinhole (jacobians.rs:115-124).
bool project_pinhole(const Intr& in, const V3& P, double uv[2]) {
    if (P.z <= 0.0) return false;
    uv[0] = in.fx * P.x / P.z + in.cx;
    uv[1] = in.fy * P.y / P.z + in.cy;
    return true;
}
// Camera::project_with_point_jacobian for the Brown–Conrady lens
// (camera.rs:449-507; only the OPENCV-with-tangential branch reaches it).
bool lens_project_with_jacobian(const Cam& c, const V3& P, double uv[2], M23* J) {
    if (!camera_project(c, P, uv)) return false;
    Intr in;
    if (!intrinsics(c, &in)) return false;
    double k[8];
    if (c.model == VBA_CAMERA_OPENCV && tangential_distortion(c)) {
        full_opencv_coeffs(c, k);
    } else {
        double k1 = 0.0, k2 = 0.0;
        if (!(radial_distortion(c, &k1, &k2) && (k1 != 0.0 || k2 != 0.0))) k1 = k2 = 0.0;
        for (int i = 0; i < 8; ++i) k[i] = 0.0;
        k[0] = k1;
        k[1] = k2;
    }
    if (!full_opencv_point_jacobian(in.fx, in.fy, k, P, J)) return false;
    return all_finite(*J);
}
// mono_project_with_jacobian (jacobians.rs:63-113).
bool mono_project_with_jacobian(const Proj& pr, const Cam& c, const Intr& in, const V3& P,
                                double uv[2], M23* J) {
    if (P.z <= 0.0) return false;
    const double fx = in.fx, fy = in.fy, cx = in.cx, cy = in.cy;
    switch (pr.kind) {
        case PROJ_PINHOLE: {
            if (!project_pinhole(in, P, uv)) return false;
            double z_inv = 1.0 / P.z;
            (*J)(0, 0) = fx * z_inv;
            (*J)(0, 1) = 0.0;
            (*J)(0, 2) = -fx * P.x * z_inv * z_inv;
            (*J)(1, 0) = 0.0;
            (*J)(1, 1) = fy * z_inv;
            (*J)(1, 2) = -fy * P.y * z_inv * z_inv;
            return true;
        }
        case PROJ_RADIAL: {
            double k1 = pr.k1, k2 = pr.k2;
            double x = P.x / P.z;
            double y = P.y / P.z;
            double r2 = x * x + y * y;
            double d = 1.0 + k1 * r2 + k2 * r2 * r2;
            uv[0] = fx * (x * d) + cx;
            uv[1] = fy * (y * d) + cy;
            double g = k1 + 2.0 * k2 * r2;
            double d11 = d + 2.0 * x * x * g;
            double d12 = 2.0 * x * y * g;
            double d22 = d + 2.0 * y * y * g;
            double z_inv = 1.0 / P.z;
            (*J)(0, 0) = fx * d11 * z_inv;
            (*J)(0, 1) = fx * d12 * z_inv;
            (*J)(0, 2) = -fx * (d11 * x + d12 * y) * z_inv;
            (*J)(1, 0) = fy * d12 * z_inv;
            (*J)(1, 1) = fy * d22 * z_inv;
            (*J)(1, 2) = -fy * (d12 * x + d22 * y) * z_inv;
            return std::isfinite(uv[0]) && std::isfinite(uv[1]) && all_finite(*J);
        }
        default:
            return lens_project_with_jacobian(c, P, uv, J);
    }
}
// pinhole_projection_jacobian (sqrt_qr.rs:445-463).
bool pinhole_projection_jacobian(const Intr& in, const V3& P, M23* J) {
    if (P.z <= 0.0) return false;
    double z_inv = 1.0 / P.z;
    double z_inv2 = z_inv * z_inv;
    (*J)(0, 0) = in.fx * z_inv;
    (*J)(0, 1) = 0.0;
    (*J)(0, 2) = -in.fx * P.x * z_inv2;
    (*J)(1, 0) = 0.0;
    (*J)(1, 1) = in.fy * z_inv;
    (*J)(1, 2) = -in.fy * P.y * z_inv2;
    return true;
}

// RobustKernel::{cost, weight} (pose_graph.rs:102-139).
struct Kernel {
    uint32_t kind;
    double scale;
    double cost(double s) const {
        if (kind == VBA_KERNEL_HUBER) {
            double d = scale, dsq = d * d;
            return s <= dsq ? s : 2.0 * d * std::sqrt(s) - dsq;
        }
        if (kind == VBA_KERNEL_CAUCHY) {
            double csq = scale * scale;
            return csq * std::log(1.0 + s / csq);
        }
        return s;
    }
    double weight(double s) const {
        if (kind == VBA_KERNEL_HUBER) {
            double d = scale, dsq = d * d;
            return s <= dsq ? 1.0 : d / std::sqrt(s);
        }
        if (kind == VBA_KERNEL_CAUCHY) {
            double csq = scale * scale;
            return 1.0 / (1.0 + s / csq);
        }
        return 1.0;
    }
};

// ---------------------------------------------------------------------------
// Problem (BundleAdjustment restricted to the terms of this path).
// ---------------------------------------------------------------------------
struct Sensor {
    Cam cam;
    SE3 sensor_from_rig;
};
enum ObsKind { OBS_MONO = 0, OBS_RIG = 1, OBS_RECTIFIED = 2, OBS_GENERAL = 3 };
struct Obs {
    uint32_t pose;  // index into poses (id order)
    uint32_t lm;    // index into landmarks (id order)
    double x, y;     // pixel (the left one of a stereo observation)
    int32_t sensor;  // rig: index into sensors; general stereo: index into pairs; else -1
    double weight = 1.0;  // external observation weight (gnc_weights[obs_idx]), 1 when none
    int kind = OBS_MONO;
    double x2 = 0.0, y2 = 0.0;  // rectified: u_right; general: the right pixel
};
struct Problem {
    Cam camera;
    std::vector<SE3> poses;
    std::vector<uint8_t> pose_fixed;
    std::vector<V3> lms;
    std::vector<uint8_t> lm_fixed;
    std::vector<Obs> obs;  // flattened: mono, rectified stereo, general stereo, rig; input order
    std::vector<Sensor> sensors;
    std::vector<Sensor> pairs;  // general stereo: right_camera and T_right<-left
    double baseline = 0.0;      // stereo_baseline
    size_t num_mono = 0;
    // input position ↔ id-sorted slot
    std::vector<uint32_t> pose_slot_of_input, lm_slot_of_input;
    std::vector<int32_t> pose_var, lm_var;  // variable index or -1
    uint32_t P = 0, L = 0;                   // variable counts
};

int32_t sorted_slots(uint64_t n, const uint64_t* ids, std::vector<uint32_t>* slot_of_input,
                     std::vector<uint64_t>* sorted_ids) {
    std::vector<uint32_t> order(n);
    for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ids[a] < ids[b]; });
    slot_of_input->assign(n, 0);
    sorted_ids->resize(n);
    for (uint64_t s = 0; s < n; ++s) {
        (*slot_of_input)[order[s]] = (uint32_t)s;
        (*sorted_ids)[s] = ids[order[s]];
        if (s > 0 && (*sorted_ids)[s] == (*sorted_ids)[s - 1]) return VBA_ERR_INVALID_ARGUMENT;
    }
    return VBA_OK;
}
bool find_slot(const std::vector<uint64_t>& sorted, uint64_t id, uint32_t* slot) {
    auto it = std::lower_bound(sorted.begin(), sorted.end(), id);
    if (it == sorted.end() || *it != id) return false;
    *slot = (uint32_t)(it - sorted.begin());
    return true;
}
SE3 pose_from7(const double* p) { return {{p[0], p[1], p[2], p[3]}, {p[4], p[5], p[6]}}; }
void pose_to7(const SE3& T, double* p) {
    p[0] = T.q.w; p[1] = T.q.i; p[2] = T.q.j; p[3] = T.q.k;
    p[4] = T.t.x; p[5] = T.t.y; p[6] = T.t.z;
}

// Validation in the order of optimize_weighted_backend (lm.rs:168-229, 234-252).
int32_t load_problem(const vba_problem* pb, Problem* out, uint64_t* error_id) {
    *error_id = 0;
    if (!pb) return VBA_ERR_INVALID_ARGUMENT;
    Problem& P = *out;
    P.camera = to_cam(pb->camera);
    Intr in;
    if (!known_model(P.camera.model) || !intrinsics(P.camera, &in)) return VBA_ERR_UNSUPPORTED_CAMERA_MODEL;
    const vba_stereo_terms* st = pb->stereo;
    const uint64_t ns = st ? st->num_stereo : 0, ng = st ? st->num_general : 0;
    // check_stereo_camera_model (problem.rs:733-746)
    if ((ns || ng) && P.camera.model != VBA_CAMERA_PINHOLE && P.camera.model != VBA_CAMERA_SIMPLE_PINHOLE)
        return VBA_ERR_UNSUPPORTED_CAMERA_MODEL;
    if (pb->num_poses == 0) return VBA_ERR_NO_POSES;
    bool has_visual = pb->num_observations > 0 || ns > 0 || ng > 0 || pb->num_rig_observations > 0;
    if (!has_visual) return VBA_ERR_NO_OBSERVATIONS;
    if (pb->num_landmarks == 0) return VBA_ERR_NO_LANDMARKS;
    std::vector<uint64_t> pose_ids, lm_ids;
    if (sorted_slots(pb->num_poses, pb->pose_ids, &P.pose_slot_of_input, &pose_ids) != VBA_OK)
        return VBA_ERR_INVALID_ARGUMENT;
    if (sorted_slots(pb->num_landmarks, pb->landmark_ids, &P.lm_slot_of_input, &lm_ids) != VBA_OK)
        return VBA_ERR_INVALID_ARGUMENT;
    P.poses.resize(pb->num_poses);
    P.pose_fixed.assign(pb->num_poses, 0);
    for (uint64_t i = 0; i < pb->num_poses; ++i) {
        uint32_t s = P.pose_slot_of_input[i];
        P.poses[s] = pose_from7(pb->poses + 7 * i);
        P.pose_fixed[s] = pb->pose_fixed ? (pb->pose_fixed[i] != 0) : 0;
    }
    P.lms.resize(pb->num_landmarks);
    P.lm_fixed.assign(pb->num_landmarks, 0);
    for (uint64_t i = 0; i < pb->num_landmarks; ++i) {
        uint32_t s = P.lm_slot_of_input[i];
        const double* x = pb->landmarks + 3 * i;
        P.lms[s] = {x[0], x[1], x[2]};
        P.lm_fixed[s] = pb->landmark_fixed ? (pb->landmark_fixed[i] != 0) : 0;
    }
    P.obs.clear();
    P.obs.reserve(pb->num_observations + ns + ng + pb->num_rig_observations);
    for (uint64_t i = 0; i < pb->num_observations; ++i) {
        Obs o;
        if (!find_slot(pose_ids, pb->obs_pose_ids[i], &o.pose)) {
            *error_id = pb->obs_pose_ids[i];
            return VBA_ERR_MISSING_POSE;
        }
        if (!find_slot(lm_ids, pb->obs_landmark_ids[i], &o.lm)) {
            *error_id = pb->obs_landmark_ids[i];
            return VBA_ERR_MISSING_LANDMARK;
        }
        o.x = pb->obs_xy[2 * i];
        o.y = pb->obs_xy[2 * i + 1];
        o.sensor = -1;
        P.obs.push_back(o);
    }
    P.num_mono = P.obs.size();
    auto resolve = [&](uint64_t pid, uint64_t lid, Obs* o) -> int32_t {
        if (!find_slot(pose_ids, pid, &o->pose)) {
            *error_id = pid;
            return VBA_ERR_MISSING_POSE;
        }
        if (!find_slot(lm_ids, lid, &o->lm)) {
            *error_id = lid;
            return VBA_ERR_MISSING_LANDMARK;
        }
        return VBA_OK;
    };
    // rectified stereo: the baseline, then the ids (lm.rs:194-206)
    if (ns) {
        if (!(std::isfinite(st->baseline) && st->baseline > 0.0)) return VBA_ERR_MISSING_STEREO_BASELINE;
        P.baseline = st->baseline;
        for (uint64_t i = 0; i < ns; ++i) {
            Obs o;
            if (int32_t e = resolve(st->stereo_pose_ids[i], st->stereo_landmark_ids[i], &o)) return e;
            o.kind = OBS_RECTIFIED;
            o.x = st->stereo_obs[3 * i];
            o.y = st->stereo_obs[3 * i + 1];
            o.x2 = st->stereo_obs[3 * i + 2];
            o.sensor = -1;
            P.obs.push_back(o);
        }
    }
    // general stereo: ids, then the right camera's intrinsics (lm.rs:207-217)
    if (ng) {
        P.pairs.resize(st->num_pairs);
        for (uint64_t q = 0; q < st->num_pairs; ++q) {
            P.pairs[q].cam = to_cam(st->pairs[q].camera);
            P.pairs[q].sensor_from_rig = pose_from7(st->pairs[q].sensor_from_rig);
        }
        for (uint64_t i = 0; i < ng; ++i) {
            Obs o;
            if (int32_t e = resolve(st->general_pose_ids[i], st->general_landmark_ids[i], &o)) return e;
            const uint32_t q = st->general_pair[i];
            if (q >= st->num_pairs) return VBA_ERR_INVALID_ARGUMENT;
            Intr ri;
            if (!known_model(P.pairs[q].cam.model) || !intrinsics(P.pairs[q].cam, &ri))
                return VBA_ERR_UNSUPPORTED_CAMERA_MODEL;
            o.kind = OBS_GENERAL;
            o.x = st->general_obs[4 * i];
            o.y = st->general_obs[4 * i + 1];
            o.x2 = st->general_obs[4 * i + 2];
            o.y2 = st->general_obs[4 * i + 3];
            o.sensor = (int32_t)q;
            P.obs.push_back(o);
        }
    }
    P.sensors.resize(pb->num_sensors);
    for (uint64_t s = 0; s < pb->num_sensors; ++s) {
        P.sensors[s].cam = to_cam(pb->sensors[s].camera);
        P.sensors[s].sensor_from_rig = pose_from7(pb->sensors[s].sensor_from_rig);
    }
    for (uint64_t i = 0; i < pb->num_rig_observations; ++i) {
        Obs o;
        if (!find_slot(pose_ids, pb->rig_pose_ids[i], &o.pose)) {
            *error_id = pb->rig_pose_ids[i];
            return VBA_ERR_MISSING_POSE;
        }
        if (!find_slot(lm_ids, pb->rig_landmark_ids[i], &o.lm)) {
            *error_id = pb->rig_landmark_ids[i];
            return VBA_ERR_MISSING_LANDMARK;
        }
        uint32_t s = pb->rig_sensor[i];
        if (s >= pb->num_sensors) return VBA_ERR_INVALID_ARGUMENT;
        Intr si;
        if (!known_model(P.sensors[s].cam.model) || !intrinsics(P.sensors[s].cam, &si))
            return VBA_ERR_UNSUPPORTED_CAMERA_MODEL;
        o.x = pb->rig_xy[2 * i];
        o.y = pb->rig_xy[2 * i + 1];
        o.sensor = (int32_t)s;
        o.kind = OBS_RIG;
        P.obs.push_back(o);
    }
    P.pose_var.assign(P.poses.size(), -1);
    P.lm_var.assign(P.lms.size(), -1);
    P.P = P.L = 0;
    for (size_t s = 0; s < P.poses.size(); ++s)
        if (!P.pose_fixed[s]) P.pose_var[s] = (int32_t)P.P++;
    for (size_t s = 0; s < P.lms.size(); ++s)
        if (!P.lm_fixed[s]) P.lm_var[s] = (int32_t)P.L++;
    if (P.P == 0 && P.L == 0) return VBA_ERR_ALL_POSES_FIXED;
    return VBA_OK;
}

// ---------------------------------------------------------------------------
// Per-observation residual and Jacobians.
// ---------------------------------------------------------------------------
struct Lin {
    double r[2];
    M26 jp;
    M23 jl;
};
// Monocular: normal_equations.rs:418-458.
bool linearize_mono(const Problem& P, const Intr& in, const Proj& pr, const Obs& o, Lin* out) {
    const SE3& T = P.poses[o.pose];
    const V3& X = P.lms[o.lm];
    M3 R = qmatrix(T.q);
    V3 xc = se3_apply(T, X);
    if (xc.z <= 0.0) return false;
    double uv[2];
    M23 jpi;
    if (!mono_project_with_jacobian(pr, P.camera, in, xc, uv, &jpi)) return false;
    out->r[0] = uv[0] - o.x;
    out->r[1] = uv[1] - o.y;
    M36 dx = M36::zeros();
    M3 nRS = mul(neg(R), skew(X));
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            dx(r, c) = R(r, c);
            dx(r, 3 + c) = nRS(r, c);
        }
    out->jp = mul(jpi, dx);
    out->jl = mul(jpi, R);
    return true;
}
// Rig: rig_residual_jacobians (sqrt_qr.rs:960-1001).
bool linearize_rig(const Problem& P, const Obs& o, Lin* out) {
    const Sensor& S = P.sensors[o.sensor];
    Intr in;
    if (!intrinsics(S.cam, &in)) return false;
    const SE3& T = P.poses[o.pose];
    const V3& X = P.lms[o.lm];
    V3 prig = se3_apply(T, X);
    V3 ps = se3_apply(S.sensor_from_rig, prig);
    Proj pr = mono_projection_for(S.cam);
    double uv[2];
    M23 proj;
    if (pr.kind == PROJ_PINHOLE) {
        if (!project_pinhole(in, ps, uv)) return false;
        if (!pinhole_projection_jacobian(in, ps, &proj)) return false;
    } else {
        if (!mono_project_with_jacobian(pr, S.cam, in, ps, uv, &proj)) return false;
    }
    out->r[0] = uv[0] - o.x;
    out->r[1] = uv[1] - o.y;
    M3 Rwr = qmatrix(T.q);
    M3 Rrs = qmatrix(S.sensor_from_rig.q);
    M36 d = M36::zeros();
    M3 nRS = mul(neg(Rwr), skew(X));
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            d(r, c) = Rwr(r, c);
            d(r, 3 + c) = nRS(r, c);
        }
    M23 pr_rs = mul(proj, Rrs);
    out->jp = mul(pr_rs, d);
    out->jl = mul(pr_rs, Rwr);
    return true;
}
bool linearize_obs(const Problem& P, const Intr& in, const Proj& pr, const Obs& o, Lin* out) {
    return o.kind == OBS_MONO ? linearize_mono(P, in, pr, o, out) : linearize_rig(P, o, out);
}
// d(T X)/d(translation, rotation) = [R | -R [X]x] (normal_equations.rs:533-539).
M36 pose_point_jacobian(const M3& R, const V3& X) {
    M36 d = M36::zeros();
    M3 nRS = mul(neg(R), skew(X));
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            d(r, c) = R(r, c);
            d(r, 3 + c) = nRS(r, c);
        }
    return d;
}
// Rectified stereo (normal_equations.rs:486-571): residual (u_l, v_l, u_r)
// with u_r = u_l - fx b / Z through the plain pinhole.
struct Lin3 {
    double r[3];
    M36 jp;
    M3 jl;
};
bool linearize_rectified(const Problem& P, const Intr& in, const Obs& o, Lin3* out) {
    const SE3& T = P.poses[o.pose];
    const V3& X = P.lms[o.lm];
    const M3 R = qmatrix(T.q);
    const V3 xc = se3_apply(T, X);
    if (xc.z <= 0.0) return false;
    double uv[2];
    if (!project_pinhole(in, xc, uv)) return false;
    const double b = P.baseline;
    const double u_r = uv[0] - in.fx * b / xc.z;
    out->r[0] = uv[0] - o.x;
    out->r[1] = uv[1] - o.y;
    out->r[2] = u_r - o.x2;
    const double z_inv = 1.0 / xc.z, z_inv2 = z_inv * z_inv;
    M3 jpi = M3::zeros();
    jpi(0, 0) = in.fx * z_inv;
    jpi(0, 2) = -in.fx * xc.x * z_inv2;
    jpi(1, 1) = in.fy * z_inv;
    jpi(1, 2) = -in.fy * xc.y * z_inv2;
    jpi(2, 0) = in.fx * z_inv;
    jpi(2, 2) = -in.fx * (xc.x - b) * z_inv2;
    out->jp = mul(jpi, pose_point_jacobian(R, X));
    out->jl = mul(jpi, R);
    return true;
}
// General stereo: general_stereo_residual_jacobians (sqrt_qr.rs:465-515),
// residual (u_l, v_l, u_r, v_r), the right camera at T_right<-left.
struct Lin4 {
    double r[4];
    M46 jp;
    M43 jl;
};
bool linearize_general(const Problem& P, const Intr& in, const Obs& o, Lin4* out) {
    const Sensor& S = P.pairs[o.sensor];
    Intr ri;
    if (!intrinsics(S.cam, &ri)) return false;
    const SE3& T = P.poses[o.pose];
    const V3& X = P.lms[o.lm];
    const V3 pl = se3_apply(T, X);
    const V3 pr = se3_apply(S.sensor_from_rig, pl);
    double ul[2], ur[2];
    if (!project_pinhole(in, pl, ul) || !project_pinhole(ri, pr, ur)) return false;
    out->r[0] = ul[0] - o.x;
    out->r[1] = ul[1] - o.y;
    out->r[2] = ur[0] - o.x2;
    out->r[3] = ur[1] - o.y2;
    M23 jl, jr;
    if (!pinhole_projection_jacobian(in, pl, &jl) || !pinhole_projection_jacobian(ri, pr, &jr)) return false;
    const M3 Rwl = qmatrix(T.q), Rlr = qmatrix(S.sensor_from_rig.q);
    const M36 d = pose_point_jacobian(Rwl, X);
    const M23 jr_lr = mul(jr, Rlr);
    const M26 jpl = mul(jl, d), jpr = mul(jr_lr, d);
    const M23 jll = mul(jl, Rwl), jlr = mul(jr_lr, Rwl);
    for (int c = 0; c < 6; ++c)
        for (int k = 0; k < 2; ++k) {
            out->jp(k, c) = jpl(k, c);
            out->jp(2 + k, c) = jpr(k, c);
        }
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 2; ++k) {
            out->jl(k, c) = jll(k, c);
            out->jl(2 + k, c) = jlr(k, c);
        }
    return true;
}
// Squared norm of a residual, summed in row order.
template <int N>
double norm_sq(const double* r) {
    double s = 0.0;
    for (int k = 0; k < N; ++k) s += r[k] * r[k];
    return s;
}
// The squared residual of an observation of any kind and whether it can be
// evaluated: the projection tests of robust_cost_weighted /
// nonprojectable_observation_count (problem.rs:252-303, 497-583).
bool squared_residual(const Problem& P, const Intr& in, const Obs& o, double* s) {
    if (o.kind == OBS_MONO) {
        const V3 xc = se3_apply(P.poses[o.pose], P.lms[o.lm]);
        double uv[2];
        if (xc.z <= 0.0 || !camera_project(P.camera, xc, uv)) return false;
        const double rx = uv[0] - o.x, ry = uv[1] - o.y;
        *s = rx * rx + ry * ry;
        return true;
    }
    if (o.kind == OBS_RECTIFIED) {
        const V3 xc = se3_apply(P.poses[o.pose], P.lms[o.lm]);
        double uv[2];
        if (xc.z <= 0.0 || !project_pinhole(in, xc, uv)) return false;
        const double u_r = uv[0] - in.fx * P.baseline / xc.z;
        const double dx = uv[0] - o.x, dy = uv[1] - o.y, dr = u_r - o.x2;
        *s = dx * dx + dy * dy + dr * dr;
        return true;
    }
    if (o.kind == OBS_GENERAL) {
        Lin4 lin;
        if (!linearize_general(P, in, o, &lin)) return false;
        *s = norm_sq<4>(lin.r);
        return true;
    }
    Lin lin;
    if (!linearize_rig(P, o, &lin)) return false;
    *s = lin.r[0] * lin.r[0] + lin.r[1] * lin.r[1];
    return true;
}

// robust_cost_weighted (problem.rs:497-583) and
// nonprojectable_observation_count (problem.rs:252-303).
// threads > 1 evaluates the per-observation terms in parallel and sums them
// in observation order, as the serial loop does.
void cost_and_nonprojectable(const Problem& P, const Kernel& K, double* cost, uint64_t* nonproj,
                             int threads = 1) {
    double total = 0.0;
    uint64_t bad = 0;
    Intr in;
    intrinsics(P.camera, &in);
    const size_t n = P.obs.size();
    if (threads <= 1 || n < 65536) {
        for (const Obs& o : P.obs) {
            double s;
            if (!squared_residual(P, in, o, &s)) {
                ++bad;
                continue;
            }
            total += o.weight * K.cost(s);
        }
    } else {
        std::vector<double> term(n);
        std::vector<uint8_t> ok(n);
        parallel_ranges(threads, n, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) {
                double s;
                ok[i] = squared_residual(P, in, P.obs[i], &s);
                if (ok[i]) term[i] = P.obs[i].weight * K.cost(s);
            }
        });
        for (size_t i = 0; i < n; ++i) {
            if (!ok[i]) {
                ++bad;
                continue;
            }
            total += term[i];
        }
    }
    *cost = total;
    *nonproj = bad;
}

// ---------------------------------------------------------------------------
// Normal equations (NormalEquationsBa, normal_equations.rs:10-18, 78-86).
// ---------------------------------------------------------------------------
struct CrossEntry {
    uint32_t pose;
    M63 w;
};
struct LandmarkBlock {
    M3 h_ll;
    double b_l[3];
    std::vector<CrossEntry> cross;
};
struct Normal {
    uint32_t P = 0;
    bool dense = true;
    std::vector<double> H;        // dense (6P)^2 row-major, when dense
    std::vector<M66> diag;        // pose blocks, when !dense
    std::vector<double> b_p;      // 6P
    std::vector<LandmarkBlock> lm;
    double& h(uint32_t r, uint32_t c) {
        return dense ? H[(size_t)r * 6 * P + c] : diag[r / 6](r % 6, c % 6);
    }
};

struct Contribution {
    bool valid = false;
    int32_t p = -1, l = -1;
    M66 hpp;
    double bp[6];
    M3 hll;
    double bl[3];
    M63 cross;
};
// compute_mono_contribution (normal_equations.rs:203-265) and its rig twin.
// The blocks of one observation's N residual rows: w J_p^T J_p, w J_p^T r,
// w J_l^T J_l, w J_l^T r and w J_p^T J_l (normal_equations.rs:546-567).
template <int N>
void contribution_from(const Problem& P, const Kernel& K, const Obs& o, const double* r,
                       const Mat<N, 6>& jp, const Mat<N, 3>& jl, Contribution* c) {
    double s = norm_sq<N>(r);
    double w = K.weight(s) * o.weight;
    c->p = P.pose_var[o.pose];
    c->l = P.lm_var[o.lm];
    if (c->p >= 0) {
        M66 h = mul(tr(jp), jp);
        double b[6];
        Mat<6, N> jpt = tr(jp);
        mulv(jpt, r, b);
        c->hpp = scale(w, h);
        for (int k = 0; k < 6; ++k) c->bp[k] = w * b[k];
    }
    if (c->l >= 0) {
        M3 h = mul(tr(jl), jl);
        double b[3];
        Mat<3, N> jlt = tr(jl);
        mulv(jlt, r, b);
        c->hll = scale(w, h);
        for (int k = 0; k < 3; ++k) c->bl[k] = w * b[k];
    }
    if (c->p >= 0 && c->l >= 0) c->cross = scale(w, mul(tr(jp), jl));
}
void compute_contribution(const Problem& P, const Intr& in, const Proj& pr, const Kernel& K,
                          const Obs& o, Contribution* c) {
    if (o.kind == OBS_RECTIFIED) {
        Lin3 lin;
        c->valid = linearize_rectified(P, in, o, &lin);
        if (c->valid) contribution_from<3>(P, K, o, lin.r, lin.jp, lin.jl, c);
        return;
    }
    if (o.kind == OBS_GENERAL) {
        Lin4 lin;
        c->valid = linearize_general(P, in, o, &lin);
        if (c->valid) contribution_from<4>(P, K, o, lin.r, lin.jp, lin.jl, c);
        return;
    }
    Lin lin;
    c->valid = linearize_obs(P, in, pr, o, &lin);
    if (c->valid) contribution_from<2>(P, K, o, lin.r, lin.jp, lin.jl, c);
}
void apply_contribution(const Contribution& c, Normal* ne) {
    if (!c.valid) return;
    if (c.p >= 0) {
        uint32_t p = (uint32_t)c.p;
        for (int r = 0; r < 6; ++r) {
            for (int cc = 0; cc < 6; ++cc) {
                double& dst = ne->h(p * 6 + r, p * 6 + cc);
                dst = dst + c.hpp(r, cc);
            }
            ne->b_p[p * 6 + r] = ne->b_p[p * 6 + r] + c.bp[r];
        }
    }
    if (c.l >= 0) {
        LandmarkBlock& L = ne->lm[c.l];
        add_assign(L.h_ll, c.hll);
        for (int k = 0; k < 3; ++k) L.b_l[k] = L.b_l[k] + c.bl[k];
    }
    if (c.p >= 0 && c.l >= 0) ne->lm[c.l].cross.push_back({(uint32_t)c.p, c.cross});
}
// build_normal_equations (normal_equations.rs:352-484, 617-651). threads > 1
// computes contributions chunk-wise in parallel and scatters serially in
// observation order (assemble_mono_observations_parallel, :267-350).
void build_normal_equations(const Problem& P, const Kernel& K, bool pose_blocks, int threads,
                            Normal* ne) {
    ne->P = P.P;
    ne->dense = !pose_blocks;
    if (ne->dense) {
        ne->H.assign((size_t)36 * P.P * P.P, 0.0);
        ne->diag.clear();
    } else {
        ne->diag.assign(P.P, M66::zeros());
        ne->H.clear();
    }
    ne->b_p.assign((size_t)6 * P.P, 0.0);
    ne->lm.assign(P.L, LandmarkBlock{M3::zeros(), {0.0, 0.0, 0.0}, {}});
    Intr in;
    intrinsics(P.camera, &in);
    Proj pr = mono_projection_for(P.camera);
    const size_t n = P.obs.size();
    if (threads <= 1 || n < 4096) {
        Contribution c;
        for (size_t i = 0; i < n; ++i) {
            compute_contribution(P, in, pr, K, P.obs[i], &c);
            apply_contribution(c, ne);
        }
        return;
    }
    const size_t chunk = 65536;
    std::vector<Contribution> buf(chunk);
    for (size_t start = 0; start < n; start += chunk) {
        size_t end = std::min(n, start + chunk);
        size_t m = end - start;
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) {
            pool.emplace_back([&, t]() {
                for (size_t i = t; i < m; i += (size_t)threads)
                    compute_contribution(P, in, pr, K, P.obs[start + i], &buf[i]);
            });
        }
        for (auto& th : pool) th.join();
        for (size_t i = 0; i < m; ++i) apply_contribution(buf[i], ne);
    }
}

// nalgebra Matrix3::try_inverse (cofactor expansion; None only for det == 0).
bool try_inverse3(const M3& m, M3* out) {
    double m11 = m(0, 0), m12 = m(0, 1), m13 = m(0, 2);
    double m21 = m(1, 0), m22 = m(1, 1), m23 = m(1, 2);
    double m31 = m(2, 0), m32 = m(2, 1), m33 = m(2, 2);
    double minor_m12_m23 = m22 * m33 - m32 * m23;
    double minor_m11_m23 = m21 * m33 - m31 * m23;
    double minor_m11_m22 = m21 * m32 - m31 * m22;
    double det = m11 * minor_m12_m23 - m12 * minor_m11_m23 + m13 * minor_m11_m22;
    if (det == 0.0) return false;
    (*out)(0, 0) = minor_m12_m23 / det;
    (*out)(0, 1) = (m13 * m32 - m33 * m12) / det;
    (*out)(0, 2) = (m12 * m23 - m22 * m13) / det;
    (*out)(1, 0) = -minor_m11_m23 / det;
    (*out)(1, 1) = (m11 * m33 - m31 * m13) / det;
    (*out)(1, 2) = (m13 * m21 - m23 * m11) / det;
    (*out)(2, 0) = minor_m11_m22 / det;
    (*out)(2, 1) = (m12 * m31 - m32 * m11) / det;
    (*out)(2, 2) = (m11 * m22 - m21 * m12) / det;
    return true;
}
M3 damped3(const M3& h, double lambda) {
    M3 d = h;
    if (lambda > 0.0) {
        d(0, 0) += lambda;
        d(1, 1) += lambda;
        d(2, 2) += lambda;
    }
    return d;
}

// nalgebra Cholesky::new on a dense n x n row-major matrix (column-oriented,
// left-looking). Returns false for a non-positive or non-finite pivot.
// nalgebra's column Cholesky: column j gets -L(j, k) L(i, k) added for every
// k < j in ascending order, then is scaled by the square root of its
// diagonal. Each entry (i, j) is accumulated on its own, row by row (the same
// operations in the same order as the column sweep, read along the rows);
// threads > 1 splits the rows of a column.
// Columns are taken in blocks of kCholBlock: first every entry (i, j) of the
// block (i >= j) gets the terms k below the block, k ascending, read against
// a contiguous copy of the block's rows; then the block's columns finish one
// by one with the terms k inside the block. Each entry sees exactly the
// column sweep's sequence of operations.
constexpr size_t kCholBlock = 64;
bool cholesky_dense(std::vector<double>& m, size_t n, int threads = 1) {
    auto at = [&](size_t r, size_t c) -> double& { return m[r * n + c]; };
    std::vector<double> panel;
    for (size_t j = 0; j < n; ++j) {
        if (j % kCholBlock == 0 && j > 0) {
            const size_t j0 = j, j1 = std::min(n, j0 + kCholBlock), B = j1 - j0;
            panel.resize(j0 * B);  // panel[k * B + jj] = L(j0 + jj, k), k < j0
            for (size_t jj = 0; jj < B; ++jj)
                for (size_t k = 0; k < j0; ++k) panel[k * B + jj] = m[(j0 + jj) * n + k];
            auto rows = [&](size_t i0, size_t i1) {
                double acc[kCholBlock];
                for (size_t i = j0 + i0; i < j0 + i1; ++i) {
                    double* ri = m.data() + i * n;
                    const size_t cols = std::min(B, i - j0 + 1);  // entries j <= i
                    for (size_t jj = 0; jj < cols; ++jj) acc[jj] = ri[j0 + jj];
                    for (size_t k = 0; k < j0; ++k) {
                        const double aik = ri[k];
                        const double* pk = panel.data() + k * B;
                        for (size_t jj = 0; jj < cols; ++jj) acc[jj] = -pk[jj] * aik + acc[jj];
                    }
                    for (size_t jj = 0; jj < cols; ++jj) ri[j0 + jj] = acc[jj];
                }
            };
            if (threads > 1 && (n - j0) * j0 >= (1u << 14)) parallel_ranges(threads, n - j0, rows);
            else rows(0, n - j0);
        }
        const size_t kb = j - j % kCholBlock;  // terms k in [kb, j) are still to come
        auto rows = [&](size_t i0, size_t i1) {
            const double* rj = m.data() + j * n;
            for (size_t i = j + i0; i < j + i1; ++i) {
                const double* ri = m.data() + i * n;
                double v = ri[j];
                for (size_t k = kb; k < j; ++k) v = -rj[k] * ri[k] + v;
                m[i * n + j] = v;
            }
        };
        if (threads > 1 && (n - j) * (j - kb) >= (1u << 13)) parallel_ranges(threads, n - j, rows);
        else rows(0, n - j);
        double diag = at(j, j);
        if (diag != 0.0 && diag >= 0.0) {
            double denom = std::sqrt(diag);
            at(j, j) = denom;
            for (size_t i = j + 1; i < n; ++i) at(i, j) = at(i, j) / denom;
            continue;
        }
        return false;
    }
    return true;
}
// Cholesky::solve: forward (column axpy) then the adjoint solve (dynamic dot).
void cholesky_solve(const std::vector<double>& L, size_t n, std::vector<double>& b) {
    auto at = [&](size_t r, size_t c) { return L[r * n + c]; };
    for (size_t i = 0; i < n; ++i) {
        double coeff = b[i] / at(i, i);
        b[i] = coeff;
        for (size_t r = i + 1; r < n; ++r) b[r] = -coeff * at(r, i) + b[r];
    }
    std::vector<double> col;
    for (size_t ii = n; ii-- > 0;) {
        size_t len = n - ii - 1;
        col.resize(len);
        for (size_t r = 0; r < len; ++r) col[r] = at(ii + 1 + r, ii);
        double dot = dyn_dot(col.data(), b.data() + ii + 1, len);
        b[ii] = (b[ii] - dot) / at(ii, ii);
    }
}
// nalgebra LU::new (partial pivoting, gauss_step / gauss_step_swap) + solve.
bool lu_solve(std::vector<double> m, size_t n, std::vector<double>& b) {
    auto at = [&](size_t r, size_t c) -> double& { return m[r * n + c]; };
    std::vector<std::pair<size_t, size_t>> perm;
    for (size_t i = 0; i < n; ++i) {
        size_t piv = i;
        double best = std::fabs(at(i, i));
        for (size_t r = i + 1; r < n; ++r)
            if (std::fabs(at(r, i)) > best) {
                best = std::fabs(at(r, i));
                piv = r;
            }
        double diag = at(piv, i);
        if (diag == 0.0) continue;
        if (piv != i) {
            perm.push_back({i, piv});
            for (size_t c = 0; c < i; ++c) std::swap(at(i, c), at(piv, c));
            // gauss_step_swap
            double inv_diag = 1.0 / diag;
            std::swap(at(i, i), at(piv, i));
            for (size_t r = i + 1; r < n; ++r) at(r, i) *= inv_diag;
            for (size_t c = i + 1; c < n; ++c) {
                std::swap(at(i, c), at(piv, c));
                double f = -at(i, c);
                for (size_t r = i + 1; r < n; ++r) at(r, c) = f * at(r, i) + at(r, c);
            }
        } else {
            double inv_diag = 1.0 / diag;
            for (size_t r = i + 1; r < n; ++r) at(r, i) *= inv_diag;
            for (size_t c = i + 1; c < n; ++c) {
                double f = -at(i, c);
                for (size_t r = i + 1; r < n; ++r) at(r, c) = f * at(r, i) + at(r, c);
            }
        }
    }
    for (auto& pr : perm) std::swap(b[pr.first], b[pr.second]);
    for (size_t i = 0; i + 1 < n; ++i) {
        double coeff = b[i] / 1.0;
        for (size_t r = i + 1; r < n; ++r) b[r] = -coeff * at(r, i) + b[r];
    }
    for (size_t ii = n; ii-- > 0;) {
        double diag = at(ii, ii);
        if (diag == 0.0) return false;
        double coeff = b[ii] / diag;
        b[ii] = coeff;
        for (size_t r = 0; r < ii; ++r) b[r] = -coeff * at(r, ii) + b[r];
    }
    return true;
}
// solve_normal_equations (pose_graph.rs:2259-2270).
// allow_lu = false: LinearSolver::Sparse, whose block Cholesky has no
// fallback (normal_equations.rs:1483-1535; its factor equals the dense one
// to rounding, tests/bundle_adjustment.rs:1160-1200).
bool solve_normal_equations(const std::vector<double>& A, size_t n, const std::vector<double>& b,
                            std::vector<double>* x, int threads = 1, bool allow_lu = true) {
    std::vector<double> L = A;
    if (cholesky_dense(L, n, threads)) {
        *x = b;
        cholesky_solve(L, n, *x);
        return true;
    }
    if (!allow_lu) return false;
    *x = b;
    return lu_solve(A, n, *x);
}

// 6x6 Cholesky::new + inverse (static size; same column algorithm).
bool cholesky_inverse6(const M66& block, M66* inv) {
    std::vector<double> m(block.a, block.a + 36);
    if (!cholesky_dense(m, 6)) return false;
    for (int c = 0; c < 6; ++c) {
        std::vector<double> e(6, 0.0);
        e[c] = 1.0;
        cholesky_solve(m, 6, e);
        for (int r = 0; r < 6; ++r) (*inv)(r, c) = e[r];
    }
    return true;
}

// ---------------------------------------------------------------------------
// Dense Schur step: solve_step_with_debug (normal_equations.rs:1314-1592).
// ---------------------------------------------------------------------------
// threads > 1: the landmark inverses and the back-substitution run per
// landmark in parallel; the Schur reduction runs per range of pose rows, each
// range visiting the landmarks in order, so every entry of S and b_S takes
// the serial sequence of updates.
int32_t solve_step_dense(Normal& ne, uint32_t L, double lambda, std::vector<double>* dp,
                         std::vector<double>* dl, int threads = 1, bool allow_lu = true) {
    const uint32_t P = ne.P;
    if (P == 0) {  // landmark-only branch (:1332-1371)
        dl->assign((size_t)3 * L, 0.0);
        for (uint32_t l = 0; l < L; ++l) {
            M3 inv;
            if (!try_inverse3(damped3(ne.lm[l].h_ll, lambda), &inv)) return VBA_ERR_SINGULAR_SYSTEM;
            double v[3];
            mulv(inv, ne.lm[l].b_l, v);
            for (int k = 0; k < 3; ++k) (*dl)[l * 3 + k] = -v[k];
        }
        dp->clear();
        return VBA_OK;
    }
    const size_t dim = (size_t)6 * P;
    std::vector<double>& S = ne.H;
    if (!ne.dense) {
        S.assign(dim * dim, 0.0);
        for (uint32_t p = 0; p < P; ++p)
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) S[(p * 6 + r) * dim + p * 6 + c] = ne.diag[p](r, c);
    }
    if (lambda > 0.0)
        for (size_t k = 0; k < dim; ++k) S[k * dim + k] += lambda;
    std::vector<double> b(dim);
    for (size_t k = 0; k < dim; ++k) b[k] = -ne.b_p[k];
    std::vector<M3> inv(L);
    std::vector<uint8_t> has_inv(L, 0);
    parallel_ranges(threads, L, [&](size_t l0, size_t l1) {
        for (size_t l = l0; l < l1; ++l) has_inv[l] = try_inverse3(damped3(ne.lm[l].h_ll, lambda), &inv[l]);
    });
    parallel_ranges(threads, P, [&](size_t p0, size_t p1) {
        for (uint32_t l = 0; l < L; ++l) {
            if (!has_inv[l]) continue;
            const LandmarkBlock& lb = ne.lm[l];
            for (const CrossEntry& a : lb.cross) {
                if (a.pose < p0 || a.pose >= p1) continue;
                M63 a_h = mul(a.w, inv[l]);
                for (const CrossEntry& bq : lb.cross) {
                    M66 blk = mul(a_h, tr(bq.w));
                    for (int r = 0; r < 6; ++r)
                        for (int c = 0; c < 6; ++c) {
                            double& dst = S[(a.pose * 6 + r) * dim + bq.pose * 6 + c];
                            dst -= blk(r, c);
                        }
                }
                double upd[6];
                mulv(a_h, lb.b_l, upd);
                for (int k = 0; k < 6; ++k) b[a.pose * 6 + k] += upd[k];
            }
        }
    });
    if (!solve_normal_equations(S, dim, b, dp, threads, allow_lu)) return VBA_ERR_SINGULAR_SYSTEM;
    dl->assign((size_t)3 * L, 0.0);
    parallel_ranges(threads, L, [&](size_t l0, size_t l1) {
        for (size_t l = l0; l < l1; ++l) {
            if (!has_inv[l]) continue;
            const LandmarkBlock& lb = ne.lm[l];
            double acc[3] = {-lb.b_l[0], -lb.b_l[1], -lb.b_l[2]};
            for (const CrossEntry& a : lb.cross) {
                double sub[3];
                mulv(tr(a.w), dp->data() + a.pose * 6, sub);
                for (int k = 0; k < 3; ++k) acc[k] -= sub[k];
            }
            double d[3];
            mulv(inv[l], acc, d);
            for (int k = 0; k < 3; ++k) (*dl)[l * 3 + k] = d[k];
        }
    });
    return VBA_OK;
}

// ---------------------------------------------------------------------------
// Matrix-free implicit Schur operator + PCG (implicit_schur.rs).
// ---------------------------------------------------------------------------
enum ImplicitErr {
    IE_OK = 0,
    IE_NONFINITE = -1,
    IE_INVALID_LAMBDA = -2,
    IE_EMPTY = -3,
    IE_NON_SPD_PRECONDITIONER = -4,
    IE_NON_POSITIVE_CURVATURE = -5,
    IE_RESIDUAL_CHECK_FAILED = -6,
    IE_MAX_ITERATIONS = -7,
    IE_INVALID_TOLERANCE = -8
};
struct Implicit {
    const Normal* ne = nullptr;
    std::vector<M66> diagonal;  // damped
    std::vector<M3> inv;
    std::vector<uint8_t> has_inv;
    std::vector<M66> precond_inv;
    std::vector<double> rhs;
    // threads > 1: each pose's cross entries in (landmark, entry) order, the
    // order in which the serial operator subtracts them
    int threads = 1;
    std::vector<size_t> pose_ptr;
    std::vector<std::pair<uint32_t, const CrossEntry*>> pose_entries;
    size_t dim() const { return diagonal.size() * 6; }
};
// ImplicitSchurOperator::new (implicit_schur.rs:133-297).
int build_implicit_serial(const Normal& ne, double lambda, Implicit* op) {
    if (!std::isfinite(lambda) || lambda < 0.0) return IE_INVALID_LAMBDA;
    if (ne.diag.empty()) return IE_EMPTY;
    op->ne = &ne;
    const uint32_t P = (uint32_t)ne.diag.size();
    if (!all_finite(ne.b_p)) return IE_NONFINITE;
    for (const M66& b : ne.diag)
        if (!all_finite(b)) return IE_NONFINITE;
    op->diagonal = ne.diag;
    if (lambda > 0.0)
        for (M66& b : op->diagonal)
            for (int k = 0; k < 6; ++k) b(k, k) += lambda;
    for (const M66& b : op->diagonal)
        if (!all_finite(b)) return IE_NONFINITE;
    const size_t L = ne.lm.size();
    op->inv.assign(L, M3::zeros());
    op->has_inv.assign(L, 0);
    op->rhs.resize((size_t)6 * P);
    for (size_t k = 0; k < op->rhs.size(); ++k) op->rhs[k] = -ne.b_p[k];
    for (size_t l = 0; l < L; ++l) {
        const LandmarkBlock& lb = ne.lm[l];
        if (!all_finite(lb.h_ll) || !std::isfinite(lb.b_l[0]) || !std::isfinite(lb.b_l[1]) ||
            !std::isfinite(lb.b_l[2]))
            return IE_NONFINITE;
        for (const CrossEntry& c : lb.cross)
            if (!all_finite(c.w)) return IE_NONFINITE;
        M3 h = lb.h_ll;
        if (lambda > 0.0)
            for (int k = 0; k < 3; ++k) h(k, k) += lambda;
        if (!all_finite(h)) return IE_NONFINITE;
        M3 inv;
        bool ok = try_inverse3(h, &inv);
        if (ok && !all_finite(inv)) return IE_NONFINITE;
        op->has_inv[l] = ok;
        if (!ok) continue;
        op->inv[l] = inv;
        for (const CrossEntry& c : lb.cross) {
            M63 ci = mul(c.w, inv);
            double upd[6];
            mulv(ci, lb.b_l, upd);
            for (int k = 0; k < 6; ++k) op->rhs[c.pose * 6 + k] += upd[k];
        }
    }
    if (!all_finite(op->rhs)) return IE_NONFINITE;
    std::vector<M66> pre = op->diagonal;
    std::vector<std::pair<uint32_t, M63>> grouped;
    for (size_t l = 0; l < L; ++l) {
        if (!op->has_inv[l]) continue;
        grouped.clear();  // BTreeMap<pose, sum of cross>: ascending pose
        for (const CrossEntry& c : ne.lm[l].cross) {
            auto it = std::find_if(grouped.begin(), grouped.end(),
                                   [&](const std::pair<uint32_t, M63>& g) { return g.first == c.pose; });
            if (it == grouped.end())
                grouped.push_back({c.pose, add(M63::zeros(), c.w)});
            else
                add_assign(it->second, c.w);
        }
        std::sort(grouped.begin(), grouped.end(),
                  [](const std::pair<uint32_t, M63>& a, const std::pair<uint32_t, M63>& b) {
                      return a.first < b.first;
                  });
        for (auto& g : grouped) sub_assign(pre[g.first], mul(mul(g.second, op->inv[l]), tr(g.second)));
    }
    for (const M66& b : pre)
        if (!all_finite(b)) return IE_NONFINITE;
    op->precond_inv.resize(P);
    for (uint32_t p = 0; p < P; ++p) {
        if (!cholesky_inverse6(pre[p], &op->precond_inv[p])) return IE_NON_SPD_PRECONDITIONER;
        if (!all_finite(op->precond_inv[p])) return IE_NONFINITE;
    }
    return IE_OK;
}
// The same operator with the landmark inverses per landmark in parallel and
// the right-hand side and preconditioner sums per range of poses, each range
// visiting the landmarks in order (so every sum has the serial order).
int build_implicit_threaded(const Normal& ne, double lambda, int threads, Implicit* op) {
    if (!std::isfinite(lambda) || lambda < 0.0) return IE_INVALID_LAMBDA;
    if (ne.diag.empty()) return IE_EMPTY;
    op->ne = &ne;
    op->threads = threads;
    const uint32_t P = (uint32_t)ne.diag.size();
    if (!all_finite(ne.b_p)) return IE_NONFINITE;
    for (const M66& b : ne.diag)
        if (!all_finite(b)) return IE_NONFINITE;
    op->diagonal = ne.diag;
    if (lambda > 0.0)
        for (M66& b : op->diagonal)
            for (int k = 0; k < 6; ++k) b(k, k) += lambda;
    for (const M66& b : op->diagonal)
        if (!all_finite(b)) return IE_NONFINITE;
    const size_t L = ne.lm.size();
    op->inv.assign(L, M3::zeros());
    op->has_inv.assign(L, 0);
    std::vector<uint8_t> bad(L, 0);
    parallel_ranges(threads, L, [&](size_t l0, size_t l1) {
        for (size_t l = l0; l < l1; ++l) {
            const LandmarkBlock& lb = ne.lm[l];
            if (!all_finite(lb.h_ll) || !std::isfinite(lb.b_l[0]) || !std::isfinite(lb.b_l[1]) ||
                !std::isfinite(lb.b_l[2])) {
                bad[l] = 1;
                continue;
            }
            for (const CrossEntry& c : lb.cross)
                if (!all_finite(c.w)) bad[l] = 1;
            if (bad[l]) continue;
            M3 h = lb.h_ll;
            if (lambda > 0.0)
                for (int k = 0; k < 3; ++k) h(k, k) += lambda;
            if (!all_finite(h)) {
                bad[l] = 1;
                continue;
            }
            M3 inv;
            bool ok = try_inverse3(h, &inv);
            if (ok && !all_finite(inv)) {
                bad[l] = 1;
                continue;
            }
            op->has_inv[l] = ok;
            if (ok) op->inv[l] = inv;
        }
    });
    for (size_t l = 0; l < L; ++l)
        if (bad[l]) return IE_NONFINITE;
    // per-pose entry lists in (landmark, entry) order
    op->pose_ptr.assign((size_t)P + 1, 0);
    for (size_t l = 0; l < L; ++l)
        for (const CrossEntry& c : ne.lm[l].cross) ++op->pose_ptr[c.pose + 1];
    for (uint32_t p = 0; p < P; ++p) op->pose_ptr[p + 1] += op->pose_ptr[p];
    op->pose_entries.resize(op->pose_ptr[P]);
    {
        std::vector<size_t> fill(op->pose_ptr.begin(), op->pose_ptr.end() - 1);
        for (size_t l = 0; l < L; ++l)
            for (const CrossEntry& c : ne.lm[l].cross) op->pose_entries[fill[c.pose]++] = {(uint32_t)l, &c};
    }
    op->rhs.resize((size_t)6 * P);
    for (size_t k = 0; k < op->rhs.size(); ++k) op->rhs[k] = -ne.b_p[k];
    std::vector<M66> pre = op->diagonal;
    parallel_ranges(threads, P, [&](size_t p0, size_t p1) {
        std::vector<std::pair<uint32_t, M63>> grouped;
        for (size_t l = 0; l < L; ++l) {
            if (!op->has_inv[l]) continue;
            const LandmarkBlock& lb = ne.lm[l];
            grouped.clear();
            for (const CrossEntry& c : lb.cross) {
                if (c.pose < p0 || c.pose >= p1) continue;
                M63 ci = mul(c.w, op->inv[l]);
                double upd[6];
                mulv(ci, lb.b_l, upd);
                for (int k = 0; k < 6; ++k) op->rhs[c.pose * 6 + k] += upd[k];
                auto it = std::find_if(grouped.begin(), grouped.end(),
                                       [&](const std::pair<uint32_t, M63>& g) { return g.first == c.pose; });
                if (it == grouped.end())
                    grouped.push_back({c.pose, add(M63::zeros(), c.w)});
                else
                    add_assign(it->second, c.w);
            }
            for (auto& g : grouped) sub_assign(pre[g.first], mul(mul(g.second, op->inv[l]), tr(g.second)));
        }
    });
    if (!all_finite(op->rhs)) return IE_NONFINITE;
    for (const M66& b : pre)
        if (!all_finite(b)) return IE_NONFINITE;
    op->precond_inv.resize(P);
    std::vector<int> code(P, IE_OK);
    parallel_ranges(threads, P, [&](size_t p0, size_t p1) {
        for (size_t p = p0; p < p1; ++p) {
            if (!cholesky_inverse6(pre[p], &op->precond_inv[p])) code[p] = IE_NON_SPD_PRECONDITIONER;
            else if (!all_finite(op->precond_inv[p])) code[p] = IE_NONFINITE;
        }
    });
    for (uint32_t p = 0; p < P; ++p)
        if (code[p] != IE_OK) return code[p];
    return IE_OK;
}
int build_implicit(const Normal& ne, double lambda, Implicit* op, int threads = 1) {
    if (threads > 1) return build_implicit_threaded(ne, lambda, threads, op);
    return build_implicit_serial(ne, lambda, op);
}
// ImplicitSchurOperator::apply (implicit_schur.rs:390-431). With threads > 1
// the landmark reductions run per landmark and the pose sums per pose, each
// pose subtracting its entries in the serial (landmark, entry) order.
int implicit_apply(const Implicit& op, const std::vector<double>& x, std::vector<double>* out) {
    if (!all_finite(x)) return IE_NONFINITE;
    out->assign(op.dim(), 0.0);
    if (op.threads > 1) {
        const Normal& ne = *op.ne;
        const size_t L = ne.lm.size();
        std::vector<double> reduced(3 * L);
        parallel_ranges(op.threads, L, [&](size_t l0, size_t l1) {
            for (size_t l = l0; l < l1; ++l) {
                if (!op.has_inv[l]) continue;
                double projected[3] = {0.0, 0.0, 0.0};
                for (const CrossEntry& c : ne.lm[l].cross) {
                    double v[3];
                    mulv(tr(c.w), x.data() + c.pose * 6, v);
                    for (int k = 0; k < 3; ++k) projected[k] += v[k];
                }
                mulv(op.inv[l], projected, reduced.data() + 3 * l);
            }
        });
        parallel_ranges(op.threads, op.diagonal.size(), [&](size_t p0, size_t p1) {
            for (size_t p = p0; p < p1; ++p) {
                mulv(op.diagonal[p], x.data() + p * 6, out->data() + p * 6);
                for (size_t e = op.pose_ptr[p]; e < op.pose_ptr[p + 1]; ++e) {
                    const uint32_t l = op.pose_entries[e].first;
                    if (!op.has_inv[l]) continue;
                    double v[6];
                    mulv(op.pose_entries[e].second->w, reduced.data() + 3 * l, v);
                    for (int k = 0; k < 6; ++k) (*out)[p * 6 + k] -= v[k];
                }
            }
        });
        if (!all_finite(*out)) return IE_NONFINITE;
        return IE_OK;
    }
    for (size_t p = 0; p < op.diagonal.size(); ++p) mulv(op.diagonal[p], x.data() + p * 6, out->data() + p * 6);
    const Normal& ne = *op.ne;
    for (size_t l = 0; l < ne.lm.size(); ++l) {
        if (!op.has_inv[l]) continue;
        double projected[3] = {0.0, 0.0, 0.0};
        for (const CrossEntry& c : ne.lm[l].cross) {
            double v[3];
            mulv(tr(c.w), x.data() + c.pose * 6, v);
            for (int k = 0; k < 3; ++k) projected[k] += v[k];
        }
        double reduced[3];
        mulv(op.inv[l], projected, reduced);
        for (const CrossEntry& c : ne.lm[l].cross) {
            double v[6];
            mulv(c.w, reduced, v);
            for (int k = 0; k < 6; ++k) (*out)[c.pose * 6 + k] -= v[k];
        }
    }
    if (!all_finite(*out)) return IE_NONFINITE;
    return IE_OK;
}
int implicit_precondition(const Implicit& op, const std::vector<double>& r, std::vector<double>* out) {
    if (!all_finite(r)) return IE_NONFINITE;
    out->assign(op.dim(), 0.0);
    for (size_t p = 0; p < op.precond_inv.size(); ++p)
        mulv(op.precond_inv[p], r.data() + p * 6, out->data() + p * 6);
    if (!all_finite(*out)) return IE_NONFINITE;
    return IE_OK;
}
struct PcgOut {
    std::vector<double> x;
    uint32_t iterations = 0;
    bool has_iterations = false;
    double residual_norm = 0.0, target = 0.0;
    bool has_diag = false;
};
int true_residual(const Implicit& op, const std::vector<double>& rhs, const std::vector<double>& x,
                  std::vector<double>* tr_out, double* norm) {
    std::vector<double> ax;
    int e = implicit_apply(op, x, &ax);
    if (e) return e;
    tr_out->resize(rhs.size());
    for (size_t i = 0; i < rhs.size(); ++i) (*tr_out)[i] = rhs[i] - ax[i];
    *norm = dyn_norm(*tr_out);
    if (!std::isfinite(*norm)) return IE_NONFINITE;
    return IE_OK;
}
// solve_pcg_with_restart_hook with max_restarts = 0 (implicit_schur.rs:522-736).
// accept_iterate: a solve that ends unconverged returns its iterate instead
// of failing (the reference GPU backend's semantics).
int solve_pcg(const Implicit& op, uint32_t max_it, double rtol, double atol, bool accept_iterate,
              PcgOut* out) {
    const std::vector<double>& rhs = op.rhs;
    if (!all_finite(rhs)) return IE_NONFINITE;
    if (!std::isfinite(rtol) || rtol < 0.0 || !std::isfinite(atol) || atol < 0.0) return IE_INVALID_TOLERANCE;
    double rhs_norm = dyn_norm(rhs);
    double target = std::max(atol, rtol * rhs_norm);
    if (!std::isfinite(target)) return IE_NONFINITE;
    out->target = target;
    std::vector<double> x(op.dim(), 0.0), r = rhs, z, ad, trv;
    double rn = dyn_norm(r);
    if (!std::isfinite(rn)) return IE_NONFINITE;
    out->has_iterations = true;
    out->iterations = 0;
    if (rn <= target) {
        double tn;
        int e = true_residual(op, rhs, x, &trv, &tn);
        if (e) return e;
        out->residual_norm = tn;
        out->has_diag = true;
        out->x = x;
        if (tn <= target || accept_iterate) return IE_OK;
        return IE_RESIDUAL_CHECK_FAILED;
    }
    if (max_it == 0) {
        out->residual_norm = rn;
        out->has_diag = true;
        if (accept_iterate) {
            out->x = x;
            return IE_OK;
        }
        return IE_MAX_ITERATIONS;
    }
    int e = implicit_precondition(op, r, &z);
    if (e) return e;
    double rho = dyn_dot(r.data(), z.data(), r.size());
    if (!std::isfinite(rho) || rho <= 0.0) return IE_NON_POSITIVE_CURVATURE;
    std::vector<double> d = z;
    for (uint32_t it = 1; it <= max_it; ++it) {
        e = implicit_apply(op, d, &ad);
        if (e) return e;
        double curv = dyn_dot(d.data(), ad.data(), d.size());
        if (!std::isfinite(curv) || curv <= 0.0) return IE_NON_POSITIVE_CURVATURE;
        double alpha = rho / curv;
        if (!std::isfinite(alpha)) return IE_NONFINITE;
        for (size_t i = 0; i < x.size(); ++i) x[i] = x[i] + alpha * d[i];
        for (size_t i = 0; i < r.size(); ++i) r[i] = r[i] - alpha * ad[i];
        rn = dyn_norm(r);
        out->iterations = it;
        if (!all_finite(x) || !std::isfinite(rn)) return IE_NONFINITE;
        if (rn <= target) {
            double tn;
            e = true_residual(op, rhs, x, &trv, &tn);
            if (e) return e;
            out->residual_norm = tn;
            out->has_diag = true;
            out->x = x;
            if (tn <= target || accept_iterate) return IE_OK;
            return IE_RESIDUAL_CHECK_FAILED;
        }
        if (it == max_it) {
            double tn;
            e = true_residual(op, rhs, x, &trv, &tn);
            if (e) return e;
            out->residual_norm = tn;
            out->has_diag = true;
            out->x = x;
            if (tn <= target || accept_iterate) return IE_OK;
            return IE_MAX_ITERATIONS;
        }
        e = implicit_precondition(op, r, &z);
        if (e) return e;
        double next_rho = dyn_dot(r.data(), z.data(), r.size());
        if (!std::isfinite(next_rho) || next_rho <= 0.0) return IE_NON_POSITIVE_CURVATURE;
        double beta = next_rho / rho;
        if (!std::isfinite(beta)) return IE_NONFINITE;
        for (size_t i = 0; i < d.size(); ++i) d[i] = z[i] + beta * d[i];
        rho = next_rho;
    }
    return IE_MAX_ITERATIONS;
}
// complete_delta (implicit_schur.rs:789-824).
int complete_delta(const Implicit& op, const std::vector<double>& dp, std::vector<double>* dl) {
    if (!all_finite(dp)) return IE_NONFINITE;
    const Normal& ne = *op.ne;
    dl->assign(ne.lm.size() * 3, 0.0);
    parallel_ranges(op.threads, ne.lm.size(), [&](size_t l0, size_t l1) {
        for (size_t l = l0; l < l1; ++l) {
            if (!op.has_inv[l]) continue;
            double acc[3] = {-ne.lm[l].b_l[0], -ne.lm[l].b_l[1], -ne.lm[l].b_l[2]};
            for (const CrossEntry& c : ne.lm[l].cross) {
                double v[3];
                mulv(tr(c.w), dp.data() + c.pose * 6, v);
                for (int k = 0; k < 3; ++k) acc[k] -= v[k];
            }
            double d[3];
            mulv(op.inv[l], acc, d);
            for (int k = 0; k < 3; ++k) (*dl)[l * 3 + k] = d[k];
        }
    });
    if (!all_finite(*dl)) return IE_NONFINITE;
    return IE_OK;
}

// ---------------------------------------------------------------------------
// Matrix-free entry validation (optimize.rs:382-626), restricted to the
// terms this path carries.
// ---------------------------------------------------------------------------
int32_t validate_matrix_free(const Problem& P, const vba_options& o, const Kernel& K) {
    if (o.refine_intrinsics) return VBA_ERR_INELIGIBLE;
    if (o.max_iterations == 0) return VBA_ERR_INVALID_CONFIGURATION;
    if (!o.has_initial_lambda) return VBA_ERR_INVALID_CONFIGURATION;
    double il = o.initial_lambda;
    if (!(std::isfinite(il) && il > 0.0)) return VBA_ERR_INVALID_CONFIGURATION;
    if (!std::isfinite(o.lambda_increase_factor) || o.lambda_increase_factor <= 1.0 ||
        !std::isfinite(o.lambda_decrease_factor) || o.lambda_decrease_factor <= 0.0 ||
        o.lambda_decrease_factor >= 1.0 || !std::isfinite(o.max_lambda) || o.max_lambda < il ||
        !std::isfinite(o.min_lambda) || o.min_lambda <= 0.0 || o.min_lambda > il ||
        !std::isfinite(o.step_tolerance) || o.step_tolerance < 0.0 || !std::isfinite(o.cost_tolerance) ||
        o.cost_tolerance < 0.0 ||
        (o.has_relative_cost_tolerance &&
         (!std::isfinite(o.relative_cost_tolerance) || o.relative_cost_tolerance < 0.0)))
        return VBA_ERR_INVALID_CONFIGURATION;
    if (o.max_pcg_iterations == 0 || !std::isfinite(o.pcg_relative_tolerance) ||
        o.pcg_relative_tolerance < 0.0 || !std::isfinite(o.pcg_absolute_tolerance) ||
        o.pcg_absolute_tolerance <= 0.0)
        return VBA_ERR_INVALID_CONFIGURATION;
    if (K.kind != VBA_KERNEL_NONE && !(std::isfinite(K.scale) && K.scale > 0.0))
        return VBA_ERR_INVALID_CONFIGURATION;
    auto finite_cam = [](const Cam& c) {
        Intr in;
        if (!intrinsics(c, &in)) return false;
        for (int i = 0; i < c.np; ++i)
            if (!std::isfinite(c.p[i])) return false;
        return std::isfinite(in.fx) && std::isfinite(in.fy) && std::isfinite(in.cx) &&
               std::isfinite(in.cy) && in.fx > 0.0 && in.fy > 0.0;
    };
    if (!finite_cam(P.camera)) return VBA_ERR_INELIGIBLE;
    bool any_fixed = false, any_var = false;
    for (uint8_t f : P.pose_fixed) (f ? any_fixed : any_var) = true;
    if (!any_fixed) return VBA_ERR_INELIGIBLE;
    if (!any_var) return VBA_ERR_INELIGIBLE;
    for (const SE3& T : P.poses) {
        const double v[7] = {T.q.w, T.q.i, T.q.j, T.q.k, T.t.x, T.t.y, T.t.z};
        for (double x : v)
            if (!std::isfinite(x)) return VBA_ERR_INELIGIBLE;
    }
    for (const V3& X : P.lms)
        if (!std::isfinite(X.x) || !std::isfinite(X.y) || !std::isfinite(X.z)) return VBA_ERR_INELIGIBLE;
    double c;
    uint64_t np;
    cost_and_nonprojectable(P, K, &c, &np);
    if (!std::isfinite(c)) return VBA_ERR_INELIGIBLE;
    auto finite_pose = [](const SE3& T) {
        const double v[7] = {T.q.w, T.q.i, T.q.j, T.q.k, T.t.x, T.t.y, T.t.z};
        for (double x : v)
            if (!std::isfinite(x)) return false;
        return true;
    };
    // optimize.rs:546-622: pixels, and the rig / stereo pair calibration
    for (const Obs& ob : P.obs) {
        if (!std::isfinite(ob.x) || !std::isfinite(ob.y)) return VBA_ERR_INELIGIBLE;
        if ((ob.kind == OBS_RECTIFIED || ob.kind == OBS_GENERAL) && !std::isfinite(ob.x2))
            return VBA_ERR_INELIGIBLE;
        if (ob.kind == OBS_GENERAL && !std::isfinite(ob.y2)) return VBA_ERR_INELIGIBLE;
        if (ob.kind == OBS_RIG || ob.kind == OBS_GENERAL) {
            const Sensor& s = ob.kind == OBS_RIG ? P.sensors[ob.sensor] : P.pairs[ob.sensor];
            if (!finite_cam(s.cam) || !finite_pose(s.sensor_from_rig)) return VBA_ERR_INELIGIBLE;
        }
    }
    return VBA_OK;
}

// ---------------------------------------------------------------------------
// LM loop: optimize_weighted_backend (lm.rs:162-1017).
// ---------------------------------------------------------------------------
double norm6_seq(const double* v) {  // Vector6::norm (generic dot, n < 8: sequential)
    double res = 0.0;
    for (int k = 0; k < 6; ++k) res += v[k] * v[k];
    return std::sqrt(res);
}

int32_t run_lm(Problem& P, const vba_options& o, int threads, vba_iteration_stats* stats,
               uint32_t cap, vba_result* res) {
    const Kernel K{o.robust_kernel, o.robust_scale};
    const bool matrix_free = o.linear_solver == VBA_SOLVER_MATRIX_FREE || o.linear_solver == VBA_SOLVER_PCG_EXPLICIT;
    const bool accept_iterate = o.pcg_maxit_policy == VBA_PCG_MAXIT_ACCEPT_ITERATE;
    double initial_cost;
    uint64_t current_nonproj;
    cost_and_nonprojectable(P, K, &initial_cost, &current_nonproj, threads);
    double current_cost = initial_cost;
    double lambda = o.has_initial_lambda ? o.initial_lambda : 0.0;
    bool converged = false;
    uint32_t written = 0, total_pcg = 0, linearizations = 0;
    auto push = [&](const vba_iteration_stats& s) {
        if (written < cap && stats) stats[written] = s;
        ++written;
    };
    Normal ne;
    for (uint32_t iteration = 0; iteration < o.max_iterations; ++iteration) {
        build_normal_equations(P, K, matrix_free, threads, &ne);
        ++linearizations;
        std::vector<SE3> saved_poses = P.poses;
        std::vector<V3> saved_lms = P.lms;
        const double cost_before = current_cost;
        std::vector<double> dp, dl;
        int32_t solve_status = VBA_OK;
        vba_iteration_stats st;
        std::memset(&st, 0, sizeof(st));
        st.iteration = iteration;
        st.pcg_iterations = -1;
        if (!matrix_free) {
            solve_status = solve_step_dense(ne, P.L, lambda, &dp, &dl, threads,
                                            o.linear_solver != VBA_SOLVER_SPARSE);
        } else {
            Implicit op;
            int e = build_implicit(ne, lambda, &op, threads);
            PcgOut pcg;
            if (e == IE_OK) e = solve_pcg(op, o.max_pcg_iterations, o.pcg_relative_tolerance,
                                          o.pcg_absolute_tolerance, accept_iterate, &pcg);
            if (pcg.has_iterations && (e == IE_OK || e == IE_MAX_ITERATIONS || e == IE_RESIDUAL_CHECK_FAILED)) {
                st.pcg_iterations = (int32_t)pcg.iterations;
                st.pcg_residual_norm = pcg.residual_norm;
                st.pcg_target = pcg.target;
                total_pcg += pcg.iterations;
            }
            if (e == IE_OK) {
                dp = pcg.x;
                e = complete_delta(op, dp, &dl);
            }
            if (e != IE_OK) solve_status = VBA_ERR_SINGULAR_SYSTEM;
        }
        if (solve_status != VBA_OK) {
            lambda = std::min(lambda * o.lambda_increase_factor, o.max_lambda);
            st.cost_before = cost_before;
            st.cost_after = cost_before;
            st.lambda = lambda;
            st.step_accepted = 0;
            st.nonprojectable_after = (uint32_t)current_nonproj;
            push(st);
            if (lambda >= o.max_lambda) break;
            continue;
        }
        double max_pose_step = 0.0;
        for (size_t s = 0; s < P.poses.size(); ++s) {
            int32_t i = P.pose_var[s];
            if (i < 0) continue;
            const double* xi = dp.data() + (size_t)i * 6;
            double step = norm6_seq(xi);
            if (step > max_pose_step) max_pose_step = step;
            P.poses[s] = se3_compose(P.poses[s], se3_exp(xi));
        }
        double max_landmark_step = 0.0;
        for (size_t s = 0; s < P.lms.size(); ++s) {
            int32_t i = P.lm_var[s];
            if (i < 0) continue;
            const double* d = dl.data() + (size_t)i * 3;
            V3 v{d[0], d[1], d[2]};
            double step = std::sqrt(norm_sq3(v));
            if (step > max_landmark_step) max_landmark_step = step;
            P.lms[s] = {P.lms[s].x + v.x, P.lms[s].y + v.y, P.lms[s].z + v.z};
        }
        double cost_after;
        uint64_t nonproj_after;
        cost_and_nonprojectable(P, K, &cost_after, &nonproj_after, threads);
        bool cost_accepted = o.has_initial_lambda ? (cost_after < cost_before) : true;
        bool feasible = nonproj_after <= current_nonproj;
        bool accepted = cost_accepted && feasible;
        st.cost_before = cost_before;
        st.cost_after = cost_after;
        st.max_pose_step = max_pose_step;
        st.max_landmark_step = max_landmark_step;
        st.nonprojectable_after = (uint32_t)nonproj_after;
        if (!accepted) {
            P.poses = saved_poses;
            P.lms = saved_lms;
            lambda = std::min(lambda * o.lambda_increase_factor, o.max_lambda);
            st.lambda = lambda;
            st.step_accepted = 0;
            push(st);
            if (!o.has_initial_lambda) break;
            if (lambda >= o.max_lambda) break;
            continue;
        }
        st.lambda = lambda;
        st.step_accepted = 1;
        push(st);
        current_cost = cost_after;
        current_nonproj = nonproj_after;
        if (o.has_initial_lambda) lambda = std::max(lambda * o.lambda_decrease_factor, o.min_lambda);
        if (max_pose_step < o.step_tolerance && max_landmark_step < o.step_tolerance) {
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
    res->initial_cost = initial_cost;
    res->final_cost = current_cost;
    res->num_iterations = std::min(written, cap);
    res->converged = converged;
    res->num_observations_used = P.obs.size();
    res->total_pcg_iterations = total_pcg;
    res->linearizations = linearizations;
    return VBA_OK;
}

// ---------------------------------------------------------------------------
// Joint pose / landmark / intrinsics refinement: optimize_joint_intrinsics
// (joint_intrinsics.rs:24-353) with its dense camera system
// (build_joint_intrinsics_system, :593-825; the monocular terms :356-465).
// ---------------------------------------------------------------------------
struct JointLm {
    M3 h_ll;
    double b_l[3];
    std::map<size_t, std::vector<double>> cross;  // column start -> rows x 3, row-major
};
void joint_add_cross(JointLm& L, size_t cs, int rows, const double* blk) {
    auto it = L.cross.find(cs);
    if (it == L.cross.end()) {
        L.cross.emplace(cs, std::vector<double>(blk, blk + 3 * rows));
        return;
    }
    for (int i = 0; i < 3 * rows; ++i) it->second[i] += blk[i];
}
// joint_mono_terms (:356-465): residual, J_pose, J_landmark and the camera
// columns J_k (2 x k_dim, row-major) of a monocular observation.
bool joint_mono_terms(const Problem& P, const Obs& o, const Intr& in, int k_dim, bool has_dist, double k1,
                      double k2, bool has_tan, double r[2], M26* jp, M23* jl, double* jk) {
    const SE3& T = P.poses[o.pose];
    const V3& X = P.lms[o.lm];
    const V3 xc = se3_apply(T, X);
    if (xc.z <= 0.0) return false;
    const double fx = in.fx, fy = in.fy;
    const double z_inv = 1.0 / xc.z;
    const double x = xc.x * z_inv, y = xc.y * z_inv;
    const double r2 = x * x + y * y;
    if (!has_dist) k1 = k2 = 0.0;
    const double d = 1.0 + k1 * r2 + k2 * r2 * r2;
    const double g = k1 + 2.0 * k2 * r2;
    double xd = x * d, yd = y * d;
    r[0] = (fx * xd + in.cx) - o.x;
    r[1] = (fy * yd + in.cy) - o.y;
    const M3 R = qmatrix(T.q);
    const double d11 = d + 2.0 * x * x * g, d12 = 2.0 * x * y * g, d22 = d + 2.0 * y * y * g;
    M23 jpi = M23::zeros();
    jpi(0, 0) = fx * d11 * z_inv;
    jpi(0, 1) = fx * d12 * z_inv;
    jpi(0, 2) = -fx * (d11 * x + d12 * y) * z_inv;
    jpi(1, 0) = fy * d12 * z_inv;
    jpi(1, 1) = fy * d22 * z_inv;
    jpi(1, 2) = -fy * (d12 * x + d22 * y) * z_inv;
    const M36 dx = pose_point_jacobian(R, X);
    *jp = mul(jpi, dx);
    *jl = mul(jpi, R);
    for (int i = 0; i < 2 * k_dim; ++i) jk[i] = 0.0;
    if (has_tan) {
        double uv[2];
        M23 J;
        if (!lens_project_with_jacobian(P.camera, xc, uv, &J)) return false;
        r[0] = uv[0] - o.x;
        r[1] = uv[1] - o.y;
        *jp = mul(J, dx);
        *jl = mul(J, R);
        const double p1 = P.camera.np > 6 ? P.camera.p[6] : 0.0, p2 = P.camera.np > 7 ? P.camera.p[7] : 0.0;
        const double xy2 = 2.0 * x * y;
        xd = x * d + p1 * xy2 + p2 * (r2 + 2.0 * x * x);
        yd = y * d + p1 * (r2 + 2.0 * y * y) + p2 * xy2;
    }
    jk[0] = xd;
    jk[2] = 1.0;
    jk[k_dim + 1] = yd;
    jk[k_dim + 3] = 1.0;
    if (k_dim >= 6) {
        jk[4] = fx * x * r2;
        jk[5] = fx * x * r2 * r2;
        jk[k_dim + 4] = fy * y * r2;
        jk[k_dim + 5] = fy * y * r2 * r2;
    }
    if (has_tan && k_dim == 8) {
        const double xy2 = 2.0 * x * y;
        jk[6] = fx * xy2;
        jk[7] = fx * (r2 + 2.0 * x * x);
        jk[k_dim + 6] = fy * (r2 + 2.0 * y * y);
        jk[k_dim + 7] = fy * xy2;
    }
    return true;
}
// One observation's blocks into the dense camera system (rows x 6 J_p,
// rows x 3 J_l, rows x kd J_k, row-major).
void joint_accumulate(const Problem& P, const Obs& o, const Kernel& K, int rows, const double* r, const double* jp,
                      const double* jl, const double* jk, int kd, size_t k_off, size_t cam_dim,
                      std::vector<double>& H, std::vector<double>& b, std::vector<JointLm>& lms) {
    double s = 0.0;
    for (int i = 0; i < rows; ++i) s += r[i] * r[i];
    const double w = K.weight(s);
    auto dot_cols = [&](const double* A, int na, int a, const double* B, int nb, int c) {
        double acc = 0.0;
        for (int i = 0; i < rows; ++i) acc += A[i * na + a] * B[i * nb + c];
        return acc;
    };
    for (int a = 0; a < kd; ++a) {
        for (int c = 0; c < kd; ++c) H[(k_off + a) * cam_dim + k_off + c] += w * dot_cols(jk, kd, a, jk, kd, c);
        double acc = 0.0;
        for (int i = 0; i < rows; ++i) acc += jk[i * kd + a] * r[i];
        b[k_off + a] += w * acc;
    }
    const int32_t p = P.pose_var[o.pose], l = P.lm_var[o.lm];
    if (p >= 0) {
        const size_t p0 = 6 * (size_t)p;
        for (int a = 0; a < 6; ++a) {
            for (int c = 0; c < 6; ++c) H[(p0 + a) * cam_dim + p0 + c] += w * dot_cols(jp, 6, a, jp, 6, c);
            double acc = 0.0;
            for (int i = 0; i < rows; ++i) acc += jp[i * 6 + a] * r[i];
            b[p0 + a] += w * acc;
            for (int c = 0; c < kd; ++c) {
                const double v = w * dot_cols(jp, 6, a, jk, kd, c);
                H[(p0 + a) * cam_dim + k_off + c] += v;
                H[(k_off + c) * cam_dim + p0 + a] += v;
            }
        }
    }
    if (l >= 0) {
        JointLm& L = lms[l];
        for (int a = 0; a < 3; ++a) {
            for (int c = 0; c < 3; ++c) L.h_ll(a, c) += w * dot_cols(jl, 3, a, jl, 3, c);
            double acc = 0.0;
            for (int i = 0; i < rows; ++i) acc += jl[i * 3 + a] * r[i];
            L.b_l[a] += w * acc;
        }
        if (p >= 0) {
            double cr[18];
            for (int a = 0; a < 6; ++a)
                for (int c = 0; c < 3; ++c) cr[a * 3 + c] = w * dot_cols(jp, 6, a, jl, 3, c);
            joint_add_cross(L, 6 * (size_t)p, 6, cr);
        }
        double ck[24];
        for (int a = 0; a < kd; ++a)
            for (int c = 0; c < 3; ++c) ck[a * 3 + c] = w * dot_cols(jk, kd, a, jl, 3, c);
        joint_add_cross(L, k_off, kd, ck);
    }
}

constexpr uint32_t kRefineIntrinsics = 1, kRefineDistortion = 2, kRefineTangential = 4, kSharedFocal = 8;

int32_t run_joint(Problem& P, const vba_options& o, vba_iteration_stats* stats, uint32_t cap, vba_result* res) {
    const Kernel K{o.robust_kernel, o.robust_scale};
    const uint32_t flags = o.refine_intrinsics;
    bool any_stereo = false;
    for (const Obs& ob : P.obs)
        if (ob.kind == OBS_RECTIFIED) any_stereo = true;
    const size_t k_off = 6 * (size_t)P.P;
    Cam& cam = P.camera;
    const bool refine_dist = (flags & kRefineDistortion) &&
                             (cam.model == VBA_CAMERA_PINHOLE || cam.model == VBA_CAMERA_OPENCV) && !any_stereo;
    if (refine_dist)
        while (cam.np < 6) cam.p[cam.np++] = 0.0;
    const bool refine_tan = refine_dist && (flags & kRefineTangential);
    if (refine_tan) {
        cam.model = VBA_CAMERA_OPENCV;
        while (cam.np < 8) cam.p[cam.np++] = 0.0;
    }
    const int k_dim = refine_tan ? 8 : refine_dist ? 6 : 4;
    const size_t cam_dim = k_off + k_dim;
    const bool shared = flags & kSharedFocal;
    if (shared) {
        const double f = 0.5 * (cam.p[0] + cam.p[1]);
        cam.p[0] = f;
        cam.p[1] = f;
    }
    double initial_cost;
    uint64_t current_nonproj;
    cost_and_nonprojectable(P, K, &initial_cost, &current_nonproj);
    double current_cost = initial_cost;
    double lambda = o.has_initial_lambda ? o.initial_lambda : 0.0;
    bool converged = false;
    uint32_t written = 0, linearizations = 0;
    auto push = [&](const vba_iteration_stats& s) {
        if (written < cap && stats) stats[written] = s;
        ++written;
    };
    for (uint32_t iteration = 0; iteration < o.max_iterations; ++iteration) {
        double k1 = 0.0, k2 = 0.0;
        const bool has_dist = radial_distortion(cam, &k1, &k2);
        const bool has_tan = refine_tan || tangential_distortion(cam);
        Intr in;
        intrinsics(cam, &in);
        // build_joint_intrinsics_system, dense
        std::vector<double> H(cam_dim * cam_dim, 0.0), bc(cam_dim, 0.0);
        std::vector<JointLm> lms(P.L);
        for (JointLm& L : lms) {
            L.h_ll = M3::zeros();
            L.b_l[0] = L.b_l[1] = L.b_l[2] = 0.0;
        }
        for (const Obs& ob : P.obs) {
            if (ob.kind != OBS_MONO) continue;
            double r[2], jk[16];
            M26 jp;
            M23 jl;
            if (!joint_mono_terms(P, ob, in, k_dim, has_dist, k1, k2, has_tan, r, &jp, &jl, jk)) continue;
            joint_accumulate(P, ob, K, 2, r, jp.a, jl.a, jk, k_dim, k_off, cam_dim, H, bc, lms);
        }
        for (const Obs& ob : P.obs) {
            if (ob.kind != OBS_RECTIFIED) continue;
            const SE3& T = P.poses[ob.pose];
            const V3& X = P.lms[ob.lm];
            const V3 xc = se3_apply(T, X);
            if (xc.z <= 0.0) continue;
            double uv[2];
            if (!project_pinhole(in, xc, uv)) continue;
            const double b = P.baseline;
            const double z_inv = 1.0 / xc.z, z_inv2 = z_inv * z_inv;
            const double u_r = uv[0] - in.fx * b * z_inv;
            const double r[3] = {uv[0] - ob.x, uv[1] - ob.y, u_r - ob.x2};
            const M3 R = qmatrix(T.q);
            M3 jpi = M3::zeros();
            jpi(0, 0) = in.fx * z_inv;
            jpi(0, 2) = -in.fx * xc.x * z_inv2;
            jpi(1, 1) = in.fy * z_inv;
            jpi(1, 2) = -in.fy * xc.y * z_inv2;
            jpi(2, 0) = in.fx * z_inv;
            jpi(2, 2) = -in.fx * (xc.x - b) * z_inv2;
            const M36 jp = mul(jpi, pose_point_jacobian(R, X));
            const M3 jl = mul(jpi, R);
            double jk[12] = {xc.x * z_inv, 0, 1, 0, 0, xc.y * z_inv, 0, 1, (xc.x - b) * z_inv, 0, 1, 0};
            joint_accumulate(P, ob, K, 3, r, jp.a, jl.a, jk, 4, k_off, cam_dim, H, bc, lms);
        }
        ++linearizations;
        // damped Schur reduction (joint_intrinsics.rs:143-185)
        std::vector<double> S = H;
        if (lambda > 0.0)
            for (size_t d = 0; d < cam_dim; ++d) S[d * cam_dim + d] += lambda;
        std::vector<double> bred(cam_dim);
        for (size_t i = 0; i < cam_dim; ++i) bred[i] = -bc[i];
        std::vector<bool> has_inv(P.L, false);
        std::vector<M3> invs(P.L);
        for (uint32_t l = 0; l < P.L; ++l) {
            M3 h = lms[l].h_ll;
            if (lambda > 0.0) {
                h(0, 0) += lambda;
                h(1, 1) += lambda;
                h(2, 2) += lambda;
            }
            has_inv[l] = try_inverse3(h, &invs[l]);
            if (!has_inv[l]) continue;
            const M3& inv = invs[l];
            for (const auto& ea : lms[l].cross) {
                const int ra = (int)ea.second.size() / 3;
                std::vector<double> ah(3 * ra);
                for (int i = 0; i < ra; ++i)
                    for (int c = 0; c < 3; ++c) {
                        double acc = 0.0;
                        for (int t = 0; t < 3; ++t) acc += ea.second[i * 3 + t] * inv(t, c);
                        ah[i * 3 + c] = acc;
                    }
                for (const auto& eb : lms[l].cross) {
                    const int rb = (int)eb.second.size() / 3;
                    for (int i = 0; i < ra; ++i)
                        for (int j = 0; j < rb; ++j) {
                            double acc = 0.0;
                            for (int t = 0; t < 3; ++t) acc += ah[i * 3 + t] * eb.second[j * 3 + t];
                            S[(ea.first + i) * cam_dim + eb.first + j] -= acc;
                        }
                }
                for (int i = 0; i < ra; ++i) {
                    double acc = 0.0;
                    for (int t = 0; t < 3; ++t) acc += ah[i * 3 + t] * lms[l].b_l[t];
                    bred[ea.first + i] += acc;
                }
            }
        }
        std::vector<double> delta;
        bool ok;
        if (shared) {
            // fy folded into fx: S' = T^T S T, b' = T^T b, delta = T delta'
            const size_t fyi = k_off + 1, n1 = cam_dim - 1;
            auto map = [&](size_t i) { return i < fyi ? i : i == fyi ? k_off : i - 1; };
            std::vector<double> Sr(n1 * n1, 0.0), br(n1, 0.0), dr;
            for (size_t i = 0; i < cam_dim; ++i) {
                br[map(i)] += bred[i];
                for (size_t j = 0; j < cam_dim; ++j) Sr[map(i) * n1 + map(j)] += S[i * cam_dim + j];
            }
            ok = solve_normal_equations(Sr, n1, br, &dr);
            if (ok) {
                delta.resize(cam_dim);
                for (size_t i = 0; i < cam_dim; ++i) delta[i] = dr[map(i)];
            }
        } else {
            ok = solve_normal_equations(S, cam_dim, bred, &delta);
        }
        vba_iteration_stats st;
        std::memset(&st, 0, sizeof(st));
        st.iteration = iteration;
        st.pcg_iterations = -1;
        const double cost_before = current_cost;
        if (!ok) {
            lambda = std::min(lambda * o.lambda_increase_factor, o.max_lambda);
            st.cost_before = st.cost_after = cost_before;
            st.lambda = lambda;
            st.nonprojectable_after = (uint32_t)current_nonproj;
            push(st);
            if (lambda >= o.max_lambda) break;
            continue;
        }
        // back-substitution (:219-233)
        std::vector<V3> dl(P.L, V3{0, 0, 0});
        for (uint32_t l = 0; l < P.L; ++l) {
            if (!has_inv[l]) continue;
            double acc[3] = {-lms[l].b_l[0], -lms[l].b_l[1], -lms[l].b_l[2]};
            for (const auto& ea : lms[l].cross) {
                const int ra = (int)ea.second.size() / 3;
                for (int c = 0; c < 3; ++c) {
                    double t = 0.0;
                    for (int i = 0; i < ra; ++i) t += ea.second[i * 3 + c] * delta[ea.first + i];
                    acc[c] -= t;
                }
            }
            const M3& inv = invs[l];
            dl[l] = {inv(0, 0) * acc[0] + inv(0, 1) * acc[1] + inv(0, 2) * acc[2],
                     inv(1, 0) * acc[0] + inv(1, 1) * acc[1] + inv(1, 2) * acc[2],
                     inv(2, 0) * acc[0] + inv(2, 1) * acc[1] + inv(2, 2) * acc[2]};
        }
        const std::vector<SE3> saved_poses = P.poses;
        const std::vector<V3> saved_lms = P.lms;
        const Cam saved_cam = cam;
        double max_pose_step = 0.0, max_landmark_step = 0.0;
        for (size_t s = 0; s < P.poses.size(); ++s) {
            const int32_t i = P.pose_var[s];
            if (i < 0) continue;
            const double* xi = delta.data() + 6 * (size_t)i;
            max_pose_step = std::max(max_pose_step, norm6_seq(xi));
            P.poses[s] = se3_compose(P.poses[s], se3_exp(xi));
        }
        for (size_t s = 0; s < P.lms.size(); ++s) {
            const int32_t i = P.lm_var[s];
            if (i < 0 || !has_inv[i]) continue;
            const V3 v = dl[i];
            max_landmark_step = std::max(max_landmark_step, std::sqrt(norm_sq3(v)));
            P.lms[s] = {P.lms[s].x + v.x, P.lms[s].y + v.y, P.lms[s].z + v.z};
        }
        for (int j = 0; j < k_dim; ++j) cam.p[j] += delta[k_off + j];
        double cost_after;
        uint64_t nonproj_after;
        cost_and_nonprojectable(P, K, &cost_after, &nonproj_after);
        const bool cost_accepted = o.has_initial_lambda ? (cost_after < cost_before) : true;
        const bool accepted = cost_accepted && nonproj_after <= current_nonproj;
        st.cost_before = cost_before;
        st.cost_after = cost_after;
        st.max_pose_step = max_pose_step;
        st.max_landmark_step = max_landmark_step;
        st.nonprojectable_after = (uint32_t)nonproj_after;
        if (!accepted) {
            P.poses = saved_poses;
            P.lms = saved_lms;
            cam = saved_cam;
            lambda = std::min(lambda * o.lambda_increase_factor, o.max_lambda);
            st.lambda = lambda;
            push(st);
            if (!o.has_initial_lambda) break;
            if (lambda >= o.max_lambda) break;
            continue;
        }
        st.lambda = lambda;
        st.step_accepted = 1;
        push(st);
        current_cost = cost_after;
        current_nonproj = nonproj_after;
        if (o.has_initial_lambda) lambda = std::max(lambda * o.lambda_decrease_factor, o.min_lambda);
        if (max_pose_step < o.step_tolerance && max_landmark_step < o.step_tolerance) {
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
    res->initial_cost = initial_cost;
    res->final_cost = current_cost;
    res->num_iterations = std::min(written, cap);
    res->converged = converged;
    res->num_observations_used = P.obs.size();
    res->linearizations = linearizations;
    return VBA_OK;
}

// Build a Normal from the C normal-system view.
void normal_from_view(const oracle_normal_system* v, Normal* ne) {
    ne->P = v->num_poses;
    ne->dense = false;
    ne->diag.resize(v->num_poses);
    for (uint32_t p = 0; p < v->num_poses; ++p)
        for (int i = 0; i < 36; ++i) ne->diag[p].a[i] = v->diag[p * 36 + i];
    ne->b_p.assign(v->b_p, v->b_p + 6 * (size_t)v->num_poses);
    ne->lm.resize(v->num_landmarks);
    for (uint32_t l = 0; l < v->num_landmarks; ++l) {
        LandmarkBlock& lb = ne->lm[l];
        for (int i = 0; i < 9; ++i) lb.h_ll.a[i] = v->h_ll[l * 9 + i];
        for (int i = 0; i < 3; ++i) lb.b_l[i] = v->b_l[l * 3 + i];
        lb.cross.clear();
        for (uint32_t e = v->cross_ptr[l]; e < v->cross_ptr[l + 1]; ++e) {
            CrossEntry c;
            c.pose = v->cross_pose[e];
            for (int i = 0; i < 18; ++i) c.w.a[i] = v->cross[e * 18 + i];
            lb.cross.push_back(c);
        }
    }
}

// reprojection_squared_residuals (problem.rs:756-...): per observation in
// the flattened order, NaN when it cannot be evaluated (no intrinsics, behind
// the camera, or no projection). Monocular observations use project_pinhole
// for a distortion-free camera and Camera::project otherwise.
std::vector<double> reprojection_squared_residuals(const Problem& P) {
    std::vector<double> out(P.obs.size(), std::numeric_limits<double>::quiet_NaN());
    Intr in;
    if (!intrinsics(P.camera, &in)) return out;
    const Proj pr = mono_projection_for(P.camera);
    for (size_t i = 0; i < P.obs.size(); ++i) {
        const Obs& o = P.obs[i];
        if (o.kind == OBS_MONO) {
            const V3 xc = se3_apply(P.poses[o.pose], P.lms[o.lm]);
            if (xc.z <= 0.0) continue;
            double uv[2];
            const bool ok = pr.kind == PROJ_PINHOLE ? project_pinhole(in, xc, uv) : camera_project(P.camera, xc, uv);
            if (!ok) continue;
            const double rx = uv[0] - o.x, ry = uv[1] - o.y;
            out[i] = rx * rx + ry * ry;
        } else {
            double s;
            if (squared_residual(P, in, o, &s)) out[i] = s;
        }
    }
    return out;
}

// GncState (gnc.rs:118-227).
constexpr double kTlsMuTerminal = 200.0;
struct GncState {
    uint32_t kernel;
    double c2, factor, mu;
    GncState(const vba_gnc_options& g, double max_residual_sq) : kernel(g.kernel), c2(g.c * g.c),
                                                                  factor(g.anneal_factor) {
        const double s_max = std::max(max_residual_sq, 0.0);
        if (kernel == VBA_GNC_GEMAN_MCCLURE) {
            mu = std::max(2.0 * s_max / c2, 1.0);
        } else {
            const double denom = 2.0 * s_max - c2;
            mu = denom <= 0.0 ? kTlsMuTerminal
                              : std::min(std::max(c2 / denom, std::numeric_limits<double>::min()), kTlsMuTerminal);
        }
    }
    void set_inlier_scale(double c) { c2 = c * c; }
    double weight(double s) const {
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
    bool is_terminal() const { return kernel == VBA_GNC_GEMAN_MCCLURE ? mu <= 1.0 : mu >= kTlsMuTerminal; }
    bool anneal() {
        if (kernel == VBA_GNC_GEMAN_MCCLURE) mu = std::max(mu / factor, 1.0);
        else mu = std::min(mu * factor, kTlsMuTerminal);
        return is_terminal();
    }
};

// median_in_place / estimate_scale_mad (gnc.rs:252-286).
double median_in_place(std::vector<double>& v) {
    std::sort(v.begin(), v.end(), [](double a, double b) {
        return a < b || (a == b && std::signbit(a) && !std::signbit(b));  // f64::total_cmp on non-NaN values
    });
    const size_t n = v.size();
    return n % 2 == 1 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}
bool estimate_scale_mad(const std::vector<double>& squared, double k, double* out) {
    std::vector<double> norms;
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

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

}  // extern "C"

namespace {

void write_state(const vba_problem* problem, const Problem& P, double* poses_out, double* landmarks_out) {
    for (uint64_t i = 0; i < problem->num_poses; ++i)
        if (poses_out) pose_to7(P.poses[P.pose_slot_of_input[i]], poses_out + 7 * i);
    for (uint64_t i = 0; i < problem->num_landmarks; ++i) {
        if (!landmarks_out) break;
        const V3& X = P.lms[P.lm_slot_of_input[i]];
        landmarks_out[3 * i] = X.x;
        landmarks_out[3 * i + 1] = X.y;
        landmarks_out[3 * i + 2] = X.z;
    }
}

// optimize / optimize_weighted (lm.rs:153-160) on a loaded problem whose
// observations carry their weights.
int32_t optimize_loaded(Problem& P, const vba_options& o, int32_t threads, vba_iteration_stats* stats,
                        uint32_t cap, vba_result* result) {
    const bool matrix_free = o.linear_solver == VBA_SOLVER_MATRIX_FREE || o.linear_solver == VBA_SOLVER_PCG_EXPLICIT;
    if (matrix_free) {
        int32_t st = validate_matrix_free(P, o, Kernel{o.robust_kernel, o.robust_scale});
        if (st != VBA_OK) return st;
    }
    return run_lm(P, o, threads, stats, cap, result);
}

}  // namespace

extern "C" {

int32_t oracle_optimize(const vba_problem* problem, const vba_options* options, double* poses_out,
                        double* landmarks_out, vba_iteration_stats* stats, uint32_t stats_capacity,
                        vba_result* result, int32_t threads) {
    if (!problem || !options || !result) return VBA_ERR_INVALID_ARGUMENT;
    std::memset(result, 0, sizeof(*result));
    if (problem->declined_features) return VBA_DECLINED;
    Problem P;
    uint64_t error_id = 0;
    int32_t st = load_problem(problem, &P, &error_id);
    result->error_id = error_id;
    if (st != VBA_OK) return st;
    st = optimize_loaded(P, *options, threads, stats, stats_capacity, result);
    if (st != VBA_OK) return st;
    write_state(problem, P, poses_out, landmarks_out);
    return VBA_OK;
}

int32_t oracle_optimize_intrinsics(const vba_problem* problem, const vba_options* options, double* poses_out,
                                   double* landmarks_out, vba_camera* camera_out, vba_iteration_stats* stats,
                                   uint32_t stats_capacity, vba_result* result, int32_t threads) {
    if (!problem || !options || !result || !camera_out) return VBA_ERR_INVALID_ARGUMENT;
    std::memset(result, 0, sizeof(*result));
    *camera_out = problem->camera;
    if (problem->declined_features) return VBA_DECLINED;
    const vba_stereo_terms* st3 = problem->stereo;
    const bool general = st3 && st3->num_general;
    const int model = problem->camera.model;
    Intr in0;
    // BundleAdjustment::optimize (optimize.rs:8-28): the joint solve for a
    // PINHOLE / OPENCV camera with intrinsics and no general stereo or rig
    // observations, else the pose / structure solve
    const bool joint = (options->refine_intrinsics & kRefineIntrinsics) && !general &&
                       problem->num_rig_observations == 0 &&
                       (model == VBA_CAMERA_PINHOLE || model == VBA_CAMERA_OPENCV) &&
                       intrinsics(to_cam(problem->camera), &in0);
    Problem P;
    uint64_t error_id = 0;
    int32_t st = load_problem(problem, &P, &error_id);
    result->error_id = error_id;
    if (!joint) {
        if (st != VBA_OK) return st;
        vba_options plain = *options;
        plain.refine_intrinsics = 0;
        st = optimize_loaded(P, plain, threads, stats, stats_capacity, result);
        if (st != VBA_OK) return st;
        write_state(problem, P, poses_out, landmarks_out);
        return VBA_OK;
    }
    // the joint solve refines the camera even with every pose and landmark fixed
    if (st != VBA_OK && st != VBA_ERR_ALL_POSES_FIXED) return st;
    st = run_joint(P, *options, stats, stats_capacity, result);
    if (st != VBA_OK) return st;
    write_state(problem, P, poses_out, landmarks_out);
    camera_out->model = P.camera.model;
    camera_out->num_params = (uint32_t)P.camera.np;
    for (int i = 0; i < 12; ++i) camera_out->params[i] = i < P.camera.np ? P.camera.p[i] : 0.0;
    return VBA_OK;
}

int32_t oracle_optimize_weighted(const vba_problem* problem, const vba_options* options, const double* weights,
                                 uint64_t num_weights, double* poses_out, double* landmarks_out,
                                 vba_iteration_stats* stats, uint32_t stats_capacity, vba_result* result,
                                 int32_t threads) {
    if (!problem || !options || !result) return VBA_ERR_INVALID_ARGUMENT;
    std::memset(result, 0, sizeof(*result));
    if (problem->declined_features & ~(uint32_t)VBA_FEATURE_OBSERVATION_WEIGHTS) return VBA_DECLINED;
    // validate_observation_weights (optimize.rs:628-650), then the intrinsics
    // refusal (:49-51)
    const uint64_t expected = problem->num_observations + problem->num_rig_observations +
                              (problem->stereo ? problem->stereo->num_stereo + problem->stereo->num_general : 0);
    if (num_weights != expected || (num_weights && !weights)) return VBA_ERR_OBSERVATION_WEIGHT_COUNT;
    for (uint64_t i = 0; i < num_weights; ++i)
        if (!std::isfinite(weights[i]) || weights[i] < 0.0) {
            result->error_id = i;
            return VBA_ERR_INVALID_OBSERVATION_WEIGHT;
        }
    if (options->refine_intrinsics) return VBA_ERR_WEIGHTS_WITH_INTRINSICS;
    Problem P;
    uint64_t error_id = 0;
    int32_t st = load_problem(problem, &P, &error_id);
    result->error_id = error_id;
    if (st != VBA_OK) return st;
    for (size_t i = 0; i < P.obs.size(); ++i) P.obs[i].weight = weights[i];
    st = optimize_loaded(P, *options, threads, stats, stats_capacity, result);
    if (st != VBA_OK) return st;
    write_state(problem, P, poses_out, landmarks_out);
    return VBA_OK;
}

void oracle_default_gnc_options(vba_gnc_options* g) {
    std::memset(g, 0, sizeof(*g));
    g->kernel = VBA_GNC_GEMAN_MCCLURE;
    g->c = 1.0;
    g->anneal_factor = 1.4;
    g->max_outer = 100;
    g->inner_iterations = 5;
}

// BundleAdjustment::optimize_gnc (lm.rs:28-142).
int32_t oracle_optimize_gnc(const vba_problem* problem, const vba_options* options, const vba_gnc_options* gnc,
                            double* poses_out, double* landmarks_out, double* weights_out, vba_gnc_result* res,
                            int32_t threads) {
    if (!problem || !options || !gnc || !res) return VBA_ERR_INVALID_ARGUMENT;
    std::memset(res, 0, sizeof(*res));
    if (problem->declined_features) return VBA_DECLINED;
    Problem P;
    uint64_t error_id = 0;
    int32_t st = load_problem(problem, &P, &error_id);
    if (st != VBA_OK) return st;
    const Kernel none{VBA_KERNEL_NONE, 0.0};
    uint64_t nonproj = 0;
    double initial_cost = 0.0;
    cost_and_nonprojectable(P, none, &initial_cost, &nonproj);
    const size_t n = P.obs.size();
    std::vector<double> sq = reprojection_squared_residuals(P);
    double s_max = 0.0;
    for (double s : sq)
        if (std::isfinite(s)) s_max = std::max(s_max, s);
    vba_gnc_options eff = *gnc;
    if (gnc->has_auto_scale) {
        double est;
        if (estimate_scale_mad(sq, gnc->auto_scale, &est)) eff.c = std::max(est, gnc->c);
    }
    double inlier_scale = eff.c;
    GncState state(eff, s_max);
    vba_options inner = *options;
    inner.robust_kernel = VBA_KERNEL_NONE;
    inner.max_iterations = std::max<uint32_t>(gnc->inner_iterations, 1);
    std::vector<double> w(n, 1.0);
    bool converged = false;
    uint32_t outer = 0;
    for (uint32_t it = 0; it < std::max<uint32_t>(gnc->max_outer, 1); ++it) {
        ++outer;
        const bool terminal = state.is_terminal();
        const std::vector<double> r = reprojection_squared_residuals(P);
        if (gnc->auto_scale_readapt && gnc->has_auto_scale) {
            double est;
            if (estimate_scale_mad(r, gnc->auto_scale, &est)) {
                const double c = std::max(est, gnc->c);
                state.set_inlier_scale(c);
                inlier_scale = c;
            }
        }
        for (size_t i = 0; i < n; ++i) {
            w[i] = std::isfinite(r[i]) ? state.weight(r[i]) : 1.0;
            P.obs[i].weight = w[i];
        }
        vba_result inner_res{};
        st = optimize_loaded(P, inner, threads, nullptr, 0, &inner_res);
        if (st != VBA_OK) return st;
        if (terminal) {
            converged = true;
            break;
        }
        state.anneal();
    }
    const std::vector<double> r = reprojection_squared_residuals(P);
    for (size_t i = 0; i < n; ++i) {
        w[i] = std::isfinite(r[i]) ? state.weight(r[i]) : std::numeric_limits<double>::quiet_NaN();
        P.obs[i].weight = w[i];
    }
    double final_cost = 0.0;
    cost_and_nonprojectable(P, none, &final_cost, &nonproj);
    for (size_t i = 0; i < n; ++i)
        P.obs[i].weight = std::isfinite(w[i]) ? (w[i] >= 0.5 ? 1.0 : 0.0) : std::numeric_limits<double>::quiet_NaN();
    double inlier_cost = 0.0;
    cost_and_nonprojectable(P, none, &inlier_cost, &nonproj);
    res->initial_cost = initial_cost;
    res->final_cost = final_cost;
    res->inlier_cost = inlier_cost;
    res->inlier_scale = inlier_scale;
    res->observation_count = n;
    res->outer_iterations = outer;
    res->converged = converged ? 1 : 0;
    write_state(problem, P, poses_out, landmarks_out);
    if (weights_out) std::copy(w.begin(), w.end(), weights_out);
    return VBA_OK;
}

int32_t oracle_gnc_state_weights(const vba_gnc_options* gnc, double max_residual_sq, uint32_t anneal_steps,
                                 const double* s, uint32_t n, double* weights, double* mu) {
    GncState st(*gnc, max_residual_sq);
    for (uint32_t i = 0; i < anneal_steps; ++i) st.anneal();
    for (uint32_t i = 0; i < n; ++i) weights[i] = st.weight(s[i]);
    if (mu) *mu = st.mu;
    return st.is_terminal() ? 1 : 0;
}

int32_t oracle_estimate_scale_mad(const double* squared, uint32_t n, double k, double* out) {
    return estimate_scale_mad(std::vector<double>(squared, squared + n), k, out) ? 1 : 0;
}

int32_t oracle_camera_project(const vba_camera* camera, const double xc[3], double uv[2]) {
    return camera_project(to_cam(*camera), V3{xc[0], xc[1], xc[2]}, uv) ? 1 : 0;
}

int32_t oracle_mono_project(const vba_camera* camera, const double xc[3], double uv[2], double j[6]) {
    Cam c = to_cam(*camera);
    Intr in;
    if (!intrinsics(c, &in)) return 0;
    M23 J;
    if (!mono_project_with_jacobian(mono_projection_for(c), c, in, V3{xc[0], xc[1], xc[2]}, uv, &J)) return 0;
    for (int i = 0; i < 6; ++i) j[i] = J.a[i];
    return 1;
}

void oracle_robust(uint32_t kernel, double scale, double s, double* cost, double* weight) {
    Kernel K{kernel, scale};
    if (cost) *cost = K.cost(s);
    if (weight) *weight = K.weight(s);
}

int32_t oracle_linearize(const vba_problem* problem, uint32_t kernel, double scale, double* r,
                         double* j_pose, double* j_landmark, double* w, uint8_t* valid) {
    Problem P;
    uint64_t eid;
    int32_t st = load_problem(problem, &P, &eid);
    if (st != VBA_OK) return st;
    for (const Obs& o : P.obs)
        if (o.kind == OBS_RECTIFIED || o.kind == OBS_GENERAL) return VBA_ERR_INVALID_ARGUMENT;  // 2-row layout only
    Kernel K{kernel, scale};
    Intr in;
    intrinsics(P.camera, &in);
    Proj pr = mono_projection_for(P.camera);
    for (size_t i = 0; i < P.obs.size(); ++i) {
        Lin lin;
        bool ok = linearize_obs(P, in, pr, P.obs[i], &lin);
        valid[i] = ok;
        if (!ok) {
            std::memset(r + 2 * i, 0, 2 * sizeof(double));
            std::memset(j_pose + 12 * i, 0, 12 * sizeof(double));
            std::memset(j_landmark + 6 * i, 0, 6 * sizeof(double));
            w[i] = 0.0;
            continue;
        }
        r[2 * i] = lin.r[0];
        r[2 * i + 1] = lin.r[1];
        for (int k = 0; k < 12; ++k) j_pose[12 * i + k] = lin.jp.a[k];
        for (int k = 0; k < 6; ++k) j_landmark[6 * i + k] = lin.jl.a[k];
        w[i] = K.weight(lin.r[0] * lin.r[0] + lin.r[1] * lin.r[1]);
    }
    return VBA_OK;
}

int32_t oracle_cost(const vba_problem* problem, uint32_t kernel, double scale, double* cost,
                    uint64_t* nonprojectable) {
    Problem P;
    uint64_t eid;
    int32_t st = load_problem(problem, &P, &eid);
    if (st != VBA_OK) return st;
    cost_and_nonprojectable(P, Kernel{kernel, scale}, cost, nonprojectable);
    return VBA_OK;
}

void oracle_se3_exp(const double xi[6], double pose[7]) { pose_to7(se3_exp(xi), pose); }
void oracle_se3_log(const double pose[7], double xi[6]) { se3_log(pose_from7(pose), xi); }
void oracle_se3_compose(const double a[7], const double b[7], double out[7]) {
    pose_to7(se3_compose(pose_from7(a), pose_from7(b)), out);
}
void oracle_camera_center(const double pose[7], double center[3]) {
    SE3 inv = se3_inverse(pose_from7(pose));
    center[0] = inv.t.x;
    center[1] = inv.t.y;
    center[2] = inv.t.z;
}

void oracle_ns_reduce_dense(const oracle_normal_system* ns, double lambda, double* s, double* rhs) {
    Normal ne;
    normal_from_view(ns, &ne);
    const uint32_t P = ne.P;
    const size_t dim = (size_t)6 * P;
    std::vector<double> S(dim * dim, 0.0);
    for (uint32_t p = 0; p < P; ++p)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) S[(p * 6 + r) * dim + p * 6 + c] = ne.diag[p](r, c);
    if (lambda > 0.0)
        for (size_t k = 0; k < dim; ++k) S[k * dim + k] += lambda;
    std::vector<double> b(dim);
    for (size_t k = 0; k < dim; ++k) b[k] = -ne.b_p[k];
    for (const LandmarkBlock& lb : ne.lm) {
        M3 inv;
        if (!try_inverse3(damped3(lb.h_ll, lambda), &inv)) continue;
        for (const CrossEntry& a : lb.cross) {
            M63 a_h = mul(a.w, inv);
            for (const CrossEntry& bq : lb.cross) {
                M66 blk = mul(a_h, tr(bq.w));
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) S[(a.pose * 6 + r) * dim + bq.pose * 6 + c] -= blk(r, c);
            }
            double upd[6];
            mulv(a_h, lb.b_l, upd);
            for (int k = 0; k < 6; ++k) b[a.pose * 6 + k] += upd[k];
        }
    }
    std::copy(S.begin(), S.end(), s);
    std::copy(b.begin(), b.end(), rhs);
}

int32_t oracle_ns_solve_dense(const oracle_normal_system* ns, double lambda, double* delta_p,
                              double* delta_l) {
    Normal ne;
    normal_from_view(ns, &ne);
    std::vector<double> dp, dl;
    int32_t st = solve_step_dense(ne, ns->num_landmarks, lambda, &dp, &dl);
    if (st != VBA_OK) return st;
    std::copy(dp.begin(), dp.end(), delta_p);
    std::copy(dl.begin(), dl.end(), delta_l);
    return VBA_OK;
}

int32_t oracle_ns_operator(const oracle_normal_system* ns, double lambda, const double* x, double* y,
                           const double* r, double* z, double* rhs) {
    Normal ne;
    normal_from_view(ns, &ne);
    Implicit op;
    int e = build_implicit(ne, lambda, &op);
    if (e) return e;
    const size_t dim = op.dim();
    if (x && y) {
        std::vector<double> xv(x, x + dim), yv;
        e = implicit_apply(op, xv, &yv);
        if (e) return e;
        std::copy(yv.begin(), yv.end(), y);
    }
    if (r && z) {
        std::vector<double> rv(r, r + dim), zv;
        e = implicit_precondition(op, rv, &zv);
        if (e) return e;
        std::copy(zv.begin(), zv.end(), z);
    }
    if (rhs) std::copy(op.rhs.begin(), op.rhs.end(), rhs);
    return 0;
}

int32_t oracle_ns_solve_pcg(const oracle_normal_system* ns, double lambda, uint32_t max_iterations,
                            double rtol, double atol, int32_t accept_at_cap, double* delta_p,
                            double* delta_l, uint32_t* iterations, double* residual_norm,
                            double* target) {
    Normal ne;
    normal_from_view(ns, &ne);
    Implicit op;
    int e = build_implicit(ne, lambda, &op);
    if (e) return e;
    PcgOut out;
    e = solve_pcg(op, max_iterations, rtol, atol, accept_at_cap != 0, &out);
    if (iterations) *iterations = out.iterations;
    if (residual_norm) *residual_norm = out.residual_norm;
    if (target) *target = out.target;
    if (e) return e;
    std::vector<double> dl;
    e = complete_delta(op, out.x, &dl);
    if (e) return e;
    std::copy(out.x.begin(), out.x.end(), delta_p);
    std::copy(dl.begin(), dl.end(), delta_l);
    return 0;
}

int32_t oracle_cholesky(uint32_t n, const double* a, double* l, int32_t threads) {
    std::vector<double> L(a, a + (size_t)n * n);
    if (!cholesky_dense(L, n, threads)) return 0;
    std::copy(L.begin(), L.end(), l);
    return 1;
}

int32_t oracle_solve_normal_equations(uint32_t n, const double* a, const double* b, double* x) {
    std::vector<double> A(a, a + (size_t)n * n), B(b, b + n), X;
    if (!solve_normal_equations(A, n, B, &X)) return 0;
    std::copy(X.begin(), X.end(), x);
    return 1;
}

// ---------------------------------------------------------------------------
// Point elimination of the COLMAP-port rig solver's reduced camera system
// (pipelines/slam/src/colmap_incremental/rig_ba_solver.rs): the off-diagonal
// blocks of eliminate_by_output (:2551-2660), with the shard partial sums of
// elimination_shard_count (:1903-1928) and balanced_shard_boundaries
// (:2902-2925), sub_outer_6x3's fused multiply-adds (:2531-2539) and the
// product hcp[k] * inv (nalgebra's column sweep); then the direct couplings,
// the run blocks and the aggregate-pair sums in edge order that
// ReducedSolveAccelerator::eliminate returns (ba_accel.rs:71-157).
// ---------------------------------------------------------------------------
int32_t oracle_reduced_eliminate(const vba_point_elimination* el, double* blocks, double* run_blocks,
                                 double* coarse_sums) {
    const uint32_t n = el->num_nodes, E = el->num_edges, NP = el->num_points;
    // shards: a pure function of the edge count and the points' track lengths
    const size_t per_shard = (size_t)std::max<uint32_t>(E, 1) * 288;
    const size_t shards = std::min<size_t>(8, std::max<size_t>(1, (size_t)1536 * 1024 * 1024 / per_shard));
    std::vector<uint64_t> prefix(NP + 1, 0);
    for (uint32_t p = 0; p < NP; ++p) {
        const uint64_t k = el->point_off[p + 1] - el->point_off[p];
        prefix[p + 1] = prefix[p] + k * (k > 0 ? k - 1 : 0) / 2;
    }
    std::vector<size_t> bounds{0};
    for (size_t sh = 1; sh < shards; ++sh) {
        const uint64_t target = prefix[NP] * sh / shards;
        size_t idx = std::partition_point(prefix.begin(), prefix.end(), [&](uint64_t v) { return v < target; }) -
                     prefix.begin();
        idx = std::min<size_t>(std::max(idx, bounds.back()), NP);
        bounds.push_back(idx);
    }
    bounds.push_back(NP);
    std::vector<uint32_t> shard_of(NP, 0);
    for (size_t sh = 0; sh + 1 < bounds.size(); ++sh)
        for (size_t p = bounds[sh]; p < bounds[sh + 1]; ++p) shard_of[p] = (uint32_t)sh;
    // block(e) in column-major order, rows t = later node, columns s
    std::fill(blocks, blocks + 36 * (size_t)E, 0.0);
    std::vector<double> acc, scaled(18);
    std::vector<uint32_t> position(n, 0);
    for (uint32_t s = 0; s < n; ++s) {
        const uint64_t r0 = el->row_off[s], r1 = el->row_off[s + 1];
        for (uint64_t e = r0; e < r1; ++e) position[el->edges[2 * e + 1]] = (uint32_t)(e - r0);
        acc.assign(36 * (r1 - r0), 0.0);
        double* total = blocks + 36 * r0;
        int64_t current = -1;
        auto flush = [&]() {
            for (size_t i = 0; i < acc.size(); ++i) {
                total[i] += acc[i];
                acc[i] = 0.0;
            }
        };
        for (uint64_t q = el->slot_off[s]; q < el->slot_off[s + 1]; ++q) {
            const uint32_t p = el->slots[2 * q], k = el->slots[2 * q + 1];
            if (!el->has_inverse[p]) continue;
            if (current != (int64_t)shard_of[p]) {
                if (current >= 0) flush();
                current = shard_of[p];
            }
            const uint64_t po = el->point_off[p];
            const double* hk = el->hcp + 18 * (po + k);
            const double* inv = el->point_inverse + 9 * (size_t)p;
            // scaled = hcp[k] * inv: column c = sum_i hcp.col(i) inv(i, c), i ascending
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 6; ++r) {
                    double v = hk[r] * inv[3 * c];
                    v = v + hk[6 + r] * inv[3 * c + 1];
                    v = v + hk[12 + r] * inv[3 * c + 2];
                    scaled[6 * c + r] = v;
                }
            const uint64_t pn = el->point_off[p + 1] - po;
            for (uint64_t j = k + 1; j < pn; ++j) {
                const uint32_t t = el->point_nodes[po + j];
                double* a = acc.data() + 36 * (size_t)position[t];
                const double* hj = el->hcp + 18 * (po + j);
                for (int c = 0; c < 6; ++c) {
                    const double b0 = -scaled[c], b1 = -scaled[6 + c], b2 = -scaled[12 + c];
                    for (int r = 0; r < 6; ++r) {
                        double& v = a[6 * c + r];
                        v = std::fma(hj[12 + r], b2, std::fma(hj[6 + r], b1, std::fma(hj[r], b0, v)));
                    }
                }
            }
        }
        if (current >= 0) flush();
    }
    for (uint32_t x = 0; x < el->num_extra; ++x) {
        double* b = blocks + 36 * (size_t)el->extra_edge[x];
        for (int i = 0; i < 36; ++i) b[i] = b[i] + el->extra_blocks[36 * (size_t)x + i];
    }
    if (run_blocks)
        for (uint32_t i = 0; i < el->num_run_edges; ++i)
            std::copy(blocks + 36 * (size_t)el->run_edges[i], blocks + 36 * (size_t)el->run_edges[i] + 36,
                      run_blocks + 36 * (size_t)i);
    if (coarse_sums && el->coarse_group > 0) {
        const uint32_t grp = el->coarse_group, nodes = el->coarse_nodes;
        const uint32_t m = (nodes + grp - 1) / grp;
        std::fill(coarse_sums, coarse_sums + 36 * (size_t)m * m, 0.0);
        for (uint32_t g = 0; g < m; ++g)
            for (uint32_t h = g; h < m; ++h) {
                double sum[36] = {0.0};
                const uint32_t t0 = h * grp, t1 = std::min((h + 1) * grp, nodes);
                for (uint32_t a = g * grp; a < std::min((g + 1) * grp, nodes); ++a)
                    for (uint64_t e = el->row_off[a]; e < el->row_off[a + 1]; ++e) {
                        const uint32_t b = el->edges[2 * e + 1];
                        if (b < t0 || b >= t1) continue;
                        for (int i = 0; i < 36; ++i) sum[i] = sum[i] + blocks[36 * e + i];
                    }
                std::copy(sum, sum + 36, coarse_sums + 36 * ((size_t)g * m + h));
            }
    }
    return VBA_OK;
}

}  // extern "C"
