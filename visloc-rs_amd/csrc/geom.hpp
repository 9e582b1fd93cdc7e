// geom.hpp — pose, lens and robust-kernel math shared by host and device code.
//
// Poses are T_world->camera as a unit quaternion (w, x, y, z) + translation,
// the layout of visloc-rs `SE3` (crates/core/src/geometry/se3.rs:3-7). The
// pose update is the right perturbation T <- T * Exp([rho; omega])
// (se3.rs:72-82; pipelines/slam/src/bundle/lm.rs:672-690).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VBA_HD __host__ __device__ inline
#else
#define VBA_HD inline
#endif

#define VBA_ISFINITE(x) __builtin_isfinite(x)

namespace vba {

template <typename T>
struct V3t {
    T x, y, z;
};
using V3 = V3t<double>;

template <typename T>
VBA_HD V3t<T> cross3(const V3t<T>& a, const V3t<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T>
VBA_HD T dot3(const V3t<T>& a, const V3t<T>& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

struct Quat {
    double w, x, y, z;
};
VBA_HD Quat qmul(const Quat& a, const Quat& b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
// Rotate p by the unit quaternion q: p + 2w (v x p) + 2 v x (v x p).
VBA_HD V3 qrot(const Quat& q, const V3& p) {
    V3 v{q.x, q.y, q.z};
    V3 c = cross3(v, p);
    V3 t{2.0 * c.x, 2.0 * c.y, 2.0 * c.z};
    V3 c2 = cross3(v, t);
    return {p.x + q.w * t.x + c2.x, p.y + q.w * t.y + c2.y, p.z + q.w * t.z + c2.z};
}
// Row-major rotation matrix of a unit quaternion.
template <typename T>
VBA_HD void qmatrix(const Quat& q, T R[9]) {
    double ww = q.w * q.w, xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    double xy = 2.0 * q.x * q.y, wz = 2.0 * q.w * q.z, wy = 2.0 * q.w * q.y;
    double xz = 2.0 * q.x * q.z, yz = 2.0 * q.y * q.z, wx = 2.0 * q.w * q.x;
    R[0] = (T)(ww + xx - yy - zz); R[1] = (T)(xy - wz); R[2] = (T)(wy + xz);
    R[3] = (T)(xy + wz); R[4] = (T)(ww - xx + yy - zz); R[5] = (T)(yz - wx);
    R[6] = (T)(xz - wy); R[7] = (T)(wx + yz); R[8] = (T)(ww - xx - yy + zz);
}

struct Pose {
    Quat q;
    V3 t;
};
VBA_HD V3 pose_apply(const Pose& T, const V3& p) {
    V3 r = qrot(T.q, p);
    return {r.x + T.t.x, r.y + T.t.y, r.z + T.t.z};
}
VBA_HD Pose pose_compose(const Pose& a, const Pose& b) {
    V3 r = qrot(a.q, b.t);
    return {qmul(a.q, b.q), {r.x + a.t.x, r.y + a.t.y, r.z + a.t.z}};
}
VBA_HD Pose pose_inverse(const Pose& a) {
    Quat qi{a.q.w, -a.q.x, -a.q.y, -a.q.z};
    V3 r = qrot(qi, a.t);
    return {qi, {-r.x, -r.y, -r.z}};
}
// SE3::exp([rho; omega]) (se3.rs:72-82): q = exp(omega/2) as a quaternion,
// t = V(omega) rho with the SO(3) left Jacobian V (se3.rs:107-120).
VBA_HD Pose se3_exp(const double xi[6]) {
    V3 w{xi[3], xi[4], xi[5]};
    V3 h{0.5 * w.x, 0.5 * w.y, 0.5 * w.z};
    double nn = dot3(h, h);
    Quat q{1.0, 0.0, 0.0, 0.0};
    if (nn > 2.220446049250313e-16 * 2.220446049250313e-16) {
        double n = sqrt(nn);
        double s = sin(n) / n;
        q = {cos(n), h.x * s, h.y * s, h.z * s};
    }
    double th2 = dot3(w, w);
    double a, b;
    if (th2 < 1e-10) {
        a = 0.5;
        b = 1.0 / 6.0;
    } else {
        double th = sqrt(th2);
        a = (1.0 - cos(th)) / th2;
        b = (th - sin(th)) / (th * th2);
    }
    // V rho = rho + a (w x rho) + b (w x (w x rho))
    V3 rho{xi[0], xi[1], xi[2]};
    V3 c1 = cross3(w, rho);
    V3 c2 = cross3(w, c1);
    return {q, {rho.x + a * c1.x + b * c2.x, rho.y + a * c1.y + b * c2.y, rho.z + a * c1.z + b * c2.z}};
}

// ---------------------------------------------------------------------------
// Lens models, resolved on the host from a vba_camera the way
// MonoProjection::for_camera does (pipelines/slam/src/bundle/jacobians.rs:
// 29-58): PINHOLE (no distortion), RADIAL (1 + k1 r^2 + k2 r^4) or the full
// Brown–Conrady OPENCV lens with tangential terms (camera.rs:852-941).
// ---------------------------------------------------------------------------
enum LensKind : int32_t { LENS_PINHOLE = 0, LENS_RADIAL = 1, LENS_OPENCV = 2 };

template <typename T>
struct Lens {
    int32_t kind;
    int32_t pad;
    T fx, fy, cx, cy, k1, k2, p1, p2;
};

// Pixel prediction and its 2x3 Jacobian w.r.t. the camera-frame point.
// Returns false where the reference's projection returns None (behind the
// camera, vanishing OPENCV depth, non-finite result).
template <typename T>
VBA_HD bool project(const Lens<T>& L, T X, T Y, T Z, T uv[2], T J[6]) {
    if (!(Z > T(0))) return false;
    if (L.kind == LENS_OPENCV && !(Z > T(1e-12))) return false;
    const T zi = T(1) / Z;
    const T x = X * zi, y = Y * zi;
    if (L.kind == LENS_PINHOLE) {
        uv[0] = L.fx * x + L.cx;
        uv[1] = L.fy * y + L.cy;
        if (J) {
            J[0] = L.fx * zi; J[1] = T(0); J[2] = -L.fx * x * zi;
            J[3] = T(0); J[4] = L.fy * zi; J[5] = -L.fy * y * zi;
        }
        return true;
    }
    const T r2 = x * x + y * y;
    const T d = T(1) + L.k1 * r2 + L.k2 * r2 * r2;
    T a, b, c, e, xd, yd;
    const T g = L.k1 + T(2) * L.k2 * r2;  // d(radial)/d(r^2)
    if (L.kind == LENS_RADIAL) {
        xd = x * d;
        yd = y * d;
        a = d + T(2) * x * x * g;
        b = T(2) * x * y * g;
        c = b;
        e = d + T(2) * y * y * g;
    } else {
        xd = x * d + T(2) * L.p1 * x * y + L.p2 * (r2 + T(2) * x * x);
        yd = y * d + L.p1 * (r2 + T(2) * y * y) + T(2) * L.p2 * x * y;
        const T cr = T(2) * x * y * g + T(2) * L.p1 * x + T(2) * L.p2 * y;
        a = d + T(2) * x * x * g + T(2) * L.p1 * y + T(6) * L.p2 * x;
        b = cr;
        c = cr;
        e = d + T(2) * y * y * g + T(6) * L.p1 * y + T(2) * L.p2 * x;
    }
    uv[0] = L.fx * xd + L.cx;
    uv[1] = L.fy * yd + L.cy;
    if (!(VBA_ISFINITE(uv[0]) && VBA_ISFINITE(uv[1]))) return false;
    if (J) {
        J[0] = L.fx * a * zi; J[1] = L.fx * b * zi; J[2] = -L.fx * (a * x + b * y) * zi;
        J[3] = L.fy * c * zi; J[4] = L.fy * e * zi; J[5] = -L.fy * (c * x + e * y) * zi;
        for (int i = 0; i < 6; ++i)
            if (!VBA_ISFINITE(J[i])) return false;
    }
    return true;
}

// RobustKernel (pipelines/slam/src/pose_graph.rs:102-139): Huber
// rho(s) = s (s <= d^2) else 2 d sqrt(s) - d^2, w = 1 or d / sqrt(s);
// Cauchy rho(s) = c^2 ln(1 + s/c^2), w = 1 / (1 + s/c^2).
template <typename T>
VBA_HD T robust_weight(uint32_t kind, T scale, T s) {
    if (kind == 1) return s <= scale * scale ? T(1) : scale / sqrt(s);
    if (kind == 2) return T(1) / (T(1) + s / (scale * scale));
    return T(1);
}
template <typename T>
VBA_HD T robust_cost(uint32_t kind, T scale, T s) {
    if (kind == 1) return s <= scale * scale ? s : T(2) * scale * sqrt(s) - scale * scale;
    if (kind == 2) return scale * scale * log(T(1) + s / (scale * scale));
    return s;
}

}  // namespace vba
