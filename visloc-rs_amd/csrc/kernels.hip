// kernels.hip — CDNA4 (gfx950) kernels of the Levenberg–Marquardt Schur path.
//
// Work units: camera tiles (a slice of one pose's c-order edges) and landmark
// tiles (a run of whole landmarks in l-order). Every reduction is a fixed
// tree inside one work-group (wave shuffles, then LDS) followed by an
// in-order sum over tiles, so results are deterministic and independent of
// scheduling; no global atomics are used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "kernels.hpp"

#include "device_common.hpp"

namespace vba {

namespace {

// Rows of a 2x3 projection Jacobian times a 3x3 matrix A, into J (3 per row).
__device__ __forceinline__ void rows_times(const double Jp[6], const double A[9], double* J) {
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            J[3 * k + c] = Jp[3 * k] * A[c] + Jp[3 * k + 1] * A[3 + c] + Jp[3 * k + 2] * A[6 + c];
}

// Residual rows r and landmark Jacobian rows J (3 per row, J_lm = J_pi R) of
// a stereo observation at xr, the point in the pose (left camera) frame, as
// four rows (a rectified observation's fourth row is zero):
//   rectified (normal_equations.rs:486-571): (u_l, v_l, u_r) with
//     u_r = u_l - fx b / Z, J_pi's third row (fx/Z, 0, -fx (X - b)/Z^2);
//   general (sqrt_qr.rs:465-515): the left pinhole rows, then the right
//     camera's pinhole rows at T_right<-left xr, J = J_pi_r R_lr R.
// Both project with the camera's plain pinhole (fx, fy, cx, cy). Returns
// false where the reference skips the observation (a point at or behind
// either camera).
template <bool JAC>
__device__ __forceinline__ bool stereo_rows(const DevProblem& P, int o, const DevSensor* sensors, const DevSensor& S,
                                            const double xr[3], const double R[9], double r[4], double J[12]) {
    const uint8_t kind = P.obs_kind[o];
    Lens<double> L = (kind == OBS_GENERAL ? sensors[S.left] : S).lens;
    L.kind = LENS_PINHOLE;
    const double ex0 = P.obs_extra[2 * (size_t)o], ex1 = P.obs_extra[2 * (size_t)o + 1];
    double uv[2], Jp[6];
    if (!project<double>(L, xr[0], xr[1], xr[2], uv, JAC ? Jp : (double*)nullptr)) return false;
    r[0] = uv[0] - P.obs_uv[2 * (size_t)o];
    r[1] = uv[1] - P.obs_uv[2 * (size_t)o + 1];
    if (JAC) rows_times(Jp, R, J);
    if (kind == OBS_RECTIFIED) {
        const double b = ex1;
        r[2] = (uv[0] - L.fx * b / xr[2]) - ex0;
        r[3] = 0.0;
        if (JAC) {
            const double zi = 1.0 / xr[2];
            const double j0 = L.fx * zi, j2 = -L.fx * (xr[0] - b) * (zi * zi);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                J[6 + c] = j0 * R[c] + j2 * R[6 + c];
                J[9 + c] = 0.0;
            }
        }
        return true;
    }
    double xs[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) xs[i] = S.R[3 * i] * xr[0] + S.R[3 * i + 1] * xr[1] + S.R[3 * i + 2] * xr[2] + S.t[i];
    double uvr[2];
    if (!project<double>(S.lens, xs[0], xs[1], xs[2], uvr, JAC ? Jp : (double*)nullptr)) return false;
    r[2] = uvr[0] - ex0;
    r[3] = uvr[1] - ex1;
    if (JAC) {
        double M[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                M[3 * i + j] = S.R[3 * i] * R[j] + S.R[3 * i + 1] * R[3 + j] + S.R[3 * i + 2] * R[6 + j];
        rows_times(Jp, M, J + 6);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Linearization: residuals, Jacobians and IRLS weights in FP64 at the current
// state, G and h per edge into c-order and l-order records, and the tile's
// share of U_p, g_p (normal_equations.rs:418-484, 617-651).
// ---------------------------------------------------------------------------
template <bool RIG>
__global__ __launch_bounds__(kBlock) void k_linearize(DevProblem P, int store32) {
    __shared__ double lds[(kBlock / 64) * 27];
    const int tile = xcd_tile(blockIdx.x, P.num_cam_tiles);
    if (tile >= P.num_cam_tiles) return;
    if (P.wstate && !P.wstate[P.cam_tile_win[tile]].relin) return;
    const CamTile ct = P.cam_tiles[tile];
    const double* pq = P.poses + 7 * (size_t)ct.pose;
    double R[9];
    qmatrix<double>(Quat{pq[0], pq[1], pq[2], pq[3]}, R);
    const double t0 = pq[4], t1 = pq[5], t2 = pq[6];
    double acc[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) acc[i] = 0.0;
    for (int e = ct.begin + (int)threadIdx.x; e < ct.end; e += kBlock) {
        const int l = P.edge_lm[e];
        const double X[3] = {P.lms[3 * (size_t)l], P.lms[3 * (size_t)l + 1], P.lms[3 * (size_t)l + 2]};
        const double xr[3] = {R[0] * X[0] + R[1] * X[1] + R[2] * X[2] + t0,
                              R[3] * X[0] + R[4] * X[1] + R[5] * X[2] + t1,
                              R[6] * X[0] + R[7] * X[1] + R[8] * X[2] + t2};
        double G[6] = {0, 0, 0, 0, 0, 0}, h[3] = {0, 0, 0};
        const int o_end = P.edge_obs[e + 1];
        for (int o = P.edge_obs[e]; o < o_end; ++o) {
            const DevSensor& S = P.sensors[RIG ? P.obs_sensor[o] : 0];
            if (RIG && P.obs_kind && P.obs_kind[o] != OBS_PIXEL) {
                double r[4], J[12];
                if (!stereo_rows<true>(P, o, P.sensors, S, xr, R, r, J)) continue;
                const double sq = ((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]) + r[3] * r[3];
                const double w = robust_weight<double>((uint32_t)P.kernel_kind, P.kernel_scale, sq) *
                                 (P.obs_w ? P.obs_w[o] : 1.0);
                double Gs[6] = {0, 0, 0, 0, 0, 0}, hs[3] = {0, 0, 0};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double a = J[3 * k], b = J[3 * k + 1], c = J[3 * k + 2];
                    Gs[0] += a * a; Gs[1] += a * b; Gs[2] += a * c;
                    Gs[3] += b * b; Gs[4] += b * c; Gs[5] += c * c;
                    hs[0] += a * r[k]; hs[1] += b * r[k]; hs[2] += c * r[k];
                }
#pragma unroll
                for (int i = 0; i < 6; ++i) G[i] += w * Gs[i];
#pragma unroll
                for (int i = 0; i < 3; ++i) h[i] += w * hs[i];
                continue;
            }
            double xs[3], M[9];
            if (RIG && !S.identity) {
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    xs[i] = S.R[3 * i] * xr[0] + S.R[3 * i + 1] * xr[1] + S.R[3 * i + 2] * xr[2] + S.t[i];
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        M[3 * i + j] = S.R[3 * i] * R[j] + S.R[3 * i + 1] * R[3 + j] + S.R[3 * i + 2] * R[6 + j];
                }
            } else {
                xs[0] = xr[0]; xs[1] = xr[1]; xs[2] = xr[2];
#pragma unroll
                for (int i = 0; i < 9; ++i) M[i] = R[i];
            }
            double uv[2], Jp[6];
            if (!project<double>(S.lens, xs[0], xs[1], xs[2], uv, Jp)) continue;
            const double r0 = uv[0] - P.obs_uv[2 * (size_t)o], r1 = uv[1] - P.obs_uv[2 * (size_t)o + 1];
            // the robust weight times the observation's external weight
            // (normal_equations.rs:458, 627), 1 when there are none
            const double w = robust_weight<double>((uint32_t)P.kernel_kind, P.kernel_scale, r0 * r0 + r1 * r1) *
                             (P.obs_w ? P.obs_w[o] : 1.0);
            double J[6];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                J[c] = Jp[0] * M[c] + Jp[1] * M[3 + c] + Jp[2] * M[6 + c];
                J[3 + c] = Jp[3] * M[c] + Jp[4] * M[3 + c] + Jp[5] * M[6 + c];
            }
            G[0] += w * (J[0] * J[0] + J[3] * J[3]);
            G[1] += w * (J[0] * J[1] + J[3] * J[4]);
            G[2] += w * (J[0] * J[2] + J[3] * J[5]);
            G[3] += w * (J[1] * J[1] + J[4] * J[4]);
            G[4] += w * (J[1] * J[2] + J[4] * J[5]);
            G[5] += w * (J[2] * J[2] + J[5] * J[5]);
            h[0] += w * (J[0] * r0 + J[3] * r1);
            h[1] += w * (J[1] * r0 + J[4] * r1);
            h[2] += w * (J[2] * r0 + J[5] * r1);
        }
        // G (FP64, and FP32 for the mixed-precision operator) and h at the
        // edge's storage position
        const size_t pl = (size_t)P.perm_cl[e];
        double2* g64 = (double2*)(P.Gl64 + 6 * pl);
        g64[0] = make_double2(G[0], G[1]);
        g64[1] = make_double2(G[2], G[3]);
        g64[2] = make_double2(G[4], G[5]);
        P.hl64[3 * pl] = h[0];
        P.hl64[3 * pl + 1] = h[1];
        P.hl64[3 * pl + 2] = h[2];
        if (store32) {
            const int pp = P.g32_pos[pl];
            *(float4*)(P.Gl32 + g32_off4(pp)) = make_float4((float)G[0], (float)G[1], (float)G[2], (float)G[3]);
            *(float2*)(P.Gl32 + g32_off2(pp)) = make_float2((float)G[4], (float)G[5]);
        }
        add_schur_block<double>(G, X, 1.0, acc);
        acc[21] += h[0];
        acc[22] += h[1];
        acc[23] += h[2];
        acc[24] += X[1] * h[2] - X[2] * h[1];
        acc[25] += X[2] * h[0] - X[0] * h[2];
        acc[26] += X[0] * h[1] - X[1] * h[0];
    }
    block_sum<kBlock, double, 27>(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < 27; ++i) P.tile27[27 * (size_t)tile + i] = acc[i];
    }
}

// U_p, g_p: in-order sum of the pose's tile partials.
__global__ void k_camera_finalize(DevProblem P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.num_pose_slots) return;
    if (P.wstate && !P.wstate[P.pose_win[p]].relin) return;
    double acc[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) acc[i] = 0.0;
    for (int t = P.cam_tile_ptr[p]; t < P.cam_tile_ptr[p + 1]; ++t)
#pragma unroll
        for (int i = 0; i < 27; ++i) acc[i] += P.tile27[27 * (size_t)t + i];
#pragma unroll
    for (int i = 0; i < 21; ++i) P.U[21 * (size_t)p + i] = acc[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) P.g[6 * (size_t)p + i] = acc[21 + i];
}

// V_l, g_l: in-order sum over the landmark's l-order edges.
__global__ __launch_bounds__(256) void k_landmark_reduce(DevProblem P) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= P.num_lm_slots) return;
    if (P.wstate && !P.wstate[P.lm_win[l]].relin) return;
    double V[6] = {0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0}, Vm[6] = {0, 0, 0, 0, 0, 0};
    const bool mixed = P.Vm != nullptr;  // V of the FP32-stored G, the one the mixed operator pairs with them
    // The landmark's edges in batches of KB: every index of a batch is loaded,
    // then every block it points to, then the sums run in edge order (the
    // order of one edge at a time), so that a batch waits for two or three
    // memory round trips instead of each edge.
    constexpr int KB = 4;
    const int eb = P.lm_edge_ptr[l], ee = P.lm_edge_ptr[l + 1];
    for (int e0 = eb; e0 < ee; e0 += KB) {
        size_t st[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) st[u] = (size_t)P.lm_store[min(e0 + u, ee - 1)];
        double G[KB][6], h[KB][3];
        int pp[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) {
#pragma unroll
            for (int i = 0; i < 6; ++i) G[u][i] = P.Gl64[6 * st[u] + i];
#pragma unroll
            for (int i = 0; i < 3; ++i) h[u][i] = P.hl64[3 * st[u] + i];
            pp[u] = mixed ? P.g32_pos[st[u]] : 0;
        }
        float G32[KB][6];
        if (mixed) {
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const float4 a = *(const float4*)(P.Gl32 + g32_off4(pp[u]));
                const float2 b = *(const float2*)(P.Gl32 + g32_off2(pp[u]));
                G32[u][0] = a.x;
                G32[u][1] = a.y;
                G32[u][2] = a.z;
                G32[u][3] = a.w;
                G32[u][4] = b.x;
                G32[u][5] = b.y;
            }
        }
#pragma unroll
        for (int u = 0; u < KB; ++u)
            if (e0 + u < ee) {
#pragma unroll
                for (int i = 0; i < 6; ++i) V[i] += G[u][i];
#pragma unroll
                for (int i = 0; i < 3; ++i) g[i] += h[u][i];
                if (mixed)
#pragma unroll
                    for (int i = 0; i < 6; ++i) Vm[i] += (double)G32[u][i];
            }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) P.V[6 * (size_t)l + i] = V[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) P.gl[3 * (size_t)l + i] = g[i];
    if (mixed) {
#pragma unroll
        for (int i = 0; i < 6; ++i) P.Vm[6 * (size_t)l + i] = Vm[i];
    }
}

// (V + lambda I)^-1 of a symmetric 3x3 by cofactors; zero when the
// determinant is exactly zero or not finite.
__device__ __forceinline__ void damped_inverse3(const double* V, double lambda, double inv[6]) {
    double a = V[0], b = V[1], c = V[2], d = V[3], e = V[4], f = V[5];
    if (lambda > 0.0) {
        a += lambda;
        d += lambda;
        f += lambda;
    }
    // symmetric [[a b c][b d e][c e f]]
    const double c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
    const double det = a * c00 + b * c01 + c * c02;
    for (int i = 0; i < 6; ++i) inv[i] = 0.0;
    if (det != 0.0 && isfinite(det)) {
        inv[0] = c00 / det;
        inv[1] = c01 / det;
        inv[2] = c02 / det;
        inv[3] = (a * f - c * c) / det;
        inv[4] = (b * c - a * e) / det;
        inv[5] = (a * d - b * b) / det;
    }
}

// Damped landmark inverse (V + lambda I)^-1 by cofactors, zero when the
// determinant is exactly zero or the landmark is fixed (the landmark then
// drops out of the Schur complement and back-substitution, as try_inverse
// returning None does in normal_equations.rs:1432-1445). The mixed operator
// gets the inverse of the V summed from its FP32-stored G, so that its
// W V^-1 W^T is the exact Schur term of one consistent set of blocks:
// pairing FP32 G with the FP64 inverse would leave rounding errors of G
// along a poorly observed landmark's depth direction amplified by V's
// condition number.
__global__ void k_landmark_prepare(DevProblem P, double lambda) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= P.num_lm_slots) return;
    if (P.wstate) {
        const WindowState& w = P.wstate[P.lm_win[l]];
        if (!w.active) return;
        lambda = w.lambda;
    }
    const bool var = P.lm_var[l] >= 0;
    double inv[6] = {0, 0, 0, 0, 0, 0};
    if (var) damped_inverse3(P.V + 6 * (size_t)l, lambda, inv);
#pragma unroll
    for (int i = 0; i < 6; ++i) P.Vinv64[6 * (size_t)l + i] = inv[i];
    if (P.VinvM) {
        if (var) damped_inverse3(P.Vm + 6 * (size_t)l, lambda, inv);
#pragma unroll
        for (int i = 0; i < 6; ++i) P.VinvM[6 * (size_t)l + i] = inv[i];
    }
}

__global__ __launch_bounds__(kBlock) void k_pose_update(DevProblem P) {
    __shared__ double lds[kBlock / 64];
    const int p = blockIdx.x * kBlock + threadIdx.x;
    double step = 0.0;
    if (p < P.num_pose_slots && !(P.wstate && !P.wstate[P.pose_win[p]].active)) {
        const double* s = P.poses + 7 * (size_t)p;
        double* o = P.poses_trial + 7 * (size_t)p;
        if (P.pose_var[p] >= 0) {
            const double* xi = P.x + 6 * (size_t)p;
            double ss = 0.0;
            for (int i = 0; i < 6; ++i) ss += xi[i] * xi[i];
            step = sqrt(ss);
            Pose T{{s[0], s[1], s[2], s[3]}, {s[4], s[5], s[6]}};
            Pose n = pose_compose(T, se3_exp(xi));
            o[0] = n.q.w; o[1] = n.q.x; o[2] = n.q.y; o[3] = n.q.z;
            o[4] = n.t.x; o[5] = n.t.y; o[6] = n.t.z;
        } else {
            for (int i = 0; i < 7; ++i) o[i] = s[i];
        }
    }
    step = block_max<kBlock>(step, lds);
    if (threadIdx.x == 0)
        P.tile_scalar[4 * (size_t)(P.num_cam_tiles + P.num_lm_tiles + blockIdx.x) + 3] = step;
}

// Robust cost and non-projectable count (problem.rs:252-303, 497-583) with a
// compensated FP64 sum per tile.
template <bool RIG>
__global__ __launch_bounds__(kBlock) void k_cost(DevProblem P, int trial) {
    __shared__ double lds[(kBlock / 64) * 3];
    const int tile = xcd_tile(blockIdx.x, P.num_cam_tiles);
    if (tile >= P.num_cam_tiles) return;
    if (P.wstate && !P.wstate[P.cam_tile_win[tile]].active) return;
    const CamTile ct = P.cam_tiles[tile];
    const double* poses = trial ? P.poses_trial : P.poses;
    const double* lms = trial ? P.lms_trial : P.lms;
    const double* pq = poses + 7 * (size_t)ct.pose;
    // the trial camera of the joint camera solve
    const DevSensor* sensors = trial && P.sensors_trial ? P.sensors_trial : P.sensors;
    double R[9];
    qmatrix<double>(Quat{pq[0], pq[1], pq[2], pq[3]}, R);
    CSum cs{0.0, 0.0};
    double bad = 0.0;
    // the next edge's landmark and observation range load with the current
    // edge's work (clamped to the tile's last edge, no branch), so that an
    // edge waits for one memory round trip, not for its indices first
    int e = ct.begin + (int)threadIdx.x;
    const int elast = ct.end - 1;
    int l_next = P.edge_lm[min(e, elast)], ob_next = P.edge_obs[min(e, elast)],
        oe_next = P.edge_obs[min(e, elast) + 1];
    for (; e < ct.end; e += kBlock) {
        const int l = l_next, ob = ob_next, oe = oe_next;
        const int en = min(e + kBlock, elast);
        l_next = P.edge_lm[en];
        ob_next = P.edge_obs[en];
        oe_next = P.edge_obs[en + 1];
        const double X[3] = {lms[3 * (size_t)l], lms[3 * (size_t)l + 1], lms[3 * (size_t)l + 2]};
        const double xr[3] = {R[0] * X[0] + R[1] * X[1] + R[2] * X[2] + pq[4],
                              R[3] * X[0] + R[4] * X[1] + R[5] * X[2] + pq[5],
                              R[6] * X[0] + R[7] * X[1] + R[8] * X[2] + pq[6]};
        for (int o = ob; o < oe; ++o) {
            const DevSensor& S = sensors[RIG ? P.obs_sensor[o] : 0];
            double sq;
            if (RIG && P.obs_kind && P.obs_kind[o] != OBS_PIXEL) {
                double r[4];
                if (!stereo_rows<false>(P, o, sensors, S, xr, R, r, (double*)nullptr)) {
                    bad += 1.0;
                    if (P.obs_s) P.obs_s[o] = __builtin_nan("");
                    continue;
                }
                sq = ((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]) + r[3] * r[3];
            } else {
                double xs[3] = {xr[0], xr[1], xr[2]};
                if (RIG && !S.identity) {
                    for (int i = 0; i < 3; ++i)
                        xs[i] = S.R[3 * i] * xr[0] + S.R[3 * i + 1] * xr[1] + S.R[3 * i + 2] * xr[2] + S.t[i];
                }
                double uv[2];
                if (!project<double>(S.lens, xs[0], xs[1], xs[2], uv, (double*)nullptr)) {
                    bad += 1.0;
                    if (P.obs_s) P.obs_s[o] = __builtin_nan("");
                    continue;
                }
                const double r0 = uv[0] - P.obs_uv[2 * (size_t)o], r1 = uv[1] - P.obs_uv[2 * (size_t)o + 1];
                sq = r0 * r0 + r1 * r1;
            }
            if (P.obs_s) P.obs_s[o] = sq;
            const double c = robust_cost<double>((uint32_t)P.kernel_kind, P.kernel_scale, sq);
            // w rho(s) with the observation's external weight (problem.rs:521-523, 583)
            csum_add(cs, P.obs_w ? P.obs_w[o] * c : c);
        }
    }
    // compensated tree: wave shuffles of (s, c) pairs, then LDS in fixed order
    for (int off = 32; off > 0; off >>= 1) {
        CSum o{__shfl_down(cs.s, off, 64), __shfl_down(cs.c, off, 64)};
        bad += __shfl_down(bad, off, 64);
        if ((threadIdx.x & 63) + off < 64) cs = csum_merge(cs, o);
    }
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
        lds[3 * w] = cs.s;
        lds[3 * w + 1] = cs.c;
        lds[3 * w + 2] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        CSum a{lds[0], lds[1]};
        double b = lds[2];
        for (int w = 1; w < kBlock / 64; ++w) {
            a = csum_merge(a, CSum{lds[3 * w], lds[3 * w + 1]});
            b += lds[3 * w + 2];
        }
        double* out = P.tile_scalar + 4 * (size_t)tile;
        out[0] = a.s;
        out[1] = a.c;
        out[2] = b;
    }
}

__global__ __launch_bounds__(kSmallBlock) void k_scalars_finalize(DevProblem P, int num_pose_blocks) {
    __shared__ double lds[(kSmallBlock / 64) * 4];
    // cost partials in tile order: thread t < parts merges a contiguous run
    // of per tiles, and thread 0 folds the parts in order. Up to
    // kSmallBlock tiles a part is one tile (the fold is then the tile-order
    // fold that k_batch_scalars does for a window); beyond, 128 parts keep
    // the serial fold short.
    const int n = P.num_cam_tiles;
    const int parts = n <= kSmallBlock ? n : 128;
    const int per = parts > 0 ? (n + parts - 1) / parts : 0;
    CSum cs{0.0, 0.0};
    double bad = 0.0, mp = 0.0, ml = 0.0;
    // (unrolled so that each thread's loads are in flight together; the
    // merge order is the loop order)
#pragma unroll 8
    for (int i = threadIdx.x * per; i < min(n, (int)(threadIdx.x + 1) * per); ++i) {
        const double* t = P.tile_scalar + 4 * (size_t)i;
        cs = csum_merge(cs, CSum{t[0], t[1]});
        bad += t[2];
    }
#pragma unroll 8
    for (int i = threadIdx.x; i < P.num_lm_tiles; i += kSmallBlock)
        ml = fmax(ml, P.tile_scalar[4 * (size_t)(n + i) + 3]);
#pragma unroll 8
    for (int i = threadIdx.x; i < num_pose_blocks; i += kSmallBlock)
        mp = fmax(mp, P.tile_scalar[4 * (size_t)(n + P.num_lm_tiles + i) + 3]);
    // ordered merge across threads through LDS (thread order = tile order)
    __shared__ double ss[kSmallBlock], sc[kSmallBlock];
    ss[threadIdx.x] = cs.s;
    sc[threadIdx.x] = cs.c;
    double v[1] = {bad};
    block_sum<kSmallBlock, double, 1>(v, lds);
    const double m1 = block_max<kSmallBlock>(mp, lds + 16);
    const double m2 = block_max<kSmallBlock>(ml, lds + 32);
    __syncthreads();
    if (threadIdx.x == 0) {
        CSum a{ss[0], sc[0]};
        for (int i = 1; i < parts; ++i) a = csum_merge(a, CSum{ss[i], sc[i]});
        StepScalars* o = P.scalars;
        o->cost_sum = a.s;
        o->cost_comp = a.c;
        o->nonprojectable = (uint64_t)v[0];
        o->max_pose_step = m1;
        o->max_landmark_step = m2;
    }
}

// ---------------------------------------------------------------------------
// Dense reduced camera system (LinearSolver::Dense): explicit S and a
// one-work-group Cholesky with an LU fallback (pose_graph.rs:2259-2270).
// ---------------------------------------------------------------------------
// One 6x6 block (p, q) of S = U + lambda I - W V^-1 W^T, written at the
// variable indices local to its system (and mirrored when p != q).
// Lanes that assemble one 6x6 block of S together.
constexpr int kBlockLanes = 8;

// One 6x6 block (p, q) of S = U + lambda I - W V^-1 W^T, written at the
// variable indices local to its system (and mirrored when p != q), by a
// group of kBlockLanes lanes: each sums a strided share of the block's terms,
// a butterfly within the group combines them in a fixed order.
__device__ void dense_assemble_group(const DevProblem& P, const DensePairs& D, int b, double* S, int n, int var_begin,
                                     double lambda, int g) {
    const int sp = D.block_pq[2 * b], sq = D.block_pq[2 * b + 1];
    const int vp = P.pose_var[sp] - var_begin, vq = P.pose_var[sq] - var_begin;
    double acc[36];
    schur_block_terms(P, D, b, g, kBlockLanes, acc);
#pragma unroll
    for (int off = 1; off < kBlockLanes; off <<= 1)
#pragma unroll
        for (int i = 0; i < 36; ++i) acc[i] += __shfl_xor(acc[i], off, kBlockLanes);
    // every lane holds the sum; lane g writes entries g, g + 8, ...
    const double* U = P.U + 21 * (size_t)sp;
#pragma unroll
    for (int idx0 = 0; idx0 < 36; idx0 += kBlockLanes) {
        const int idx = idx0 + g;
        if (idx < 36) {
            const int i = idx / 6, j = idx % 6;
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < 36; ++t)
                if (t == idx) v = acc[t];
            if (sp == sq) {
                const double u = U[i <= j ? up6(i, j) : up6(j, i)] + (i == j && lambda > 0.0 ? lambda : 0.0);
                S[(size_t)(6 * vp + i) * n + 6 * vq + j] = u - v;
            } else {
                S[(size_t)(6 * vp + i) * n + 6 * vq + j] = -v;
                S[(size_t)(6 * vq + j) * n + 6 * vp + i] = -v;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_dense_assemble(DevProblem P, DensePairs D, double* S, int n,
                                                           double lambda) {
    const int b = blockIdx.x * (kBlock / kBlockLanes) + (int)threadIdx.x / kBlockLanes;
    if (b >= D.num_blocks) return;
    dense_assemble_group(P, D, b, S, n, 0, lambda, threadIdx.x % kBlockLanes);
}

typedef double v4d __attribute__((ext_vector_type(4)));

// S(r, c) of an n x n row-major matrix, extended by the identity to whole
// 16 x 16 blocks.
__device__ __forceinline__ double padded_at(const double* S, int n, int r, int c) {
    return (r < n && c < n) ? S[(size_t)r * n + c] : (r == c ? 1.0 : 0.0);
}

// Blocked right-looking Cholesky S = L L^T in place (lower triangle,
// row-major), 16 x 16 blocks: the diagonal block is factored in LDS, the
// panel below it by one forward substitution per row, and the trailing
// blocks are updated as C -= L_i L_j^T on FP64 MFMA (v_mfma_f64_16x16x4:
// operand lane l holds A[l%16][l/16] and B[l/16][l%16]; result register q of
// lane l is C[4q + l/16][l%16]). One work-group of NT >= 256 threads.
// *fail = 1 on a pivot that is not positive and finite.
template <int NT>
__device__ void chol_blocked(double* S, int n, double* dblk, int* fail) {
    constexpr int NW = NT / 64;
    const int nb = (n + 15) >> 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti = threadIdx.x >> 4, tc = threadIdx.x & 15;  // element of the diagonal block (threads < 256)
    for (int k = 0; k < nb; ++k) {
        const int k0 = 16 * k;
        if (threadIdx.x < 256) dblk[ti * 17 + tc] = padded_at(S, n, k0 + ti, k0 + tc);
        __syncthreads();
        for (int j = 0; j < 16; ++j) {
            const double d = dblk[j * 17 + j];
            if (!(d > 0.0) || !isfinite(d)) {
                if (threadIdx.x == 0) *fail = 1;
                break;
            }
            const double ljj = sqrt(d);
            double v = 0.0;
            const bool col = threadIdx.x < 256 && tc == j && ti >= j;
            if (col) v = ti == j ? ljj : dblk[ti * 17 + j] / ljj;
            __syncthreads();
            if (col) dblk[ti * 17 + j] = v;
            __syncthreads();
            if (threadIdx.x < 256 && tc > j && tc <= ti) dblk[ti * 17 + tc] -= dblk[ti * 17 + j] * dblk[tc * 17 + j];
            __syncthreads();
        }
        __syncthreads();
        if (*fail) return;
        if (threadIdx.x < 256 && tc <= ti && k0 + ti < n) S[(size_t)(k0 + ti) * n + k0 + tc] = dblk[ti * 17 + tc];
        // panel: L[R][k0..k0+15] = S[R][k0..] L_kk^-T
        for (int R = k0 + 16 + (int)threadIdx.x; R < n; R += NT) {
            double x[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) x[c] = k0 + c < n ? S[(size_t)R * n + k0 + c] : 0.0;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                double t = x[c];
#pragma unroll
                for (int jj = 0; jj < c; ++jj) t -= x[jj] * dblk[c * 17 + jj];
                x[c] = t / dblk[c * 17 + c];
            }
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (k0 + c < n) S[(size_t)R * n + k0 + c] = x[c];
        }
        __syncthreads();
        // trailing blocks (i, j), k < j <= i < nb
        const int m = nb - k - 1;
        for (int bb = wave; bb < m * (m + 1) / 2; bb += NW) {
            int i = (int)((sqrt(8.0 * bb + 1.0) - 1.0) * 0.5);
            while ((i + 1) * (i + 2) / 2 <= bb) ++i;
            while (i * (i + 1) / 2 > bb) --i;
            const int j = bb - i * (i + 1) / 2;
            const int bi = 16 * (k + 1 + i), bj = 16 * (k + 1 + j);
            v4d acc;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = padded_at(S, n, bi + 4 * q + (lane >> 4), bj + (lane & 15));
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const double a = -padded_at(S, n, bi + (lane & 15), k0 + 4 * kk + (lane >> 4));
                const double bv = padded_at(S, n, bj + (lane & 15), k0 + 4 * kk + (lane >> 4));
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = bi + 4 * q + (lane >> 4), c = bj + (lane & 15);
                if (r < n && c < n) S[(size_t)r * n + c] = acc[q];
            }
        }
        __syncthreads();
    }
}

// Lane l's value of v in every lane; l must be the same in every lane (a
// scalar read of one lane, where a shuffle would go through LDS).
__device__ __forceinline__ double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// y = L^-1 y then y = L^-T y for the factor of chol_blocked, by 16-row
// blocks: the diagonal block, staged in LDS (dblk, 16 x 17), by wave 0 (one
// row per lane, shuffles), the rest as one short dot product per row.
template <int NT>
__device__ void chol_blocked_solve(const double* S, int n, double* y, double* dblk) {
    static_assert(NT >= 256, "one thread per entry of the diagonal block");
    const int nb = (n + 15) >> 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti = threadIdx.x >> 4, tc = threadIdx.x & 15;
    auto stage = [&](int k0) {
        if (threadIdx.x < 256)
            dblk[ti * 17 + tc] = (k0 + ti < n && tc <= ti) ? S[(size_t)(k0 + ti) * n + k0 + tc] : 0.0;
        __syncthreads();
    };
    for (int k = 0; k < nb; ++k) {  // forward: L y = b
        const int k0 = 16 * k;
        stage(k0);
        if (wave == 0) {
            const int r = k0 + lane;
            double v = lane < 16 && r < n ? y[r] : 0.0;
            for (int j = 0; j < 16 && k0 + j < n; ++j) {
                if (lane == j) v = v / dblk[j * 17 + j];
                const double yj = lane_value(v, j);
                if (lane > j && lane < 16 && r < n) v -= dblk[lane * 17 + j] * yj;
            }
            if (lane < 16 && r < n) y[r] = v;
        }
        __syncthreads();
        for (int R = k0 + 16 + (int)threadIdx.x; R < n; R += NT) {
            double t = y[R];
            for (int c = 0; c < 16; ++c) t -= S[(size_t)R * n + k0 + c] * y[k0 + c];
            y[R] = t;
        }
        __syncthreads();
    }
    for (int k = nb - 1; k >= 0; --k) {  // backward: L^T x = y
        const int k0 = 16 * k;
        stage(k0);
        if (wave == 0) {
            const int r = k0 + lane;
            double v = lane < 16 && r < n ? y[r] : 0.0;
            for (int j = 15; j >= 0; --j) {
                if (k0 + j >= n) continue;
                if (lane == j) v = v / dblk[j * 17 + j];
                const double xj = lane_value(v, j);
                if (lane < j && r < n) v -= dblk[j * 17 + lane] * xj;
            }
            if (lane < 16 && r < n) y[r] = v;
        }
        __syncthreads();
        for (int R = (int)threadIdx.x; R < k0; R += NT) {
            double t = y[R];
            for (int c = 0; c < 16 && k0 + c < n; ++c) t -= S[(size_t)(k0 + c) * n + R] * y[k0 + c];
            y[R] = t;
        }
        __syncthreads();
    }
}

// Where a dense solve's right-hand side comes from and its solution goes:
// unknown i < nv is row var_slot[i / 6] * 6 + i % 6 of b / x (row i when
// var_slot is null), unknown i >= nv is row i - nv of tail_b / tail_x (the
// camera block of the joint solve).
struct DenseRows {
    const double* b;
    double* x;
    const int32_t* var_slot;
    int nv;
    const double* tail_b;
    double* tail_x;
    __device__ __forceinline__ size_t row(int i) const {
        return var_slot ? 6 * (size_t)var_slot[i / 6] + (i % 6) : (size_t)i;
    }
    __device__ __forceinline__ double rhs(int i) const { return i < nv ? b[row(i)] : tail_b[i - nv]; }
    __device__ __forceinline__ void put(int i, double v) const {
        if (i < nv) x[row(i)] = v;
        else tail_x[i - nv] = v;
    }
};
__host__ __device__ __forceinline__ DenseRows pose_rows(const DevProblem& P, const int32_t* var_slot, int n) {
    // the last joint_k unknowns are the camera block of the joint solve
    return DenseRows{P.b, P.x, var_slot, n - P.joint_k, P.joint_rhs, P.joint_delta};
}

// LU with partial pivoting of the n x n row-major A2 (overwritten), solving
// A2 y = y in place; *fail_sh = 2 when A2 is singular (a zero pivot column
// survives to the back substitution). One work-group of NT threads.
template <int NT>
__device__ void lu_fallback_body(double* A2, double* y, int n, int* fail_sh, double* piv_val, int* piv_idx) {
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bi = k;
        for (int i = k + threadIdx.x; i < n; i += NT) {
            const double v = fabs(A2[(size_t)i * n + k]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        piv_val[threadIdx.x] = best;
        piv_idx[threadIdx.x] = bi;
        __syncthreads();
        if (threadIdx.x == 0) {
            double bv = piv_val[0];
            int bidx = piv_idx[0];
            for (int t = 1; t < NT; ++t)
                if (piv_val[t] > bv || (piv_val[t] == bv && piv_idx[t] < bidx)) {
                    bv = piv_val[t];
                    bidx = piv_idx[t];
                }
            piv_idx[0] = bidx;
            piv_val[0] = bv;
        }
        __syncthreads();
        const int pr = piv_idx[0];
        const double pv = A2[(size_t)pr * n + k];
        __syncthreads();
        if (pv == 0.0) continue;
        if (pr != k) {
            for (int c = threadIdx.x; c < n; c += NT) {
                const double t = A2[(size_t)k * n + c];
                A2[(size_t)k * n + c] = A2[(size_t)pr * n + c];
                A2[(size_t)pr * n + c] = t;
            }
            if (threadIdx.x == 0) {
                const double t = y[k];
                y[k] = y[pr];
                y[pr] = t;
            }
        }
        __syncthreads();
        const double inv = 1.0 / A2[(size_t)k * n + k];
        for (int i = k + 1 + threadIdx.x; i < n; i += NT) A2[(size_t)i * n + k] *= inv;
        __syncthreads();
        const int rows = n - k - 1;
        for (int idx = threadIdx.x; idx < rows * rows; idx += NT) {
            const int i = k + 1 + idx / rows, c = k + 1 + idx % rows;
            A2[(size_t)i * n + c] -= A2[(size_t)i * n + k] * A2[(size_t)k * n + c];
        }
        if (threadIdx.x == 0) {
            const double yk = y[k];
            for (int i = k + 1; i < n; ++i) y[i] -= A2[(size_t)i * n + k] * yk;
        }
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {
        const double dj = A2[(size_t)j * n + j];
        if (dj == 0.0) {
            if (threadIdx.x == 0) *fail_sh = 2;
            break;
        }
        if (threadIdx.x == 0) y[j] = y[j] / dj;
        __syncthreads();
        const double xj = y[j];
        for (int i = threadIdx.x; i < j; i += NT) y[i] -= A2[(size_t)i * n + j] * xj;
        __syncthreads();
    }
}

// Cholesky (blocked, MFMA trailing updates), then forward/back
// substitution; on a non-positive pivot an LU with partial pivoting on the
// saved copy (allow_lu, solve_normal_equations), or singular (the Sparse
// solver's block Cholesky has no fallback). Solves S x = b for the n
// unknowns of `io`. One work-group of NT threads.
template <int NT>
__device__ void dense_solve_body(const DenseRows& io, double* S, double* work, int n, int* status, int* fail_sh,
                                 double* piv_val, int* piv_idx, double* dblk, bool copied = false,
                                 bool allow_lu = true) {
    double* A2 = work;                 // n*n copy for the LU fallback
    double* y = work + (size_t)n * n;  // n
    // the saved copy for the LU fallback, unless the whole device made it
    // before this work-group started
    if (!copied && allow_lu)
        for (size_t k = threadIdx.x; k < (size_t)n * n; k += NT) A2[k] = S[k];
    for (int i = threadIdx.x; i < n; i += NT) y[i] = io.rhs(i);
    if (threadIdx.x == 0) *fail_sh = 0;
    __syncthreads();
    chol_blocked<NT>(S, n, dblk, fail_sh);
    __syncthreads();
    if (!*fail_sh) {
        chol_blocked_solve<NT>(S, n, y, dblk);
    } else if (allow_lu) {
        // LU with partial pivoting on A2; y holds b again.
        for (int i = threadIdx.x; i < n; i += NT) y[i] = io.rhs(i);
        __syncthreads();
        lu_fallback_body<NT>(A2, y, n, fail_sh, piv_val, piv_idx);
    } else {
        if (threadIdx.x == 0) *fail_sh = 2;
    }
    __syncthreads();
    const int f = *fail_sh;
    for (int i = threadIdx.x; i < n; i += NT) io.put(i, (f != 2) ? y[i] : 0.0);
    if (threadIdx.x == 0) *status = (f == 2) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// The same solve across the whole device, for systems larger than one
// work-group factors quickly: a right-looking Cholesky over 64-column
// panels, two launches per panel. k_chol_panel: every work-group factors the
// 64 x 64 diagonal block in LDS (the same arithmetic in each, so every one
// sees the same pivots) and solves its 256 rows of the panel below,
// L_R = A_R L_kk^-T; work-group 0 writes the diagonal block back.
// k_chol_update: the trailing lower triangle A_ij -= L_i L_j^T in 16 x 16
// tiles, one wave each, on FP64 MFMA. flags[1] records a failed pivot (set
// by work-group 0; later launches exit at once); k_dense_finish then solves
// with the factor, or runs the LU fallback on the saved copy, in one
// work-group.
// ---------------------------------------------------------------------------
constexpr int kPanel = 64;


__global__ void k_dense_flags_clear(int* flags) {
    if (threadIdx.x < 2) flags[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_chol_panel(double* S, int n, int k0, int* flags) {
    if (__hip_atomic_load(&flags[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    __shared__ double L[kPanel][kPanel + 1];
    const int w = min(kPanel, n - k0);
    for (int t = threadIdx.x; t < kPanel * kPanel; t += 256) {
        const int r = t / kPanel, c = t % kPanel;
        L[r][c] = (r < w && c < w) ? S[(size_t)(k0 + r) * n + k0 + c] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    // the diagonal block's factor by wave 0, lane i holding row i in registers:
    // column j's pivot and entries come from lane j and lane c by lane reads.
    // Then lane c forms column c of L^-1 the same way, into LDS (Li).
    __shared__ int failed;
    __shared__ double Li[kPanel][kPanel + 1];
    if (threadIdx.x < kPanel) {
        const int i = threadIdx.x;
        double a[kPanel];
#pragma unroll
        for (int c = 0; c < kPanel; ++c) a[c] = L[i][c];
        int bad = 0;
#pragma unroll
        for (int j = 0; j < kPanel; ++j) {
            const double d = lane_value(a[j], j);
            if (!(d > 0.0) || !isfinite(d)) {  // the same in every lane and work-group
                bad = 1;
                break;
            }
            const double ljj = sqrt(d);
            if (i == j) a[j] = ljj;
            else if (i > j) a[j] = a[j] / ljj;
            const double lij = a[j];
#pragma unroll
            for (int c = j + 1; c < kPanel; ++c) {
                const double lcj = lane_value(a[j], c);
                if (c <= i) a[c] -= lij * lcj;
            }
        }
        if (i == 0) failed = bad;
        if (!bad) {
            if (blockIdx.x == 0 && i < w)
#pragma unroll
                for (int c = 0; c < kPanel; ++c)
                    if (c <= i) S[(size_t)(k0 + i) * n + k0 + c] = a[c];
            if (k0 + kPanel < n) {
                // column c = i of L^-1: y_r = (e_c - sum_{k<r} L[r][k] y_k) / L[r][r]
                const int c = i;
                double yv[kPanel];
#pragma unroll
                for (int r = 0; r < kPanel; ++r) {
                    double t0 = 0.0, t1 = 0.0;
#pragma unroll
                    for (int k = 0; k < r; ++k) {
                        const double lrk = lane_value(a[k], r);
                        if (k & 1) t1 += lrk * yv[k];
                        else t0 += lrk * yv[k];
                    }
                    const double lrr = lane_value(a[r], r);
                    yv[r] = r < c ? 0.0 : ((r == c ? 1.0 : 0.0) - (t0 + t1)) / lrr;
                }
#pragma unroll
                for (int r = 0; r < kPanel; ++r) Li[r][c] = yv[r];
            }
        }
    }
    __syncthreads();
    if (failed) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            __hip_atomic_store(&flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    if (k0 + kPanel >= n) return;  // no panel rows below
    // panel rows: X = A_R L_kk^-T on FP64 MFMA, 16 x 16 output tiles; each
    // wave holds the B operands (L^-1 transposed) of all 4 column tiles and
    // takes the rows of its 64-row slice of this work-group's 256 rows.
    // Operand lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; result
    // register q holds C[4q + l / 16][l % 16].
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double bop[4][kPanel / 4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int kk = 0; kk < kPanel / 4; ++kk) bop[ct][kk] = Li[16 * ct + (lane & 15)][4 * kk + (lane >> 4)];
    const int rbase = k0 + kPanel + (int)blockIdx.x * 256 + 64 * wave;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        const int r0 = rbase + 16 * rt;
        if (r0 >= n) break;
        const int ra = r0 + (lane & 15);
        double aop[kPanel / 4];
#pragma unroll
        for (int kk = 0; kk < kPanel / 4; ++kk)
            aop[kk] = ra < n ? S[(size_t)ra * n + k0 + 4 * kk + (lane >> 4)] : 0.0;
        v4d acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            acc[ct] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < kPanel / 4; ++kk)
                acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[kk], bop[ct][kk], acc[ct], 0, 0, 0);
        }
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r0 + 4 * q + (lane >> 4);
                if (r < n) S[(size_t)r * n + k0 + 16 * ct + (lane & 15)] = acc[ct][q];
            }
    }
}

__global__ __launch_bounds__(256) void k_chol_update(double* S, int n, int k0, const int* flags) {
    if (__hip_atomic_load(&flags[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const int k1 = k0 + kPanel;
    const int m = (n - k1 + 15) >> 4;
    const int lane = threadIdx.x & 63;
    const long long bb = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m <= 0 || bb >= (long long)m * (m + 1) / 2) return;
    long long i = (long long)((sqrt(8.0 * (double)bb + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= bb) ++i;
    while (i * (i + 1) / 2 > bb) --i;
    const int j = (int)(bb - i * (i + 1) / 2);
    const int bi = k1 + 16 * (int)i, bj = k1 + 16 * j;
    v4d acc;
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = padded_at(S, n, bi + 4 * q + (lane >> 4), bj + (lane & 15));
#pragma unroll
    for (int kk = 0; kk < kPanel / 4; ++kk) {
        const double a = -padded_at(S, n, bi + (lane & 15), k0 + 4 * kk + (lane >> 4));
        const double bv = padded_at(S, n, bj + (lane & 15), k0 + 4 * kk + (lane >> 4));
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = bi + 4 * q + (lane >> 4), c = bj + (lane & 15);
        if (r < n && c < n) S[(size_t)r * n + c] = acc[q];
    }
}

// After the device-wide factor: forward / back substitution with it, or the
// LU fallback on the saved copy in work, or singular without a fallback.
// flags[0] = 1 when singular.
__global__ __launch_bounds__(kSmallBlock) void k_dense_finish(DenseRows io, double* S, double* work, int n,
                                                              int* flags, int allow_lu) {
    __shared__ int fail;
    __shared__ double piv_val[kSmallBlock];
    __shared__ int piv_idx[kSmallBlock];
    __shared__ double dblk[16 * 17];
    double* y = work + (size_t)n * n;
    for (int i = threadIdx.x; i < n; i += kSmallBlock) y[i] = io.rhs(i);
    if (threadIdx.x == 0) fail = __hip_atomic_load(&flags[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (!fail) chol_blocked_solve<kSmallBlock>(S, n, y, dblk);
    else if (allow_lu) lu_fallback_body<kSmallBlock>(work, y, n, &fail, piv_val, piv_idx);
    else if (threadIdx.x == 0) fail = 2;
    __syncthreads();
    const int f = fail;
    for (int i = threadIdx.x; i < n; i += kSmallBlock) io.put(i, (f != 2) ? y[i] : 0.0);
    if (threadIdx.x == 0) flags[0] = (f == 2) ? 1 : 0;
}

__global__ __launch_bounds__(kSmallBlock) void k_dense_solve(DevProblem P, DensePairs D, double* S, double* work,
                                                             int n, int* status, int allow_lu) {
    __shared__ int fail;
    __shared__ double piv_val[kSmallBlock];
    __shared__ int piv_idx[kSmallBlock];
    __shared__ double dblk[16 * 17];
    dense_solve_body<kSmallBlock>(pose_rows(P, D.var_slot, n), S, work, n, status, &fail, piv_val, piv_idx, dblk,
                                  true, allow_lu != 0);
}

// S x = b for a caller's row-major S (overwritten); work holds n * n + n.
__global__ __launch_bounds__(kSmallBlock) void k_dense_solve_plain(double* S, double* work, const double* b,
                                                                   double* x, int n, int* status) {
    __shared__ int fail;
    __shared__ double piv_val[kSmallBlock];
    __shared__ int piv_idx[kSmallBlock];
    __shared__ double dblk[16 * 17];
    const DenseRows io{b, x, nullptr, n, nullptr, nullptr};
    dense_solve_body<kSmallBlock>(io, S, work, n, status, &fail, piv_val, piv_idx, dblk, true);
}

// ---------------------------------------------------------------------------
// Batches of independent windows: the LM loop of lm.rs:162-1017 run for
// every window at once, with per-window lambda, cost and accept/reject kept
// on the device (no host round trip per iteration).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_batch_dense_assemble(DevProblem P, DensePairs D, BatchDev B) {
    const int b = blockIdx.x * (kBlock / kBlockLanes) + (int)threadIdx.x / kBlockLanes;
    if (b >= D.num_blocks) return;
    const int w = P.pose_win[D.block_pq[2 * b]];
    const WindowState& st = B.state[w];
    if (!st.active) return;
    const WindowLayout L = B.layout[w];
    dense_assemble_group(P, D, b, B.S + L.s_offset, 6 * (L.var_end - L.var_begin), L.var_begin, st.lambda,
                         threadIdx.x % kBlockLanes);
}

__global__ __launch_bounds__(kBlock) void k_batch_dense_zero(DevProblem P, BatchDev B) {
    const int w = blockIdx.y;
    if (!B.state[w].active) return;
    const WindowLayout L = B.layout[w];
    const size_t n = 6 * (size_t)(L.var_end - L.var_begin);
    double* S = B.S + L.s_offset;
    for (size_t k = blockIdx.x * (size_t)kBlock + threadIdx.x; k < n * n; k += (size_t)gridDim.x * kBlock) S[k] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_batch_dense_solve(DevProblem P, DensePairs D, BatchDev B) {
    __shared__ int fail;
    __shared__ double piv_val[kBlock];
    __shared__ int piv_idx[kBlock];
    const int w = blockIdx.x;
    WindowState& st = B.state[w];
    if (!st.active) return;
    const WindowLayout L = B.layout[w];
    const int n = 6 * (L.var_end - L.var_begin);
    if (n == 0) {
        if (threadIdx.x == 0) st.solve_status = 0;
        return;
    }
    __shared__ double dblk[16 * 17];
    dense_solve_body<kBlock>(pose_rows(P, D.var_slot + L.var_begin, n), B.S + L.s_offset, B.work + B.work_offset[w],
                             n, &st.solve_status, &fail, piv_val, piv_idx, dblk);
}

// Per window: compensated cost and non-projectable count over its camera
// tiles (tile order), and the largest pose / landmark step.
__global__ __launch_bounds__(kBlock) void k_batch_scalars(DevProblem P, BatchDev B) {
    __shared__ double ss[kBlock], sc[kBlock];
    __shared__ double lds[(kBlock / 64) * 3];
    const int w = blockIdx.x;
    WindowState& st = B.state[w];
    if (!st.active) return;
    const WindowLayout L = B.layout[w];
    const int nt = L.tile_end - L.tile_begin;
    const int per = (nt + kBlock - 1) / kBlock;
    CSum cs{0.0, 0.0};
    double v[3] = {0.0, 0.0, 0.0};  // non-projectable, max pose step, max landmark step
    for (int i = L.tile_begin + threadIdx.x * per; i < min(L.tile_end, L.tile_begin + (int)(threadIdx.x + 1) * per); ++i) {
        const double* t = P.tile_scalar + 4 * (size_t)i;
        cs = csum_merge(cs, CSum{t[0], t[1]});
        v[0] += t[2];
    }
    for (int p = L.pose_begin + threadIdx.x; p < L.pose_end; p += kBlock) {
        if (P.pose_var[p] < 0) continue;
        const double* xi = P.x + 6 * (size_t)p;
        double s2 = 0.0;
        for (int i = 0; i < 6; ++i) s2 += xi[i] * xi[i];
        v[1] = fmax(v[1], sqrt(s2));
    }
    for (int l = L.lm_begin + threadIdx.x; l < L.lm_end; l += kBlock) {
        const double* d = P.delta_l + 3 * (size_t)l;
        v[2] = fmax(v[2], sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
    }
    ss[threadIdx.x] = cs.s;
    sc[threadIdx.x] = cs.c;
    double cnt[1] = {v[0]};
    block_sum<kBlock, double, 1>(cnt, lds);
    const double mp = block_max<kBlock>(v[1], lds);
    const double ml = block_max<kBlock>(v[2], lds);
    __syncthreads();
    if (threadIdx.x == 0) {
        CSum a{ss[0], sc[0]};
        for (int i = 1; i < kBlock; ++i) a = csum_merge(a, CSum{ss[i], sc[i]});
        st.cost_after = a.s + a.c;
        st.nonproj_after = (uint64_t)cnt[0];
        st.max_pose_step = mp;
        st.max_landmark_step = ml;
    }
}

// Initial LM state (lm.rs:323-333) from the cost of the current state.
__global__ void k_batch_init(BatchDev B, double initial_lambda) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= B.num_windows) return;
    WindowState& st = B.state[w];
    if (!st.active) return;
    st.current_cost = st.cost_after;
    st.initial_cost = st.cost_after;
    st.current_nonproj = st.nonproj_after;
    st.lambda = initial_lambda;
    st.relin = 1;
    st.accept = 0;
    st.converged = 0;
    st.iterations = 0;
}

struct ControlOptions {
    int32_t has_initial_lambda, has_relative_cost_tolerance;
    double increase, decrease, max_lambda, min_lambda, step_tolerance, cost_tolerance, relative_cost_tolerance;
};

// Accept / reject and lambda update of one LM iteration per window
// (lm.rs:610-662, 804-1008).
__global__ void k_batch_control(BatchDev B, ControlOptions o, int iteration) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= B.num_windows) return;
    WindowState& st = B.state[w];
    st.accept = 0;
    st.relin = 0;
    if (!st.active) return;
    vba_iteration_stats rec;
    rec.iteration = (uint32_t)iteration;
    rec.step_accepted = 0;
    rec.cost_before = st.current_cost;
    rec.pcg_iterations = -1;
    rec.pcg_residual_norm = 0.0;
    rec.pcg_target = 0.0;
    bool stop = false;
    if (st.solve_status != 0) {  // singular reduced system: a rejected step
        st.lambda = fmin(st.lambda * o.increase, o.max_lambda);
        rec.cost_after = st.current_cost;
        rec.max_pose_step = 0.0;
        rec.max_landmark_step = 0.0;
        rec.lambda = st.lambda;
        rec.nonprojectable_after = (uint32_t)st.current_nonproj;
        stop = st.lambda >= o.max_lambda;
    } else {
        const double cost_before = st.current_cost;
        const bool cost_ok = o.has_initial_lambda ? (st.cost_after < cost_before) : true;
        const bool accepted = cost_ok && st.nonproj_after <= st.current_nonproj;
        rec.cost_after = st.cost_after;
        rec.max_pose_step = st.max_pose_step;
        rec.max_landmark_step = st.max_landmark_step;
        rec.nonprojectable_after = (uint32_t)st.nonproj_after;
        if (!accepted) {
            st.lambda = fmin(st.lambda * o.increase, o.max_lambda);
            rec.lambda = st.lambda;
            stop = !o.has_initial_lambda || st.lambda >= o.max_lambda;
        } else {
            rec.lambda = st.lambda;
            rec.step_accepted = 1;
            st.accept = 1;
            st.relin = 1;
            st.current_cost = st.cost_after;
            st.current_nonproj = st.nonproj_after;
            if (o.has_initial_lambda) st.lambda = fmax(st.lambda * o.decrease, o.min_lambda);
            bool conv = (st.max_pose_step < o.step_tolerance && st.max_landmark_step < o.step_tolerance) ||
                        fabs(cost_before - st.cost_after) < o.cost_tolerance;
            if (!conv && o.has_relative_cost_tolerance && isfinite(o.relative_cost_tolerance) &&
                o.relative_cost_tolerance >= 0.0 &&
                (cost_before - st.cost_after) / fmax(fabs(cost_before), 2.220446049250313e-16) <
                    o.relative_cost_tolerance)
                conv = true;
            if (conv) {
                st.converged = 1;
                stop = true;
            }
        }
    }
    if (st.iterations < B.stats_capacity) B.stats[(size_t)w * B.stats_capacity + st.iterations] = rec;
    st.iterations += 1;
    if (stop) st.active = 0;
}

// Copy the trial state of accepted windows into the current state.
__global__ void k_batch_commit(DevProblem P, BatchDev B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.num_pose_slots && B.state[P.pose_win[i]].accept)
        for (int k = 0; k < 7; ++k) P.poses[7 * (size_t)i + k] = P.poses_trial[7 * (size_t)i + k];
    if (i < P.num_lm_slots && B.state[P.lm_win[i]].accept)
        for (int k = 0; k < 3; ++k) P.lms[3 * (size_t)i + k] = P.lms_trial[3 * (size_t)i + k];
}

// ---------------------------------------------------------------------------
// Joint pose / landmark / camera solve (optimize_joint_intrinsics,
// joint_intrinsics.rs:24-353): the camera parameters (fx, fy, cx, cy[, k1,
// k2[, p1, p2]]) are K extra unknowns bordering the reduced camera system.
// Every observation's pose Jacobian is J_l [I | -[X]x], so its pose-camera
// block is [E; X x E] with E = sum w J_l^T J_k over the edge; the landmark
// block W_lk is the sum of its edges' E.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int joint_tri(int K) { return K * (K + 1) / 2; }
__host__ __device__ constexpr int joint_na(int K) { return 6 * K + joint_tri(K) + K; }

// Residual rows r, landmark Jacobian rows Jl (3 per row) and camera columns
// Jk (K per row) of one observation (joint_mono_terms, :356-465; the
// rectified stereo terms, :702-790). Returns the row count, 0 where the
// reference skips the observation.
template <int K>
__device__ int joint_rows(const DevProblem& P, int o, const double xr[3], const double R[9], double r[3],
                          double Jl[9], double Jk[3 * K]) {
#pragma unroll
    for (int i = 0; i < 3 * K; ++i) Jk[i] = 0.0;
    const DevSensor& S = P.sensors[0];
    if (P.obs_kind && P.obs_kind[o] == OBS_RECTIFIED) {
        double r4[4], J12[12];
        if (!stereo_rows<true>(P, o, P.sensors, S, xr, R, r4, J12)) return 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) r[i] = r4[i];
#pragma unroll
        for (int i = 0; i < 9; ++i) Jl[i] = J12[i];
        // u_r = fx (X - b) / Z + cx: d u_r / d fx = (X - b) / Z, d u_r / d cx = 1
        const double zi = 1.0 / xr[2], b = P.obs_extra[2 * (size_t)o + 1];
        Jk[0] = xr[0] * zi;
        Jk[2] = 1.0;
        Jk[K + 1] = xr[1] * zi;
        Jk[K + 3] = 1.0;
        Jk[2 * K] = (xr[0] - b) * zi;
        Jk[2 * K + 2] = 1.0;
        return 3;
    }
    const Lens<double>& L = S.lens;
    double uv[2], Jp[6];
    if (!project<double>(L, xr[0], xr[1], xr[2], uv, Jp)) return 0;
    r[0] = uv[0] - P.obs_uv[2 * (size_t)o];
    r[1] = uv[1] - P.obs_uv[2 * (size_t)o + 1];
    r[2] = 0.0;
    rows_times(Jp, R, Jl);
    Jl[6] = Jl[7] = Jl[8] = 0.0;
    const double zi = 1.0 / xr[2];
    const double x = xr[0] * zi, y = xr[1] * zi, r2 = x * x + y * y;
    const double d = 1.0 + L.k1 * r2 + L.k2 * r2 * r2;
    double xd = x * d, yd = y * d;
    const double xy2 = 2.0 * x * y;
    if (L.kind == LENS_OPENCV) {
        xd = x * d + L.p1 * xy2 + L.p2 * (r2 + 2.0 * x * x);
        yd = y * d + L.p1 * (r2 + 2.0 * y * y) + L.p2 * xy2;
    }
    Jk[0] = xd;
    Jk[2] = 1.0;
    Jk[K + 1] = yd;
    Jk[K + 3] = 1.0;
    if (K >= 6) {
        Jk[4] = L.fx * x * r2;
        Jk[5] = L.fx * x * r2 * r2;
        Jk[K + 4] = L.fy * y * r2;
        Jk[K + 5] = L.fy * y * r2 * r2;
    }
    if (K == 8) {
        Jk[6] = L.fx * xy2;
        Jk[7] = L.fx * (r2 + 2.0 * x * x);
        Jk[K + 6] = L.fy * (r2 + 2.0 * y * y);
        Jk[K + 7] = L.fy * xy2;
    }
    return 2;
}

// Per camera tile: E per edge, and the tile's J_p^T J_k, J_k^T J_k (upper)
// and J_k^T r (build_joint_intrinsics_system, :593-790).
template <int K>
__global__ __launch_bounds__(kBlock) void k_joint_linearize(DevProblem P) {
    constexpr int NA = joint_na(K), T0 = 6 * K, B0 = 6 * K + joint_tri(K);
    __shared__ double lds[(kBlock / 64) * NA];
    const int tile = xcd_tile(blockIdx.x, P.num_cam_tiles);
    if (tile >= P.num_cam_tiles) return;
    const CamTile ct = P.cam_tiles[tile];
    const double* pq = P.poses + 7 * (size_t)ct.pose;
    double R[9];
    qmatrix<double>(Quat{pq[0], pq[1], pq[2], pq[3]}, R);
    double acc[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) acc[i] = 0.0;
    for (int e = ct.begin + (int)threadIdx.x; e < ct.end; e += kBlock) {
        const int l = P.edge_lm[e];
        const double X[3] = {P.lms[3 * (size_t)l], P.lms[3 * (size_t)l + 1], P.lms[3 * (size_t)l + 2]};
        const double xr[3] = {R[0] * X[0] + R[1] * X[1] + R[2] * X[2] + pq[4],
                              R[3] * X[0] + R[4] * X[1] + R[5] * X[2] + pq[5],
                              R[6] * X[0] + R[7] * X[1] + R[8] * X[2] + pq[6]};
        double E[3 * K];
#pragma unroll
        for (int i = 0; i < 3 * K; ++i) E[i] = 0.0;
        for (int o = P.edge_obs[e]; o < P.edge_obs[e + 1]; ++o) {
            double r[3], Jl[9], Jk[3 * K];
            const int rows = joint_rows<K>(P, o, xr, R, r, Jl, Jk);
            if (!rows) continue;
            double sq = 0.0;
            for (int i = 0; i < rows; ++i) sq += r[i] * r[i];
            const double w = robust_weight<double>((uint32_t)P.kernel_kind, P.kernel_scale, sq);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < K; ++c)
                    E[a * K + c] += w * (Jl[a] * Jk[c] + Jl[3 + a] * Jk[K + c] + Jl[6 + a] * Jk[2 * K + c]);
            int t = T0;
#pragma unroll
            for (int a = 0; a < K; ++a) {
#pragma unroll
                for (int c = a; c < K; ++c)
                    acc[t++] += w * (Jk[a] * Jk[c] + Jk[K + a] * Jk[K + c] + Jk[2 * K + a] * Jk[2 * K + c]);
                acc[B0 + a] += w * (Jk[a] * r[0] + Jk[K + a] * r[1] + Jk[2 * K + a] * r[2]);
            }
        }
        double* Es = P.joint_E + (size_t)3 * K * P.perm_cl[e];
#pragma unroll
        for (int i = 0; i < 3 * K; ++i) Es[i] = E[i];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const double v0 = E[c], v1 = E[K + c], v2 = E[2 * K + c];
            acc[c] += v0;
            acc[K + c] += v1;
            acc[2 * K + c] += v2;
            acc[3 * K + c] += X[1] * v2 - X[2] * v1;
            acc[4 * K + c] += X[2] * v0 - X[0] * v2;
            acc[5 * K + c] += X[0] * v1 - X[1] * v0;
        }
    }
    block_sum<kBlock, double, NA>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NA; ++i) P.joint_tile[(size_t)NA * tile + i] = acc[i];
}

// Per pose slot its J_p^T J_k over its tiles in order; per landmark W_lk
// over its edges in order; thread 0.. of block 0 also sums the camera
// terms over all tiles in order.
template <int K>
__global__ void k_joint_reduce(DevProblem P) {
    constexpr int NA = joint_na(K);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.num_pose_slots)
        for (int a = 0; a < 6 * K; ++a) {
            double s = 0.0;
            for (int t = P.cam_tile_ptr[i]; t < P.cam_tile_ptr[i + 1]; ++t) s += P.joint_tile[(size_t)NA * t + a];
            P.joint_hpk[(size_t)6 * K * i + a] = s;
        }
    if (i < P.num_lm_slots)
        for (int a = 0; a < 3 * K; ++a) {
            double s = 0.0;
            for (int e = P.lm_edge_ptr[i]; e < P.lm_edge_ptr[i + 1]; ++e)
                s += P.joint_E[(size_t)3 * K * P.lm_store[e] + a];
            P.joint_Wk[(size_t)3 * K * i + a] = s;
        }
    if (i < NA - 6 * K) {
        double s = 0.0;
        for (int t = 0; t < P.num_cam_tiles; ++t) s += P.joint_tile[(size_t)NA * t + 6 * K + i];
        P.joint_kk[i] = s;
    }
}

// Border rows of S: S_pk = H_pk - sum over the pose's edges of
// [G; X x G] V^-1 W_lk, one thread per variable pose (joint_intrinsics.rs:
// 143-185).
template <int K>
__global__ void k_joint_border(DevProblem P, double* S, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.num_pose_slots) return;
    const int v = P.pose_var[p];
    if (v < 0) return;
    const int nv = n - K;
    double acc[6 * K];
#pragma unroll
    for (int i = 0; i < 6 * K; ++i) acc[i] = 0.0;
    for (int t = P.cam_tile_ptr[p]; t < P.cam_tile_ptr[p + 1]; ++t) {
        const CamTile ct = P.cam_tiles[t];
        for (int e = ct.begin; e < ct.end; ++e) {
            const int l = P.edge_lm[e];
            const double* G = P.Gl64 + 6 * (size_t)P.perm_cl[e];
            const double* Vi = P.Vinv64 + 6 * (size_t)l;
            const double* W = P.joint_Wk + (size_t)3 * K * l;
            const double X[3] = {P.lms[3 * (size_t)l], P.lms[3 * (size_t)l + 1], P.lms[3 * (size_t)l + 2]};
#pragma unroll
            for (int c = 0; c < K; ++c) {
                const double w0 = W[c], w1 = W[K + c], w2 = W[2 * K + c];
                const double t0 = Vi[0] * w0 + Vi[1] * w1 + Vi[2] * w2;
                const double t1 = Vi[1] * w0 + Vi[3] * w1 + Vi[4] * w2;
                const double t2 = Vi[2] * w0 + Vi[4] * w1 + Vi[5] * w2;
                const double y0 = G[0] * t0 + G[1] * t1 + G[2] * t2;
                const double y1 = G[1] * t0 + G[3] * t1 + G[4] * t2;
                const double y2 = G[2] * t0 + G[4] * t1 + G[5] * t2;
                acc[c] += y0;
                acc[K + c] += y1;
                acc[2 * K + c] += y2;
                acc[3 * K + c] += X[1] * y2 - X[2] * y1;
                acc[4 * K + c] += X[2] * y0 - X[0] * y2;
                acc[5 * K + c] += X[0] * y1 - X[1] * y0;
            }
        }
    }
    const double* H = P.joint_hpk + (size_t)6 * K * p;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const double s = H[a * K + c] - acc[a * K + c];
            S[(size_t)(6 * v + a) * n + nv + c] = s;
            S[(size_t)(nv + c) * n + 6 * v + a] = s;
        }
}

// Camera-block Schur terms sum_l W_lk^T V^-1 W_lk (upper) and
// W_lk^T V^-1 g_l, per work-group over a contiguous landmark range.
template <int K>
__global__ __launch_bounds__(kBlock) void k_joint_kk_partial(DevProblem P) {
    constexpr int NB = joint_tri(K) + K;
    __shared__ double lds[(kBlock / 64) * NB];
    double acc[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) acc[i] = 0.0;
    const int l = blockIdx.x * kBlock + threadIdx.x;
    if (l < P.num_lm_slots) {
        const double* Vi = P.Vinv64 + 6 * (size_t)l;
        const double* W = P.joint_Wk + (size_t)3 * K * l;
        const double* g = P.gl + 3 * (size_t)l;
        double T[3 * K];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const double w0 = W[c], w1 = W[K + c], w2 = W[2 * K + c];
            T[c] = Vi[0] * w0 + Vi[1] * w1 + Vi[2] * w2;
            T[K + c] = Vi[1] * w0 + Vi[3] * w1 + Vi[4] * w2;
            T[2 * K + c] = Vi[2] * w0 + Vi[4] * w1 + Vi[5] * w2;
        }
        int t = 0;
#pragma unroll
        for (int a = 0; a < K; ++a) {
#pragma unroll
            for (int c = a; c < K; ++c) acc[t++] = W[a] * T[c] + W[K + a] * T[K + c] + W[2 * K + a] * T[2 * K + c];
            acc[joint_tri(K) + a] = T[a] * g[0] + T[K + a] * g[1] + T[2 * K + a] * g[2];
        }
    }
    block_sum<kBlock, double, NB>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NB; ++i) P.joint_part[(size_t)NB * blockIdx.x + i] = acc[i];
}

// The camera block S_kk = H_kk + lambda I - sum and its right-hand side
// -b_k + sum, the partials summed in order; then, for a shared focal, fy
// folded into fx (S' = T^T S T with fy's row and column left as the
// identity, joint_intrinsics.rs:187-209). One work-group.
template <int K>
__global__ __launch_bounds__(kSmallBlock) void k_joint_kk_final(DevProblem P, double* S, int n, double lambda,
                                                                int num_parts) {
    constexpr int NB = joint_tri(K) + K;
    const int nv = n - K;
    if (threadIdx.x < NB) {
        const int i = threadIdx.x;
        double s = 0.0;
        for (int b = 0; b < num_parts; ++b) s += P.joint_part[(size_t)NB * b + i];
        if (i < joint_tri(K)) {
            int a = 0, rem = i;
            while (rem >= K - a) {
                rem -= K - a;
                ++a;
            }
            const int c = a + rem;
            double h = P.joint_kk[i];
            if (a == c && lambda > 0.0) h += lambda;
            S[(size_t)(nv + a) * n + nv + c] = h - s;
            S[(size_t)(nv + c) * n + nv + a] = h - s;
        } else {
            P.joint_rhs[i - joint_tri(K)] = -P.joint_kk[i] + s;
        }
    }
    if (!P.joint_shared) return;
    __syncthreads();
    const int fx = nv, fy = nv + 1;
    for (int j = threadIdx.x; j < n; j += kSmallBlock) S[(size_t)fx * n + j] += S[(size_t)fy * n + j];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kSmallBlock) S[(size_t)i * n + fx] += S[(size_t)i * n + fy];
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += kSmallBlock) {
        S[(size_t)fy * n + j] = j == fy ? 1.0 : 0.0;
        if (j != fy) S[(size_t)j * n + fy] = 0.0;
    }
    if (threadIdx.x == 0) {
        P.joint_rhs[0] += P.joint_rhs[1];
        P.joint_rhs[1] = 0.0;
    }
}

// After the solve: a shared focal's step written to fy too; the trial
// camera; the landmark gradient shifted by the camera step, which
// back-substitution reads (delta_l = -V^-1 (g_l + W_lk delta_k + W_pl^T dp)).
template <int K>
__global__ void k_joint_step(DevProblem P) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    double dk[K];
#pragma unroll
    for (int c = 0; c < K; ++c) dk[c] = P.joint_delta[c];
    if (P.joint_shared) dk[1] = dk[0];
    if (l == 0) {
        DevSensor s = P.sensors[0];
        s.lens.fx += dk[0];
        s.lens.fy += dk[1];
        s.lens.cx += dk[2];
        s.lens.cy += dk[3];
        if (K >= 6) {
            s.lens.k1 += dk[4];
            s.lens.k2 += dk[5];
        }
        if (K == 8) {
            s.lens.p1 += dk[6];
            s.lens.p2 += dk[7];
        }
        P.sensors_trial[0] = s;
    }
    if (l < P.num_lm_slots) {
        const double* W = P.joint_Wk + (size_t)3 * K * l;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double t = 0.0;
#pragma unroll
            for (int c = 0; c < K; ++c) t += W[i * K + c] * dk[c];
            P.joint_gl[3 * (size_t)l + i] = P.gl[3 * (size_t)l + i] + t;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

void launch_linearize(const DevProblem& P, bool store32, hipStream_t s) {
    if (P.num_cam_tiles == 0) return;
    if (P.any_rig)
        hipLaunchKernelGGL(k_linearize<true>, dim3(xcd_grid(P.num_cam_tiles)), dim3(kBlock), 0, s, P, (int)store32);
    else
        hipLaunchKernelGGL(k_linearize<false>, dim3(xcd_grid(P.num_cam_tiles)), dim3(kBlock), 0, s, P, (int)store32);
}
void launch_camera_finalize(const DevProblem& P, hipStream_t s) {
    hipLaunchKernelGGL(k_camera_finalize, dim3(blocks_for(P.num_pose_slots, 256)), dim3(256), 0, s, P);
}
void launch_landmark_reduce(const DevProblem& P, hipStream_t s) {
    hipLaunchKernelGGL(k_landmark_reduce, dim3(blocks_for(P.num_lm_slots, 256)), dim3(256), 0, s, P);
}
void launch_landmark_prepare(const DevProblem& P, double lambda, hipStream_t s) {
    hipLaunchKernelGGL(k_landmark_prepare, dim3(blocks_for(P.num_lm_slots, 256)), dim3(256), 0, s, P, lambda);
}
void launch_pose_update(const DevProblem& P, hipStream_t s) {
    hipLaunchKernelGGL(k_pose_update, dim3(blocks_for(P.num_pose_slots, kBlock)), dim3(kBlock), 0, s, P);
}
void launch_cost(const DevProblem& P, bool trial, hipStream_t s) {
    if (P.num_cam_tiles == 0) return;
    if (P.any_rig)
        hipLaunchKernelGGL(k_cost<true>, dim3(xcd_grid(P.num_cam_tiles)), dim3(kBlock), 0, s, P, (int)trial);
    else
        hipLaunchKernelGGL(k_cost<false>, dim3(xcd_grid(P.num_cam_tiles)), dim3(kBlock), 0, s, P, (int)trial);
}
void launch_scalars_finalize(const DevProblem& P, hipStream_t s) {
    hipLaunchKernelGGL(k_scalars_finalize, dim3(1), dim3(kSmallBlock), 0, s, P, blocks_for(P.num_pose_slots, kBlock));
}
template <int K>
void joint_linearize_k(const DevProblem& P, hipStream_t s) {
    if (P.num_cam_tiles)
        hipLaunchKernelGGL(k_joint_linearize<K>, dim3(xcd_grid(P.num_cam_tiles)), dim3(kBlock), 0, s, P);
    const int n = std::max(std::max(P.num_pose_slots, P.num_lm_slots), joint_na(K));
    hipLaunchKernelGGL(k_joint_reduce<K>, dim3(blocks_for(n, 256)), dim3(256), 0, s, P);
}
template <int K>
void joint_border_k(const DevProblem& P, double* S, int32_t n, double lambda, hipStream_t s) {
    if (P.num_pose_slots)
        hipLaunchKernelGGL(k_joint_border<K>, dim3(blocks_for(P.num_pose_slots, 64)), dim3(64), 0, s, P, S, n);
    const int parts = blocks_for(P.num_lm_slots, kBlock);
    if (parts) hipLaunchKernelGGL(k_joint_kk_partial<K>, dim3(parts), dim3(kBlock), 0, s, P);
    hipLaunchKernelGGL(k_joint_kk_final<K>, dim3(1), dim3(kSmallBlock), 0, s, P, S, n, lambda, parts);
}
template <int K>
void joint_step_k(const DevProblem& P, hipStream_t s) {
    hipLaunchKernelGGL(k_joint_step<K>, dim3(blocks_for(std::max(P.num_lm_slots, 1), 256)), dim3(256), 0, s, P);
}
void launch_joint_linearize(const DevProblem& P, hipStream_t s) {
    if (P.joint_k == 4) joint_linearize_k<4>(P, s);
    else if (P.joint_k == 6) joint_linearize_k<6>(P, s);
    else if (P.joint_k == 8) joint_linearize_k<8>(P, s);
}
void launch_joint_border(const DevProblem& P, double* S, int32_t n, double lambda, hipStream_t s) {
    if (P.joint_k == 4) joint_border_k<4>(P, S, n, lambda, s);
    else if (P.joint_k == 6) joint_border_k<6>(P, S, n, lambda, s);
    else if (P.joint_k == 8) joint_border_k<8>(P, S, n, lambda, s);
}
void launch_joint_step(const DevProblem& P, hipStream_t s) {
    if (P.joint_k == 4) joint_step_k<4>(P, s);
    else if (P.joint_k == 6) joint_step_k<6>(P, s);
    else if (P.joint_k == 8) joint_step_k<8>(P, s);
}
size_t joint_part_size(const DevProblem& P) {
    const int K = P.joint_k;
    return (size_t)std::max(blocks_for(P.num_lm_slots, kBlock), 1) * (K * (K + 1) / 2 + K);
}
void launch_dense_assemble(const DevProblem& P, const DensePairs& D, double* S, int32_t n, double lambda,
                           hipStream_t s) {
    if (D.num_blocks)
        hipLaunchKernelGGL(k_dense_assemble, dim3(blocks_for(D.num_blocks, kBlock / kBlockLanes)), dim3(kBlock), 0, s, P,
                           D, S, n, lambda);
}
namespace {
// The device-wide factor and the finishing solve (n > kDenseOneGroupMax).
hipError_t dense_solve_device_wide(const DenseRows& io, double* S, double* work, int32_t n, int32_t* status,
                                   bool allow_lu, hipStream_t s) {
    hipLaunchKernelGGL(k_dense_flags_clear, dim3(1), dim3(64), 0, s, status);
    for (int k0 = 0; k0 < n; k0 += kPanel) {
        const int below = n - k0 - kPanel;
        hipLaunchKernelGGL(k_chol_panel, dim3(below > 0 ? (below + 255) / 256 : 1), dim3(256), 0, s, S, n, k0,
                           status);
        if (below > 0) {
            const long long m = (below + 15) / 16, tiles = m * (m + 1) / 2;
            hipLaunchKernelGGL(k_chol_update, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s, S, n, k0,
                               (const int*)status);
        }
    }
    hipLaunchKernelGGL(k_dense_finish, dim3(1), dim3(kSmallBlock), 0, s, io, S, work, n, status, (int)allow_lu);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_dense_solve(const DevProblem& P, const DensePairs& D, double* S, double* work, int32_t n,
                              int32_t* status, hipStream_t s, bool allow_lu) {
    allow_lu = allow_lu && n <= kLuFallbackMax;
    // the LU fallback's copy of S at full device bandwidth
    if (allow_lu) {
        const hipError_t e = hipMemcpyAsync(work, S, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
    }
    if (n > kDenseOneGroupMax) return dense_solve_device_wide(pose_rows(P, D.var_slot, n), S, work, n, status, allow_lu, s);
    hipLaunchKernelGGL(k_dense_solve, dim3(1), dim3(kSmallBlock), 0, s, P, D, S, work, n, status, (int)allow_lu);
    return hipGetLastError();
}
hipError_t launch_dense_solve_plain(double* S, double* work, const double* b, double* x, int32_t n, int32_t* status,
                                    hipStream_t s) {
    const hipError_t e = hipMemcpyAsync(work, S, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    if (n > kDenseOneGroupMax)
        return dense_solve_device_wide(DenseRows{b, x, nullptr, n, nullptr, nullptr}, S, work, n, status,
                                       n <= kLuFallbackMax, s);
    hipLaunchKernelGGL(k_dense_solve_plain, dim3(1), dim3(kSmallBlock), 0, s, S, work, b, x, n, status);
    return hipGetLastError();
}
void launch_batch_dense_assemble(const DevProblem& P, const DensePairs& D, const BatchDev& B, hipStream_t s) {
    if (D.num_blocks)
        hipLaunchKernelGGL(k_batch_dense_assemble, dim3(blocks_for(D.num_blocks, kBlock / kBlockLanes)), dim3(kBlock),
                           0, s, P, D, B);
}
void launch_batch_dense_zero(const DevProblem& P, const BatchDev& B, hipStream_t s) {
    if (B.num_windows) hipLaunchKernelGGL(k_batch_dense_zero, dim3(16, B.num_windows), dim3(kBlock), 0, s, P, B);
}
void launch_batch_dense_solve(const DevProblem& P, const DensePairs& D, const BatchDev& B, hipStream_t s) {
    if (B.num_windows) hipLaunchKernelGGL(k_batch_dense_solve, dim3(B.num_windows), dim3(kBlock), 0, s, P, D, B);
}
void launch_batch_scalars(const DevProblem& P, const BatchDev& B, hipStream_t s) {
    if (B.num_windows) hipLaunchKernelGGL(k_batch_scalars, dim3(B.num_windows), dim3(kBlock), 0, s, P, B);
}
void launch_batch_init(const DevProblem& P, const BatchDev& B, const vba_options& o, hipStream_t s) {
    (void)P;
    if (B.num_windows)
        hipLaunchKernelGGL(k_batch_init, dim3(blocks_for(B.num_windows, 256)), dim3(256), 0, s, B,
                           o.has_initial_lambda ? o.initial_lambda : 0.0);
}
void launch_batch_control(const DevProblem& P, const BatchDev& B, const vba_options& o, int32_t iteration,
                          hipStream_t s) {
    (void)P;
    ControlOptions c;
    c.has_initial_lambda = (int32_t)o.has_initial_lambda;
    c.has_relative_cost_tolerance = (int32_t)o.has_relative_cost_tolerance;
    c.increase = o.lambda_increase_factor;
    c.decrease = o.lambda_decrease_factor;
    c.max_lambda = o.max_lambda;
    c.min_lambda = o.min_lambda;
    c.step_tolerance = o.step_tolerance;
    c.cost_tolerance = o.cost_tolerance;
    c.relative_cost_tolerance = o.relative_cost_tolerance;
    if (B.num_windows)
        hipLaunchKernelGGL(k_batch_control, dim3(blocks_for(B.num_windows, 256)), dim3(256), 0, s, B, c, iteration);
}
void launch_batch_commit(const DevProblem& P, const BatchDev& B, hipStream_t s) {
    const int n = std::max(P.num_pose_slots, P.num_lm_slots);
    hipLaunchKernelGGL(k_batch_commit, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, B);
}

}  // namespace vba
