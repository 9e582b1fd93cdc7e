// explicit_schur.hip — the reduced camera matrix S = U + lambda I - W V^-1 W^T
// assembled explicitly as 6x6 blocks between variable poses that share a
// landmark, and the block-Jacobi PCG of implicit_schur.rs:522-761 run on it.
// The reference states this matrix in its operator tests
// (implicit_schur.rs:831-850, explicit_schur; compared with the implicit
// operator at :952-975 and :1190-1197); here it is a solver of its own: after
// one assembly per LM iteration each PCG product reads 36 values per stored
// block instead of streaming every edge's G.
//
// Assembly: a group of 8 lanes per upper block sums the block's term list
// (one term per landmark the two poses share, in landmark order) with a
// fixed-order butterfly, then writes S_pq and its transpose S_qp. Products:
// one wave per block row, 48 lanes as 6 rows x 8 column-block groups.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device_common.hpp"
#include "pcg_device.hpp"

namespace vba {

namespace {

constexpr int kTermLanes = 8;  // lanes per block in the assembly

// One upper block per 8-lane group: S_pq = [p == q](U_p + lambda I) - sum of
// its terms, stored row-major at block_pos[2b] and transposed at
// block_pos[2b + 1]. The diagonal block keeps its upper triangle and mirrors
// it, so that it is exactly symmetric.
__global__ __launch_bounds__(kBlock) void k_explicit_assemble(DevProblem P, DensePairs D, ExplicitDev X,
                                                               double lambda) {
    const int b = blockIdx.x * (kBlock / kTermLanes) + (int)threadIdx.x / kTermLanes;
    if (b >= D.num_blocks) return;
    const int g = threadIdx.x % kTermLanes;
    double acc[36];
    schur_block_terms(P, D, b, g, kTermLanes, acc);
#pragma unroll
    for (int off = 1; off < kTermLanes; off <<= 1)
#pragma unroll
        for (int i = 0; i < 36; ++i) acc[i] += __shfl_xor(acc[i], off, kTermLanes);
    const int sp = D.block_pq[2 * b], sq = D.block_pq[2 * b + 1];
    const int pos = X.block_pos[2 * b], pos_t = X.block_pos[2 * b + 1];
    const double* U = P.U + 21 * (size_t)sp;
#pragma unroll
    for (int idx0 = 0; idx0 < 36; idx0 += kTermLanes) {
        const int idx = idx0 + g;
        if (idx < 36) {
            const int i = idx / 6, j = idx % 6;
            double v = 0.0;
            if (sp == sq) {
                const int src = i <= j ? idx : 6 * j + i;
#pragma unroll
                for (int t = 0; t < 36; ++t)
                    if (t == src) v = acc[t];
                v = U[i <= j ? up6(i, j) : up6(j, i)] + (i == j && lambda > 0.0 ? lambda : 0.0) - v;
            } else {
#pragma unroll
                for (int t = 0; t < 36; ++t)
                    if (t == idx) v = acc[t];
                v = -v;
            }
            X.S64[36 * (size_t)pos + idx] = v;
            if (X.S32) X.S32[36 * (size_t)pos + idx] = (float)v;
            if (pos_t >= 0) {
                X.S64[36 * (size_t)pos_t + 6 * j + i] = v;
                if (X.S32) X.S32[36 * (size_t)pos_t + 6 * j + i] = (float)v;
            }
        }
    }
}

// Block-Jacobi inverse of each variable pose's diagonal block.
__global__ __launch_bounds__(kBlock) void k_explicit_precond(DevProblem P, ExplicitDev X) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P.num_pose_slots) return;
    double* out = P.Minv + 36 * (size_t)p;
    const int pos = X.diag_pos[p];
    if (pos < 0) {
        for (int i = 0; i < 36; ++i) out[i] = 0.0;
        return;
    }
    double M[36], inv[36];
    for (int i = 0; i < 36; ++i) M[i] = X.S64[36 * (size_t)pos + i];
    if (!cholesky_inverse6(M, inv)) {
        P.scalars->precond_failed = 1;
        for (int i = 0; i < 36; ++i) out[i] = 0.0;
        return;
    }
    for (int i = 0; i < 36; ++i) out[i] = inv[i];
}

// (S v)_p for v = v1 + beta v2 (v2 may be null): lane 8i + g sums row i of
// the row's blocks g, g + 8, ...; a butterfly over g leaves the row sum in
// all 8 lanes of row i. Lanes 48..63 return 0.
template <typename TS>
__device__ __forceinline__ double row_product(const ExplicitDev& X, const TS* __restrict__ S, int p,
                                              const double* __restrict__ v1, const double* __restrict__ v2,
                                              double beta, int lane) {
    const int i = lane >> 3, g = lane & 7;
    double acc = 0.0;
    if (lane < 48) {
        // blocks g, g + 8, ... in that order, four at a time: their loads are
        // issued together before the products
        constexpr int KB = 4;
        const int end = X.row_ptr[p + 1];
        for (int t0 = X.row_ptr[p] + g; t0 < end; t0 += 8 * KB) {
            int q[KB];
#pragma unroll
            for (int u = 0; u < KB; ++u) q[u] = t0 + 8 * u < end ? X.col[t0 + 8 * u] : 0;
            TS s[KB][6];
            double v[KB][6];
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const int t = t0 + 8 * u < end ? t0 + 8 * u : t0;
                load6(S + 36 * (size_t)t + 6 * i, s[u]);
                load6(v1 + 6 * (size_t)q[u], v[u]);
                if (v2) {
                    double w[6];
                    load6(v2 + 6 * (size_t)q[u], w);
#pragma unroll
                    for (int j = 0; j < 6; ++j) v[u][j] = v[u][j] + beta * w[j];
                }
            }
#pragma unroll
            for (int u = 0; u < KB; ++u)
                if (t0 + 8 * u < end)
#pragma unroll
                    for (int j = 0; j < 6; ++j) acc += (double)s[u][j] * v[u][j];
        }
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    acc += __shfl_xor(acc, 4, 64);
    return acc;
}

// d_new = z + beta d_old and q = S d_new for one pose per wave; d.q partials
// per work-group. The work-groups first derive the iteration's scalars.
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_pcg_explicit(DevProblem P, ExplicitDev X, const TS* __restrict__ S,
                                                         const double* __restrict__ d_old, double* __restrict__ d_new,
                                                         int k, int fresh) {
    __shared__ double lds[(kBlock / 64) * 2];
    double beta;
    if (!pcg_prologue(P, k, fresh, &beta)) return;
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    double v[2] = {0.0, 0.0};
    if (p < P.num_pose_slots && P.pose_var[p] >= 0) {
        const double y = row_product<TS>(X, S, p, P.z, fresh ? nullptr : d_old, beta, lane);
        const int c = lane < 6 ? lane : 0;
        const double qc = __shfl(y, 8 * c, 64);
        const size_t kk = 6 * (size_t)p + c;
        const double dc = fresh ? P.z[kk] : P.z[kk] + beta * d_old[kk];
        if (lane < 6) {
            d_new[kk] = dc;
            P.q[kk] = qc;
            v[0] = dc * qc;
        }
        v[0] = wave_sum(v[0]);
    }
    wave_values_to_partial(v, lds, P.bpart + 2 * (size_t)blockIdx.x);
}

// r_t = b - S64 x; with restart r = r_t, z = M^-1 r_t, d = 0. Partials
// (r_t.r_t, r_t.z) per work-group, as the implicit operator's check.
__global__ __launch_bounds__(kBlock) void k_explicit_true_pose(DevProblem P, ExplicitDev X, int restart) {
    __shared__ double lds[(kBlock / 64) * 2];
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    double v[2] = {0.0, 0.0};
    if (p < P.num_pose_slots && P.pose_var[p] >= 0) {
        const double y = row_product<double>(X, X.S64, p, P.x, nullptr, 0.0, lane);
        const int c = lane < 6 ? lane : 0;
        const double sx = __shfl(y, 8 * c, 64);
        const size_t k = 6 * (size_t)p + c;
        const double rt = P.b[k] - sx;
        double zc = 0.0;
        if (restart) zc = pose_minv(P, p, c, 0, rt, lane < 6);
        if (lane < 6) {
            v[0] = rt * rt;
            if (restart) {
                P.r[k] = rt;
                P.z[k] = zc;
                P.d[k] = 0.0;
                v[1] = rt * zc;
            }
        }
        v[0] = wave_sum(v[0]);
        v[1] = wave_sum(v[1]);
    }
    wave_values_to_partial(v, lds, P.bpart + 2 * (size_t)blockIdx.x);
}

}  // namespace

void launch_explicit_assemble(const DevProblem& P, const DensePairs& D, const ExplicitDev& X, double lambda,
                              hipStream_t s) {
    if (D.num_blocks)
        hipLaunchKernelGGL(k_explicit_assemble, dim3(blocks_for(D.num_blocks, kBlock / kTermLanes)), dim3(kBlock), 0,
                           s, P, D, X, lambda);
    hipLaunchKernelGGL(k_explicit_precond, dim3(blocks_for(P.num_pose_slots, kBlock)), dim3(kBlock), 0, s, P, X);
}

void launch_pcg_explicit(const DevProblem& P, const ExplicitDev& X, bool fp32, const double* d_old, double* d_new,
                         int32_t pcg_iter, int32_t fresh, hipStream_t s) {
    const dim3 grid(blocks_for(P.num_pose_slots, kBlock / 64));
    if (fp32)
        hipLaunchKernelGGL(k_pcg_explicit<float>, grid, dim3(kBlock), 0, s, P, X, (const float*)X.S32, d_old, d_new,
                           (int)pcg_iter, (int)fresh);
    else
        hipLaunchKernelGGL(k_pcg_explicit<double>, grid, dim3(kBlock), 0, s, P, X, (const double*)X.S64, d_old, d_new,
                           (int)pcg_iter, (int)fresh);
}

void launch_explicit_true_residual(const DevProblem& P, const ExplicitDev& X, int32_t restart, hipStream_t s) {
    hipLaunchKernelGGL(k_explicit_true_pose, dim3(blocks_for(P.num_pose_slots, kBlock / 64)), dim3(kBlock), 0, s, P,
                       X, (int)restart);
    launch_pcg_true_finish(P, restart, s);
}

}  // namespace vba
