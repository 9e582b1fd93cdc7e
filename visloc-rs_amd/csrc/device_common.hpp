// device_common.hpp — device helpers shared by the kernel files: XCD-aware
// tile mapping, fixed-order wave / work-group reductions, compensated sums
// and the small symmetric-block algebra of the Schur path.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"

namespace vba {
namespace {

// Work-groups are dispatched round-robin over the 8 XCDs. Give XCD k a
// contiguous range of tiles so that neighbouring cameras, which share
// landmarks, share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int bid, int n) {
    const int per = (n + 7) >> 3;
    return (bid & 7) * per + (bid >> 3);
}
inline int xcd_grid(int n) { return ((n + 7) >> 3) << 3; }

// Float offsets of G0..G3 and of G4, G5 of the position with index pp in the
// lane-major FP32 G copy (its layout is described with kG32Chunk).
static_assert(kTileEdges == 256 && kG32Chunk == 6 * kTileEdges, "lane-major G chunk of one tile");
__device__ __forceinline__ size_t g32_off4(int pp) {
    return (size_t)(pp >> 8) * kG32Chunk + (pp & 3) * 384 + ((pp & 255) >> 2) * 4;
}
__device__ __forceinline__ size_t g32_off2(int pp) {
    return (size_t)(pp >> 8) * kG32Chunk + (pp & 3) * 384 + 256 + ((pp & 255) >> 2) * 2;
}
inline int blocks_for(int n, int per) { return (n + per - 1) / per; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

// Sum N values over an NT-thread work-group; the result is valid in thread 0
// and the order of additions is fixed. lds must hold (NT/64)*N values.
template <int NT, typename T, int N>
__device__ __forceinline__ void block_sum(T (&v)[N], T* lds) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) lds[wave * N + i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T s = lds[i];
            for (int w = 1; w < NW; ++w) s += lds[w * N + i];
            v[i] = s;
        }
    }
    __syncthreads();
}
// Sum broadcast to every thread.
template <int NT, typename T, int N>
__device__ __forceinline__ void block_sum_all(T (&v)[N], T* lds) {
    block_sum<NT, T, N>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) lds[i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = lds[i];
    __syncthreads();
}
template <int NT>
__device__ __forceinline__ double block_max(double v, double* lds) {
    constexpr int NW = NT / 64;
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lds[0];
        for (int w = 1; w < NW; ++w) m = fmax(m, lds[w]);
        v = m;
    }
    __syncthreads();
    return v;
}

// Neumaier compensated pair (sum, compensation) and its exact-order merge.
struct CSum {
    double s, c;
};
__device__ __forceinline__ void csum_add(CSum& a, double x) {
    double t = a.s + x;
    if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
    else a.c += (x - t) + a.s;
    a.s = t;
}
__device__ __forceinline__ CSum csum_merge(CSum a, CSum b) {
    CSum r;
    r.s = a.s + b.s;
    double bp = r.s - a.s;
    double e = (a.s - (r.s - bp)) + (b.s - bp);
    r.c = a.c + b.c + e;
    return r;
}

// Packed upper-triangle index of a symmetric 6x6.
__host__ __device__ constexpr int up6(int i, int j) { return i * 6 - i * (i - 1) / 2 + (j - i); }

// Accumulate s * [[A, -A K], [K A, -K A K]] (upper triangle) with K = [X]x,
// A symmetric (a00 a01 a02 a11 a12 a22).
template <typename T>
__device__ __forceinline__ void add_schur_block(const T A[6], const T X[3], T s, T out[21]) {
    const T a[9] = {A[0], A[1], A[2], A[1], A[3], A[4], A[2], A[4], A[5]};
    const T K[9] = {T(0), -X[2], X[1], X[2], T(0), -X[0], -X[1], X[0], T(0)};
    T AK[9], KAK[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) AK[3 * i + j] = a[3 * i] * K[j] + a[3 * i + 1] * K[3 + j] + a[3 * i + 2] * K[6 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) KAK[3 * i + j] = K[3 * i] * AK[j] + K[3 * i + 1] * AK[3 + j] + K[3 * i + 2] * AK[6 + j];
    out[up6(0, 0)] += s * a[0];
    out[up6(0, 1)] += s * a[1];
    out[up6(0, 2)] += s * a[2];
    out[up6(1, 1)] += s * a[4];
    out[up6(1, 2)] += s * a[5];
    out[up6(2, 2)] += s * a[8];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[up6(i, 3 + j)] -= s * AK[3 * i + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) out[up6(3 + i, 3 + j)] -= s * KAK[3 * i + j];
}

// y = sym(A) x for packed symmetric 3x3.
template <typename T>
__device__ __forceinline__ void sym3_mul(const T A[6], const T x[3], T y[3]) {
    y[0] = A[0] * x[0] + A[1] * x[1] + A[2] * x[2];
    y[1] = A[1] * x[0] + A[3] * x[1] + A[4] * x[2];
    y[2] = A[2] * x[0] + A[4] * x[1] + A[5] * x[2];
}
// y = sym(U) x for packed symmetric 6x6.
__device__ __forceinline__ void sym6_mul(const double* U, const double* x, double* y) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) s += U[i <= j ? up6(i, j) : up6(j, i)] * x[j];
        y[i] = s;
    }
}

// Sum over the terms k = first, first + stride, ... of block b's term list
// (landmark l, edge of p, edge of q) of the Schur terms
// W_p V_l^-1 W_q^T = [[C, -C K], [K C, -K C K]], C = G_p V_l^-1 G_q,
// K = [X_l]x, into acc (row-major 6x6, zeroed here).
__device__ __forceinline__ void schur_block_terms(const DevProblem& P, const DensePairs& D, int b, int first,
                                                  int stride, double acc[36]) {
#pragma unroll
    for (int i = 0; i < 36; ++i) acc[i] = 0.0;
    for (int k = D.block_ptr[b] + first; k < D.block_ptr[b + 1]; k += stride) {
        const int l = D.terms[3 * (size_t)k], ep = D.terms[3 * (size_t)k + 1], eq = D.terms[3 * (size_t)k + 2];
        const double* Vi = P.Vinv64 + 6 * (size_t)l;
        const double vi[9] = {Vi[0], Vi[1], Vi[2], Vi[1], Vi[3], Vi[4], Vi[2], Vi[4], Vi[5]};
        const double* A = P.Gl64 + 6 * (size_t)ep;
        const double* B = P.Gl64 + 6 * (size_t)eq;
        const double a[9] = {A[0], A[1], A[2], A[1], A[3], A[4], A[2], A[4], A[5]};
        const double bb[9] = {B[0], B[1], B[2], B[1], B[3], B[4], B[2], B[4], B[5]};
        double T[9], C[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) T[3 * i + j] = a[3 * i] * vi[j] + a[3 * i + 1] * vi[3 + j] + a[3 * i + 2] * vi[6 + j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) C[3 * i + j] = T[3 * i] * bb[j] + T[3 * i + 1] * bb[3 + j] + T[3 * i + 2] * bb[6 + j];
        const double* X = P.lms + 3 * (size_t)l;
        const double K[9] = {0.0, -X[2], X[1], X[2], 0.0, -X[0], -X[1], X[0], 0.0};
        double CK[9], KC[9], KCK[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                CK[3 * i + j] = C[3 * i] * K[j] + C[3 * i + 1] * K[3 + j] + C[3 * i + 2] * K[6 + j];
                KC[3 * i + j] = K[3 * i] * C[j] + K[3 * i + 1] * C[3 + j] + K[3 * i + 2] * C[6 + j];
            }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) KCK[3 * i + j] = KC[3 * i] * K[j] + KC[3 * i + 1] * K[3 + j] + KC[3 * i + 2] * K[6 + j];
        // [[C, -C K], [K C, -K C K]]
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                acc[6 * i + j] += C[3 * i + j];
                acc[6 * i + 3 + j] -= CK[3 * i + j];
                acc[6 * (3 + i) + j] += KC[3 * i + j];
                acc[6 * (3 + i) + 3 + j] -= KCK[3 * i + j];
            }
    }
}

// 6x6 Cholesky inverse; false for a non-positive or non-finite pivot.
// Every loop is unrolled, so that L and y stay in registers (dynamic
// indexing would put them in scratch memory).
__device__ inline bool cholesky_inverse6(double A[36], double out[36]) {
    double L[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) L[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = A[6 * j + j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[6 * j + k] * L[6 * j + k];
        if (!(d > 0.0) || !isfinite(d)) return false;
        const double ljj = sqrt(d);
        L[6 * j + j] = ljj;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double s = A[6 * i + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[6 * i + k] * L[6 * j + k];
            L[6 * i + j] = s / ljj;
        }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        double y[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < i; ++k) s -= L[6 * i + k] * y[k];
            y[i] = s / L[6 * i + i];
        }
#pragma unroll
        for (int i = 5; i >= 0; --i) {
            double s = y[i];
#pragma unroll
            for (int k = i + 1; k < 6; ++k) s -= L[6 * k + i] * y[k];
            y[i] = s / L[6 * i + i];
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) out[6 * i + c] = y[i];
    }
    bool finite = true;
#pragma unroll
    for (int i = 0; i < 36; ++i) finite = finite && isfinite(out[i]);
    return finite;
}

}  // namespace
}  // namespace vba
