// pcg_device.hpp — device pieces shared by the PCG kernels of the implicit
// (schur.hip) and explicit (explicit_schur.hip) reduced camera operators:
// vector loads, the stop flag, the per-iteration scalars derived inside each
// work-group, fixed-order partial sums and the block-Jacobi product.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace vba {
namespace {

template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
    using type = float2;
};
template <>
struct Vec2<double> {
    using type = double2;
};

// Six consecutive values at an address aligned to two of them.
template <typename T>
__device__ __forceinline__ void load6(const T* __restrict__ p, T v[6]) {
    using V2 = typename Vec2<T>::type;
    const V2* q = reinterpret_cast<const V2*>(p);
    const V2 a = q[0], b = q[1], c = q[2];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
    v[4] = c.x;
    v[5] = c.y;
}

__device__ __forceinline__ void set_stop(PcgState* st, int status) {
    st->status = status;
    __hip_atomic_store(&st->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The PCG stop flag, read by each wave on its own. During a launch it can
// only be set by work-group 0 for a status that every wave derives itself
// from the same inputs (pcg_prologue), so a wave that reads it late still
// takes the same branch as the others.
__device__ __forceinline__ int wave_flag(int32_t* flag) {
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ int wave_stop(PcgState* st) { return wave_flag(&st->stop); }

__device__ __forceinline__ double from_halves(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Lane `src`'s value (src the same in every lane), without LDS.
__device__ __forceinline__ double lane_value(double v, int src) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
    return from_halves(lo, hi);
}

// The value of the partner lane under DPP control CTRL (within a row).
template <int CTRL>
__device__ __forceinline__ double dpp_partner(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return from_halves(lo, hi);
}

// Sum over a wave with the same result in every lane, in VALU lane moves
// only (no LDS): each step adds a lane's value and its partner's, and the
// two lanes of a pair add the same two values. Partners: lane ^ 1, lane ^ 2
// (quad permutes), the mirror lane within 8 and within 16 (row mirrors),
// then the two rows of a row pair and the two halves of the wave (gfx950
// permlane swaps, which leave the lower row's or half's value in the first
// result and the upper one's in the second, in both lanes of a pair). All
// lanes must be active.
__device__ __forceinline__ double wave_allsum(double v) {
    v = v + dpp_partner<0xB1>(v);   // quad_perm [1,0,3,2]
    v = v + dpp_partner<0x4E>(v);   // quad_perm [2,3,0,1]
    v = v + dpp_partner<0x141>(v);  // row_half_mirror
    v = v + dpp_partner<0x140>(v);  // row_mirror
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const auto l16 = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
    const auto h16 = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    v = from_halves(l16[0], h16[0]) + from_halves(l16[1], h16[1]);
    const unsigned long long c = (unsigned long long)__double_as_longlong(v);
    const auto l32 = __builtin_amdgcn_permlane32_swap((unsigned)c, (unsigned)c, false, false);
    const auto h32 = __builtin_amdgcn_permlane32_swap((unsigned)(c >> 32), (unsigned)(c >> 32), false, false);
    return from_halves(l32[0], h32[0]) + from_halves(l32[1], h32[1]);
}

// Sum of pairs (a[2i], a[2i+1]) over i < n, the same in every lane of every
// wave: lanes stride the pairs, then a butterfly.
__device__ __forceinline__ void wave_pair_total(const double* __restrict__ a, int n, double* s0, double* s1) {
    const int lane = threadIdx.x & 63;
    double v0 = 0.0, v1 = 0.0;
    // lane k adds pairs k, k + 64, ... in that order, eight loads at a time
    for (int i0 = lane; i0 < n; i0 += 512) {
        double a0[8], a1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + 64 * u;
            a0[u] = i < n ? a[2 * (size_t)i] : 0.0;
            a1[u] = i < n ? a[2 * (size_t)i + 1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + 64 * u < n) {
                v0 += a0[u];
                v1 += a1[u];
            }
    }
    *s0 = wave_allsum(v0);
    *s1 = wave_allsum(v1);
}

// Inputs of the PCG scalars of iteration k, loaded with no branch in
// between so that a kernel can issue them together with its own loads: the
// stop flag, the step kernel's (r.z, r.r) partial `lane` and the state fields
// the stop test reads.
struct PrologueIn {
    int stop, max_iter;
    double rz0, rr0;  // partial min(lane, num_cpart - 1); lanes past the end drop it
    double rnorm, target, rho_prev;
};

// No value is used here: a use next to its load would wait for it.
__device__ __forceinline__ PrologueIn pcg_prologue_load(const DevProblem& P, int k) {
    const PcgState* st = P.pcg;
    const int i = max(0, min((int)(threadIdx.x & 63), P.num_cpart - 1));
    PrologueIn in;
    in.stop = __hip_atomic_load(&P.pcg->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    in.rz0 = P.cpart[2 * (size_t)i];
    in.rr0 = P.cpart[2 * (size_t)i + 1];
    in.rnorm = st->rnorm;
    in.target = st->target;
    in.rho_prev = st->rho[(k - 1) & 1];
    in.max_iter = st->max_iter;
    return in;
}

// Scalars of PCG iteration k, derived by every wave on its own without
// synchronisation: r.z and r.r of the current residual from the step
// kernel's partials (or the totals the init / restart kernels left in entry
// 0), the stop test on the recursive residual (implicit_schur.rs:640-700)
// and beta = rho_k / rho_{k-1}. Thread 0 of work-group 0 records them.
// Returns false when the solve has stopped; the result is the same in every
// wave of the launch. The partials are summed as wave_pair_total sums them.
__device__ __forceinline__ bool pcg_prologue_finish(const DevProblem& P, int k, int fresh, const PrologueIn& in,
                                                    double* beta_out) {
    PcgState* st = P.pcg;
    const int lane = threadIdx.x & 63;
    const bool own = lane < P.num_cpart;
    double v0 = 0.0 + (own ? in.rz0 : 0.0), v1 = 0.0 + (own ? in.rr0 : 0.0);
    for (int i = lane + 64; i < P.num_cpart; i += 64) {
        v0 += P.cpart[2 * (size_t)i];
        v1 += P.cpart[2 * (size_t)i + 1];
    }
    const double rz = wave_allsum(v0), rr = wave_allsum(v1);
    if (__builtin_amdgcn_readfirstlane(in.stop)) return false;
    double rn, beta = 0.0;
    int status = PCG_RUNNING;
    if (fresh) {
        rn = in.rnorm;
    } else {
        rn = sqrt(rr);
        if (!isfinite(rn)) status = PCG_NONFINITE;
        else if (rn <= in.target) status = PCG_CONVERGED;
        else if (k >= in.max_iter) status = PCG_MAXIT;
        else if (!isfinite(rz) || !(rz > 0.0)) status = PCG_NONPOSITIVE;
        beta = rz / in.rho_prev;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rho[k & 1] = rz;
        st->beta = beta;
        st->iter = k;
        st->rnorm = rn;
        if (status != PCG_RUNNING) set_stop(st, status);
    }
    *beta_out = beta;
    return status == PCG_RUNNING;
}

__device__ bool pcg_prologue(const DevProblem& P, int k, int fresh, double* beta_out) {
    return pcg_prologue_finish(P, k, fresh, pcg_prologue_load(P, k), beta_out);
}

// Sum of one value per wave into a per-work-group partial, in wave order.
__device__ __forceinline__ void wave_values_to_partial(const double v[2], double* lds, double* out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        lds[2 * wave] = v[0];
        lds[2 * wave + 1] = v[1];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = lds[0], b = lds[1];
        for (int w = 1; w < kBlock / 64; ++w) {
            a += lds[2 * w];
            b += lds[2 * w + 1];
        }
        out[0] = a;
        out[1] = b;
    }
}

// z_c = sum_j Minv_cj r_j on lane c < 6 (r_j shuffled from lanes base + j).
__device__ __forceinline__ double pose_minv(const DevProblem& P, int p, int c, int base, double rc, bool valid) {
    double z = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const double rj = __shfl(rc, min(base + j, 63), 64);
        if (valid) z += P.Minv[36 * (size_t)p + 6 * c + j] * rj;
    }
    return z;
}

}  // namespace
}  // namespace vba
