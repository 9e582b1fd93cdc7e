// reduced.hip — the COLMAP-port rig solver's iterative reduced-camera solve
// on the device: the `ReducedSolveAccelerator::solve` hook
// (pipelines/slam/src/ba_accel.rs:46-157) running the CPU solver's
// preconditioned conjugate gradients (pcg_core, colmap_incremental/
// rig_ba_solver.rs:1406-1490) with the same preconditioner and stopping
// rule. The system comes as damped 6x6 diagonal blocks, an edge list
// (a, b), a < b, whose FP32 blocks B are the (row b, column a) blocks, and
// the right-hand side; optionally dense inverses over runs of k consecutive
// nodes (cluster_inverses, :1112-1166) and the inverse of an aggregated
// coarse matrix (coarse_correction, :1235-1275).
//
// Arithmetic is FP64 throughout; the blocks are read as the FP32 values the
// caller stores, as the CPU solver's default product does (:1069-1072).
// Every reduction has a fixed order, so a solve is deterministic. One PCG
// iteration is a fixed sequence of kernels that decide the stop themselves
// (each work-group derives the same scalars from the same partials); the
// host reads the state once per ten iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "pcg_device.hpp"
#include "reduced.hpp"
#include "solver.hpp"

namespace vba {

namespace {

enum RsStatus : int32_t { RS_RUNNING = 0, RS_CONVERGED = 1, RS_MAXIT = 2, RS_ZERO = 3, RS_FAILED = -1 };

struct RsState {
    double b_norm, tol;
    double rz[2];  // r.z by iteration parity
    int32_t iter, status, stop, max_iter;
};

struct RsDev {
    int32_t n, num_edges, cluster_k, num_clusters;
    int32_t coarse_group, coarse_nodes, coarse_m, pad;
    const double* diag;        // 36 per node, column-major
    const uint32_t* edges;     // 2 per edge
    const float* off;          // 36 per edge, (row b, column a) block, column-major
    const double* rhs;         // 6 per node
    const int32_t* in_ptr;     // node i: edges with b == i are in_edge[in_ptr[i] .. in_ptr[i+1]]
    const int32_t* in_edge;
    const int32_t* in_from;    // the other node (a) of each in_edge entry
    float* off_in;             // the blocks in in_edge order, row-major (B's rows contiguous)
    const int32_t* out_ptr;    // node i: edges with a == i are out_ptr[i] .. out_ptr[i+1]
    const double* cl_inv;      // cluster inverses, column-major, concatenated
    const int64_t* cl_off;     // offset of cluster c's inverse (doubles)
    const double* cinv;        // coarse inverse, column-major (6M)^2
    double *x, *r, *z, *p, *ap, *minv, *rc, *yc;
    double* part_a;            // per-work-group partials of p.ap (matvec) and r.z
    double* part_b;            // per-work-group partials of r.r (update)
    RsState* st;
};

constexpr int kRsThreads = 256;
constexpr int kRsMaxCluster = 256;      // nodes per cluster run
constexpr int kRsMaxAggregates = 512;   // coarse aggregates

// Block-Jacobi inverses (the CPU computes them for every solve, clusters or
// not, and gives up on a block that is not positive definite), x = 0,
// r = b and per-work-group b.b partials.
__global__ __launch_bounds__(kRsThreads) void k_rs_prepare(RsDev D) {
    __shared__ double lds[(kRsThreads / 64) * 2];
    const int t = blockIdx.x * kRsThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (t < 6 * D.n) {
        const double b = D.rhs[t];
        D.x[t] = 0.0;
        D.r[t] = b;
        v[0] = b * b;
    }
    if (t < D.n) {
        double M[36], inv[36];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) M[6 * r + c] = D.diag[36 * (size_t)t + 6 * c + r];
        if (!cholesky_inverse6(M, inv)) {
            D.st->status = RS_FAILED;
            for (int i = 0; i < 36; ++i) inv[i] = 0.0;
        }
        for (int i = 0; i < 36; ++i) D.minv[36 * (size_t)t + i] = inv[i];
    }
    v[0] = wave_sum(v[0]);
    wave_values_to_partial(v, lds, D.part_b + 2 * (size_t)blockIdx.x);
}

__global__ __launch_bounds__(64) void k_rs_start(RsDev D, int num_part, double tol, int max_iter) {
    double bb, unused;
    wave_pair_total(D.part_b, num_part, &bb, &unused);
    if (threadIdx.x == 0) {
        RsState* st = D.st;
        st->b_norm = sqrt(bb);
        st->tol = tol;
        st->rz[0] = st->rz[1] = 0.0;
        st->iter = 0;
        st->max_iter = max_iter;
        if (st->status == RS_FAILED) st->stop = 1;
        else if (st->b_norm == 0.0) st->status = RS_ZERO, st->stop = 1;
        else if (max_iter <= 0) st->status = RS_MAXIT, st->stop = 1;
        else st->status = RS_RUNNING, st->stop = 0;
    }
}

// z = M^-1 r for the nodes' local preconditioner: block-Jacobi (thread per
// component) or one cluster inverse per work-group. At iteration k >= 0 the
// work-groups first apply the stop test of pcg_core to the r.r partials.
// Without a coarse level the work-groups also leave r.z partials.
__device__ bool rs_check(const RsDev& D, int k, int num_part) {
    if (wave_flag(&D.st->stop)) return false;
    if (k < 0) return true;
    double rr, unused;
    wave_pair_total(D.part_b, num_part, &rr, &unused);
    int status = RS_RUNNING;
    if (sqrt(rr) <= D.st->tol * D.st->b_norm) status = RS_CONVERGED;
    else if (k + 1 >= D.st->max_iter) status = RS_MAXIT;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        D.st->iter = k + 1;
        if (status != RS_RUNNING) {
            D.st->status = status;
            __hip_atomic_store(&D.st->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    return status == RS_RUNNING;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_jacobi(RsDev D, int k, int num_part, int dot) {
    __shared__ double lds[(kRsThreads / 64) * 2];
    if (!rs_check(D, k, num_part)) return;
    const int t = blockIdx.x * kRsThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (t < 6 * D.n) {
        const int i = t / 6, c = t % 6;
        double z = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) z += D.minv[36 * (size_t)i + 6 * c + j] * D.r[6 * (size_t)i + j];
        D.z[t] = z;
        v[0] = D.r[t] * z;
    }
    if (dot) {
        v[0] = wave_sum(v[0]);
        wave_values_to_partial(v, lds, D.part_a + 2 * (size_t)blockIdx.x);
    }
}

__global__ __launch_bounds__(kRsThreads) void k_rs_cluster(RsDev D, int k, int num_part, int dot) {
    __shared__ double sr[6 * kRsMaxCluster];
    __shared__ double lds[(kRsThreads / 64) * 2];
    if (!rs_check(D, k, num_part)) return;
    const int c = blockIdx.x;
    const int s = c * D.cluster_k;
    const int m = min(D.cluster_k, D.n - s), rows = 6 * m;
    for (int t = threadIdx.x; t < rows; t += kRsThreads) sr[t] = D.r[6 * (size_t)s + t];
    __syncthreads();
    const double* inv = D.cl_inv + D.cl_off[c];
    double v[2] = {0.0, 0.0};
    for (int t = threadIdx.x; t < rows; t += kRsThreads) {
        double z = 0.0;
        for (int col = 0; col < rows; ++col) z += inv[(size_t)col * rows + t] * sr[col];
        D.z[6 * (size_t)s + t] = z;
        v[0] += sr[t] * z;
    }
    if (dot) {
        v[0] = wave_sum(v[0]);
        wave_values_to_partial(v, lds, D.part_a + 2 * (size_t)blockIdx.x);
    }
}

// Coarse level (CoarseCorrection::add_to, rig_ba_solver.rs:1196-1214):
// rc = P^T r over the leading nodes, yc = C^-1 rc, z += P yc.
__global__ __launch_bounds__(kRsThreads) void k_rs_restrict(RsDev D) {
    if (wave_flag(&D.st->stop)) return;
    const int t = blockIdx.x * kRsThreads + threadIdx.x;
    if (t >= 6 * D.coarse_m) return;
    const int g = t / 6, c = t % 6;
    const int i1 = min(D.coarse_nodes, (g + 1) * D.coarse_group);
    double s = 0.0;
    for (int i = g * D.coarse_group; i < i1; ++i) s += D.r[6 * (size_t)i + c];
    D.rc[t] = s;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_coarse(RsDev D) {
    __shared__ double src[6 * kRsMaxAggregates];
    if (wave_flag(&D.st->stop)) return;
    const int rows = 6 * D.coarse_m;
    for (int t = threadIdx.x; t < rows; t += kRsThreads) src[t] = D.rc[t];
    __syncthreads();
    const int t = blockIdx.x * kRsThreads + threadIdx.x;
    if (t >= rows) return;
    double y = 0.0;
    for (int col = 0; col < rows; ++col) y += D.cinv[(size_t)col * rows + t] * src[col];
    D.yc[t] = y;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_prolong(RsDev D) {
    __shared__ double lds[(kRsThreads / 64) * 2];
    if (wave_flag(&D.st->stop)) return;
    const int t = blockIdx.x * kRsThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (t < 6 * D.n) {
        const int i = t / 6, c = t % 6;
        double z = D.z[t];
        if (i < D.coarse_nodes) z += D.yc[6 * (i / D.coarse_group) + c];
        D.z[t] = z;
        v[0] = D.r[t] * z;
    }
    v[0] = wave_sum(v[0]);
    wave_values_to_partial(v, lds, D.part_a + 2 * (size_t)blockIdx.x);
}

// p = z + beta p with beta = r.z / (previous r.z); at the start (k < 0)
// p = z. Records r.z for iteration k + 1.
__global__ __launch_bounds__(kRsThreads) void k_rs_direction(RsDev D, int k, int num_part) {
    if (wave_flag(&D.st->stop)) return;
    double rz, unused;
    wave_pair_total(D.part_a, num_part, &rz, &unused);
    const int t = blockIdx.x * kRsThreads + threadIdx.x;
    const int next = (k + 1) & 1;
    if (t < 6 * D.n) {
        if (k < 0) {
            D.p[t] = D.z[t];
        } else {
            const double beta = rz / D.st->rz[k & 1];
            D.p[t] = D.z[t] + beta * D.p[t];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) D.st->rz[next] = rz;
}

// ap = A p, one wave per node: 48 lanes as 6 rows x 8 edge groups, the
// diagonal block on group 0; in-edges add B p_a, out-edges B^T p_b; a
// butterfly over the groups. A group takes its edges four at a time, their
// loads issued together. Nodes are spread over the XCDs in contiguous
// ranges. Per-work-group p.ap partials.
__global__ __launch_bounds__(kRsThreads) void k_rs_matvec(RsDev D) {
    __shared__ double lds[(kRsThreads / 64) * 2];
    if (wave_flag(&D.st->stop)) return;
    constexpr int KB = 4;
    const int lane = threadIdx.x & 63;
    const int wg = xcd_tile(blockIdx.x, (D.n + kRsThreads / 64 - 1) / (kRsThreads / 64));
    const int node = wg * (kRsThreads / 64) + (threadIdx.x >> 6);
    double v[2] = {0.0, 0.0};
    if (node < D.n) {
        const int i = lane >> 3, g = lane & 7;
        double acc = 0.0;
        if (lane < 48) {
            if (g == 0) {
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    acc += D.diag[36 * (size_t)node + 6 * c + i] * D.p[6 * (size_t)node + c];
            }
            const int in_end = D.in_ptr[node + 1];
            for (int q0 = D.in_ptr[node] + g; q0 < in_end; q0 += 8 * KB) {
                int a[KB];
#pragma unroll
                for (int u = 0; u < KB; ++u) a[u] = D.in_from[q0 + 8 * u < in_end ? q0 + 8 * u : q0];
                float B[KB][6];
                double x[KB][6];
#pragma unroll
                for (int u = 0; u < KB; ++u) {
                    const int q = q0 + 8 * u < in_end ? q0 + 8 * u : q0;
                    load6(D.off_in + 36 * (size_t)q + 6 * i, B[u]);
                    load6(D.p + 6 * (size_t)a[u], x[u]);
                }
#pragma unroll
                for (int u = 0; u < KB; ++u)
                    if (q0 + 8 * u < in_end)
#pragma unroll
                        for (int c = 0; c < 6; ++c) acc += (double)B[u][c] * x[u][c];
            }
            const int out_end = D.out_ptr[node + 1];
            for (int e0 = D.out_ptr[node] + g; e0 < out_end; e0 += 8 * KB) {
                int b[KB];
#pragma unroll
                for (int u = 0; u < KB; ++u) b[u] = (int)D.edges[2 * (size_t)(e0 + 8 * u < out_end ? e0 + 8 * u : e0) + 1];
                float B[KB][6];
                double x[KB][6];
#pragma unroll
                for (int u = 0; u < KB; ++u) {
                    const int e = e0 + 8 * u < out_end ? e0 + 8 * u : e0;
                    load6(D.off + 36 * (size_t)e + 6 * i, B[u]);
                    load6(D.p + 6 * (size_t)b[u], x[u]);
                }
#pragma unroll
                for (int u = 0; u < KB; ++u)
                    if (e0 + 8 * u < out_end)
#pragma unroll
                        for (int r = 0; r < 6; ++r) acc += (double)B[u][r] * x[u][r];
            }
        }
        acc += __shfl_xor(acc, 1, 64);
        acc += __shfl_xor(acc, 2, 64);
        acc += __shfl_xor(acc, 4, 64);
        const int c = lane < 6 ? lane : 0;
        const double y = __shfl(acc, 8 * c, 64);
        if (lane < 6) {
            D.ap[6 * (size_t)node + c] = y;
            v[0] = D.p[6 * (size_t)node + c] * y;
        }
        v[0] = wave_sum(v[0]);
    }
    wave_values_to_partial(v, lds, D.part_a + 2 * (size_t)blockIdx.x);
}

// The blocks in in_edge order, row-major, so that a node reads the blocks of
// its in-edges as one contiguous stream (its out-edges' blocks already are).
__global__ __launch_bounds__(kRsThreads) void k_rs_gather_in(RsDev D) {
    const size_t t = (size_t)blockIdx.x * kRsThreads + threadIdx.x;
    if (t >= 36 * (size_t)D.num_edges) return;
    const size_t q = t / 36;
    const int k = (int)(t % 36), i = k / 6, c = k % 6;
    D.off_in[t] = D.off[36 * (size_t)D.in_edge[q] + 6 * c + i];
}

// alpha = r.z / p.ap (the solve gives up when p.ap is not positive and
// finite), x += alpha p, r -= alpha ap; per-work-group r.r partials.
__global__ __launch_bounds__(kRsThreads) void k_rs_update(RsDev D, int k, int num_part) {
    __shared__ double lds[(kRsThreads / 64) * 2];
    if (wave_flag(&D.st->stop)) return;
    double pap, unused;
    wave_pair_total(D.part_a, num_part, &pap, &unused);
    if (!isfinite(pap) || !(pap > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            D.st->status = RS_FAILED;
            D.st->iter = k + 1;
            __hip_atomic_store(&D.st->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    const double alpha = D.st->rz[k & 1] / pap;
    const int t = blockIdx.x * kRsThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (t < 6 * D.n) {
        D.x[t] = D.x[t] + alpha * D.p[t];
        const double r = D.r[t] - alpha * D.ap[t];
        D.r[t] = r;
        v[0] = r * r;
    }
    v[0] = wave_sum(v[0]);
    wave_values_to_partial(v, lds, D.part_b + 2 * (size_t)blockIdx.x);
}

// ---------------------------------------------------------------------------
// Point elimination (ReducedSolveAccelerator::eliminate, ba_accel.rs:71-157):
// the blocks -sum_p hcp[p, j] V_p^-1 hcp[p, k]^T of the CPU rig solver's
// eliminate_by_output (rig_ba_solver.rs:2551-2660), in its arithmetic: the
// product hcp[k] V^-1 with separately rounded multiplies and adds (a column
// sweep), each entry a chain of three fused multiply-adds per point over the
// points in ascending order, with a fresh partial sum per point shard added
// to the total when the shard ends. The pair index lists every (point, k < j)
// term under its edge, points ascending.
// ---------------------------------------------------------------------------
struct ElDev {
    int32_t num_edges, num_slots, num_extra, num_run_edges;
    int32_t coarse_group, coarse_nodes, coarse_m, pad;
    const int64_t* term_ptr;     // per edge: its terms
    const uint32_t* terms;       // 3 per term: slot index of k, slot index of j, point
    const uint8_t* shard;        // per point
    const uint32_t* point_of;    // per slot index: its point
    const double* hcp;           // 18 per slot, 6x3 column-major
    const double* vinv;          // 9 per point, column-major
    const uint8_t* has;          // per point
    double* scaled;              // 18 per slot: hcp V^-1
    double* b64;                 // 36 per edge, column-major
    float* off;                  // the solve's FP32 blocks
    const uint32_t* extra_edge;
    const double* extra_blocks;
    const uint32_t* run_edges;
    double* run_out;
    const uint32_t* edges;
    const int32_t* out_ptr;      // per node: its edges (s, _)
    double* coarse_out;
};

__global__ __launch_bounds__(kRsThreads) void k_el_scaled(ElDev D) {
    // every product and sum rounded on its own, as the CPU's (no fusion)
#pragma clang fp contract(off)
    const int q = blockIdx.x * kRsThreads + threadIdx.x;
    if (q >= D.num_slots) return;
    const uint32_t p = D.point_of[q];
    if (!D.has[p]) return;
    const double* h = D.hcp + 18 * (size_t)q;
    const double* v = D.vinv + 9 * (size_t)p;
    double* o = D.scaled + 18 * (size_t)q;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double x = h[r] * v[3 * c];
            x = x + h[6 + r] * v[3 * c + 1];
            x = x + h[12 + r] * v[3 * c + 2];
            o[6 * c + r] = x;
        }
}

// One wave per edge block; lane q < 36 forms entry q (column q / 6, row
// q % 6) over the edge's terms in point order.
__global__ __launch_bounds__(kRsThreads) void k_el_blocks(ElDev D) {
    const int lane = threadIdx.x & 63;
    const int wg = xcd_tile(blockIdx.x, (D.num_edges + kRsThreads / 64 - 1) / (kRsThreads / 64));
    const int e = wg * (kRsThreads / 64) + (threadIdx.x >> 6);
    if (e >= D.num_edges || lane >= 36) return;
    const int c = lane / 6, r = lane % 6;
    double total = 0.0, acc = 0.0;
    int cur = -1;
    const int64_t t1 = D.term_ptr[e + 1];
    for (int64_t t = D.term_ptr[e]; t < t1; ++t) {
        const uint32_t sk = D.terms[3 * t], sj = D.terms[3 * t + 1], p = D.terms[3 * t + 2];
        if (!D.has[p]) continue;
        const int sh = D.shard[p];
        if (sh != cur) {
            if (cur >= 0) {
                total = total + acc;
                acc = 0.0;
            }
            cur = sh;
        }
        const double* b = D.scaled + 18 * (size_t)sk;
        const double* a = D.hcp + 18 * (size_t)sj;
        acc = fma(a[12 + r], -b[12 + c], fma(a[6 + r], -b[6 + c], fma(a[r], -b[c], acc)));
    }
    if (cur >= 0) total = total + acc;
    D.b64[36 * (size_t)e + lane] = total;
}

// Direct couplings (one per edge at most), then the FP32 blocks of the solve.
__global__ __launch_bounds__(kRsThreads) void k_el_extra(ElDev D) {
    const size_t i = (size_t)blockIdx.x * kRsThreads + threadIdx.x;
    if (i >= 36 * (size_t)D.num_extra) return;
    const size_t k = 36 * (size_t)D.extra_edge[i / 36] + i % 36;
    D.b64[k] = D.b64[k] + D.extra_blocks[i];
}
__global__ __launch_bounds__(kRsThreads) void k_el_store(ElDev D) {
    const size_t i = (size_t)blockIdx.x * kRsThreads + threadIdx.x;
    if (i < 36 * (size_t)D.num_edges) D.off[i] = __double2float_rn(D.b64[i]);
    if (i < 36 * (size_t)D.num_run_edges) D.run_out[i] = D.b64[36 * (size_t)D.run_edges[i / 36] + i % 36];
}

// Aggregate pair (g, h), g <= h: the sum of the blocks of edges (a, b), a in
// g and b in h (both below coarse_nodes), in edge order. One thread per pair.
__global__ __launch_bounds__(kRsThreads) void k_el_coarse(ElDev D) {
    const int m = D.coarse_m;
    const int idx = blockIdx.x * kRsThreads + threadIdx.x;
    if (idx >= m * m) return;
    const int g = idx / m, h = idx % m;
    double* o = D.coarse_out + 36 * (size_t)idx;
    double sum[36];
#pragma unroll
    for (int q = 0; q < 36; ++q) sum[q] = 0.0;
    if (g <= h) {
        const int grp = D.coarse_group, nodes = D.coarse_nodes;
        const uint32_t t0 = (uint32_t)(h * grp), t1 = (uint32_t)min((h + 1) * grp, nodes);
        for (int a = g * grp; a < min((g + 1) * grp, nodes); ++a) {
            int lo = D.out_ptr[a], hi = D.out_ptr[a + 1];
            while (lo < hi) {  // first edge of row a with b >= t0
                const int mid = (lo + hi) >> 1;
                if (D.edges[2 * (size_t)mid + 1] < t0) lo = mid + 1;
                else hi = mid;
            }
            for (int e = lo; e < D.out_ptr[a + 1] && D.edges[2 * (size_t)e + 1] < t1; ++e)
#pragma unroll
                for (int q = 0; q < 36; ++q) sum[q] = sum[q] + D.b64[36 * (size_t)e + q];
        }
    }
#pragma unroll
    for (int q = 0; q < 36; ++q) o[q] = sum[q];
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct ReducedState {
    uint64_t structure_id = 0;
    int32_t n = -1, num_edges = -1;
    // the point elimination's pair index (for elim_structure) and the
    // structure whose blocks the last elimination left in the solve's buffer
    uint64_t elim_structure = 0, blocks_of = 0;
    std::vector<void*> bufs;
    std::vector<size_t> sizes;
};

namespace {

enum RsBuf {
    R_EDGES, R_IN_PTR, R_IN_EDGE, R_IN_FROM, R_OUT_PTR, R_DIAG, R_OFF, R_OFF_IN, R_RHS, R_CL_INV, R_CL_OFF, R_CINV, R_X, R_R, R_Z, R_P,
    R_AP, R_MINV, R_RC, R_YC, R_PART_A, R_PART_B, R_STATE, R_EL_TERM_PTR, R_EL_TERMS, R_EL_SHARD, R_EL_POINT_OF,
    R_EL_HCP, R_EL_VINV, R_EL_HAS, R_EL_SCALED, R_EL_B64, R_EL_EXTRA_E, R_EL_EXTRA_B, R_EL_RUN_E, R_EL_RUN_OUT,
    R_EL_COARSE, R_COUNT
};

struct RsFail {
    int32_t code;
    std::string msg;
};

void rs_check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw RsFail{e == hipErrorOutOfMemory ? VBA_ERR_OUT_OF_MEMORY : VBA_ERR_DEVICE,
                     std::string(what) + ": " + hipGetErrorString(e)};
}

void* rs_buf(ReducedState& R, int id, size_t bytes) {
    if (R.bufs.size() < R_COUNT) {
        R.bufs.resize(R_COUNT, nullptr);
        R.sizes.resize(R_COUNT, 0);
    }
    if (bytes == 0) bytes = 64;
    if (R.sizes[id] < bytes) {
        if (R.bufs[id]) rs_check_hip(hipFree(R.bufs[id]), "hipFree");
        R.bufs[id] = nullptr;
        R.sizes[id] = 0;
        rs_check_hip(hipMalloc(&R.bufs[id], bytes), "hipMalloc");
        R.sizes[id] = bytes;
    }
    return R.bufs[id];
}

template <typename T>
T* rs_upload(ReducedState& R, int id, const T* src, size_t count, hipStream_t s) {
    T* p = (T*)rs_buf(R, id, count * sizeof(T));
    if (count) rs_check_hip(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    return p;
}

// The solve's edge structure (in/out lists of every node), rebuilt unless the
// caller's structure id (nonzero) and the sizes are those it was built for.
// A rebuild forgets the blocks of an earlier elimination.
void ensure_structure(ReducedState& R, uint64_t structure_id, uint32_t n, uint32_t E, const uint32_t* edges,
                      hipStream_t s) {
    if (structure_id != 0 && structure_id == R.structure_id && R.n == (int32_t)n && R.num_edges == (int32_t)E)
        return;
    std::vector<int32_t> in_ptr(n + 1, 0), in_edge(E), out_ptr(n + 1, 0);
    for (uint32_t e = 0; e < E; ++e) {
        ++out_ptr[edges[2 * (size_t)e] + 1];
        ++in_ptr[edges[2 * (size_t)e + 1] + 1];
    }
    for (uint32_t i = 0; i < n; ++i) {
        out_ptr[i + 1] += out_ptr[i];
        in_ptr[i + 1] += in_ptr[i];
    }
    std::vector<int32_t> cur(in_ptr.begin(), in_ptr.end() - 1), in_from(E);
    for (uint32_t e = 0; e < E; ++e) {
        const int32_t q = cur[edges[2 * (size_t)e + 1]]++;
        in_edge[q] = (int32_t)e;
        in_from[q] = (int32_t)edges[2 * (size_t)e];
    }
    rs_upload(R, R_EDGES, edges, 2 * (size_t)E, s);
    rs_upload(R, R_IN_PTR, in_ptr.data(), in_ptr.size(), s);
    rs_upload(R, R_IN_EDGE, in_edge.data(), in_edge.size(), s);
    rs_upload(R, R_IN_FROM, in_from.data(), in_from.size(), s);
    rs_upload(R, R_OUT_PTR, out_ptr.data(), out_ptr.size(), s);
    rs_check_hip(hipStreamSynchronize(s), "sync");  // the host vectors go out of scope
    R.structure_id = structure_id;
    R.n = (int32_t)n;
    R.num_edges = (int32_t)E;
    R.blocks_of = 0;
}

int32_t decline(std::string* msg, const char* why) {
    *msg = why;
    return VBA_DECLINED;
}

}  // namespace

void release_reduced(Context& ctx) {
    if (!ctx.reduced) return;
    for (void* p : ctx.reduced->bufs)
        if (p) (void)hipFree(p);
    delete ctx.reduced;
    ctx.reduced = nullptr;
}

int32_t reduced_solve(Context& ctx, const vba_reduced_system& sys, double* step, uint32_t* iterations,
                      std::string* msg) {
    const uint32_t n = sys.num_nodes, E = sys.num_edges;
    // Shapes the solve supports (as the reference GPU backend's checks,
    // crates/ba-gpu/src/reduced.rs:654-760, decline the rest).
    if (n == 0 || !sys.diag || !sys.rhs || (E && !sys.edges) || E >= (1u << 31) || n > (1u << 24))
        return decline(msg, "reduced system: empty, missing arrays or too large");
    if (E && !sys.offdiag &&
        !(ctx.reduced && sys.structure_id != 0 && ctx.reduced->blocks_of == sys.structure_id &&
          ctx.reduced->structure_id == sys.structure_id && ctx.reduced->n == (int32_t)n &&
          ctx.reduced->num_edges == (int32_t)E))
        return decline(msg, "reduced system: no offdiag blocks, and no elimination of this structure on the device");
    for (uint32_t e = 0; e < E; ++e) {
        const uint32_t a = sys.edges[2 * (size_t)e], b = sys.edges[2 * (size_t)e + 1];
        if (!(a < b && b < n)) return decline(msg, "reduced system: an edge is not (a, b) with a < b < n");
        if (e && !(sys.edges[2 * (size_t)e - 2] < a || (sys.edges[2 * (size_t)e - 2] == a && sys.edges[2 * (size_t)e - 1] < b)))
            return decline(msg, "reduced system: edges are not sorted");
    }
    for (size_t i = 0; i < 6 * (size_t)n; ++i)
        if (!std::isfinite(sys.rhs[i])) return decline(msg, "reduced system: non-finite right-hand side");
    if (!(std::isfinite(sys.tol) && sys.tol >= 0.0)) return decline(msg, "reduced system: invalid tolerance");
    const bool clusters = sys.cluster_size > 1;
    const int32_t k = clusters ? (int32_t)sys.cluster_size : 1;
    if (clusters && (!sys.cluster_inverses || sys.cluster_size > (uint32_t)kRsMaxCluster))
        return decline(msg, "reduced system: cluster runs longer than 256 nodes or missing inverses");
    const bool coarse = sys.coarse_group > 0;
    int32_t cm = 0;
    if (coarse) {
        if (!sys.coarse_inverse || sys.coarse_nodes > n || sys.coarse_nodes == 0)
            return decline(msg, "reduced system: invalid coarse level");
        cm = (int32_t)((sys.coarse_nodes + sys.coarse_group - 1) / sys.coarse_group);
        if (cm > kRsMaxAggregates) return decline(msg, "reduced system: more than 512 coarse aggregates");
    }

    if (!ctx.reduced) ctx.reduced = new ReducedState();
    ReducedState& R = *ctx.reduced;
    hipStream_t s = ctx.stream;
    try {
        RsDev D{};
        D.n = (int32_t)n;
        D.num_edges = (int32_t)E;
        D.cluster_k = k;
        D.num_clusters = clusters ? (int32_t)((n + k - 1) / k) : 0;
        D.coarse_group = coarse ? (int32_t)sys.coarse_group : 1;
        D.coarse_nodes = coarse ? (int32_t)sys.coarse_nodes : 0;
        D.coarse_m = cm;
        ensure_structure(R, sys.structure_id, n, E, sys.edges, s);
        D.edges = (const uint32_t*)R.bufs[R_EDGES];
        D.in_ptr = (const int32_t*)R.bufs[R_IN_PTR];
        D.in_edge = (const int32_t*)R.bufs[R_IN_EDGE];
        D.in_from = (const int32_t*)R.bufs[R_IN_FROM];
        D.out_ptr = (const int32_t*)R.bufs[R_OUT_PTR];
        D.diag = rs_upload(R, R_DIAG, sys.diag, 36 * (size_t)n, s);
        if (sys.offdiag || !E) {
            D.off = rs_upload(R, R_OFF, sys.offdiag, 36 * (size_t)E, s);
            R.blocks_of = 0;
        } else {
            D.off = (const float*)R.bufs[R_OFF];  // formed by reduced_eliminate for this structure
        }
        D.off_in = (float*)rs_buf(R, R_OFF_IN, 36 * (size_t)E * sizeof(float));
        D.rhs = rs_upload(R, R_RHS, sys.rhs, 6 * (size_t)n, s);
        std::vector<int64_t> cl_off;
        if (clusters) {
            cl_off.resize(D.num_clusters + 1);
            int64_t total = 0;
            for (int32_t c = 0; c < D.num_clusters; ++c) {
                const int64_t m = std::min<int64_t>(k, (int64_t)n - (int64_t)c * k);
                cl_off[c] = total;
                total += 36 * m * m;
            }
            cl_off[D.num_clusters] = total;
            D.cl_inv = rs_upload(R, R_CL_INV, sys.cluster_inverses, (size_t)total, s);
            D.cl_off = rs_upload(R, R_CL_OFF, cl_off.data(), cl_off.size(), s);
        }
        if (coarse) D.cinv = rs_upload(R, R_CINV, sys.coarse_inverse, 36 * (size_t)cm * cm, s);
        const size_t n6 = 6 * (size_t)n;
        D.x = (double*)rs_buf(R, R_X, n6 * sizeof(double));
        D.r = (double*)rs_buf(R, R_R, n6 * sizeof(double));
        D.z = (double*)rs_buf(R, R_Z, n6 * sizeof(double));
        D.p = (double*)rs_buf(R, R_P, n6 * sizeof(double));
        D.ap = (double*)rs_buf(R, R_AP, n6 * sizeof(double));
        D.minv = (double*)rs_buf(R, R_MINV, 36 * (size_t)n * sizeof(double));
        D.rc = (double*)rs_buf(R, R_RC, 6 * (size_t)std::max(cm, 1) * sizeof(double));
        D.yc = (double*)rs_buf(R, R_YC, 6 * (size_t)std::max(cm, 1) * sizeof(double));
        const int wg_vec = (int)((n6 + kRsThreads - 1) / kRsThreads);          // thread per component
        const int wg_node = (int)((n + kRsThreads / 64 - 1) / (kRsThreads / 64));  // wave per node
        const int wg_cl = D.num_clusters;
        const int wg_coarse = (int)((6 * (size_t)cm + kRsThreads - 1) / kRsThreads);
        const int num_part = std::max({wg_vec, xcd_grid(wg_node), wg_cl, 1});
        D.part_a = (double*)rs_buf(R, R_PART_A, 2 * (size_t)num_part * sizeof(double));
        D.part_b = (double*)rs_buf(R, R_PART_B, 2 * (size_t)num_part * sizeof(double));
        D.st = (RsState*)rs_buf(R, R_STATE, sizeof(RsState));
        rs_check_hip(hipMemsetAsync(D.st, 0, sizeof(RsState), s), "memset");
        // partials of the local preconditioner's r.z: one per work-group of
        // the kernel that finishes z
        const int zpart = coarse ? wg_vec : (clusters ? wg_cl : wg_vec);

        auto precond = [&](int it, int check_parts) {
            if (clusters)
                hipLaunchKernelGGL(k_rs_cluster, dim3(wg_cl), dim3(kRsThreads), 0, s, D, it, check_parts, (int)!coarse);
            else
                hipLaunchKernelGGL(k_rs_jacobi, dim3(wg_vec), dim3(kRsThreads), 0, s, D, it, check_parts, (int)!coarse);
            if (coarse) {
                hipLaunchKernelGGL(k_rs_restrict, dim3(wg_coarse), dim3(kRsThreads), 0, s, D);
                hipLaunchKernelGGL(k_rs_coarse, dim3(wg_coarse), dim3(kRsThreads), 0, s, D);
                hipLaunchKernelGGL(k_rs_prolong, dim3(wg_vec), dim3(kRsThreads), 0, s, D);
            }
            hipLaunchKernelGGL(k_rs_direction, dim3(wg_vec), dim3(kRsThreads), 0, s, D, it, zpart);
        };
        if (E)
            hipLaunchKernelGGL(k_rs_gather_in, dim3((unsigned)((36 * (size_t)E + kRsThreads - 1) / kRsThreads)),
                               dim3(kRsThreads), 0, s, D);
        hipLaunchKernelGGL(k_rs_prepare, dim3(wg_vec), dim3(kRsThreads), 0, s, D);
        hipLaunchKernelGGL(k_rs_start, dim3(1), dim3(64), 0, s, D, wg_vec, sys.tol, (int)sys.max_iterations);
        precond(-1, 0);  // z = M^-1 b, p = z, r.z

        if (!ctx.host_pcg)  // pinned staging shared with the LM driver's PCG reads
            rs_check_hip(hipHostMalloc((void**)&ctx.host_pcg, sizeof(PcgState), hipHostMallocDefault), "hipHostMalloc");
        static_assert(sizeof(RsState) <= sizeof(PcgState), "state staging");
        RsState* hs = (RsState*)ctx.host_pcg;
        const int chunk = 10;
        int it = 0;
        while (true) {
            rs_check_hip(hipMemcpyAsync(hs, D.st, sizeof(RsState), hipMemcpyDeviceToHost, s), "D2H");
            rs_check_hip(hipStreamSynchronize(s), "sync");
            if (hs->stop) break;
            if (it >= (int)sys.max_iterations) throw RsFail{VBA_ERR_DEVICE, "reduced PCG did not stop"};
            const int end = std::min<int>(it + chunk, (int)sys.max_iterations);
            for (; it < end; ++it) {
                hipLaunchKernelGGL(k_rs_matvec, dim3(xcd_grid(wg_node)), dim3(kRsThreads), 0, s, D);
                hipLaunchKernelGGL(k_rs_update, dim3(wg_vec), dim3(kRsThreads), 0, s, D, it, xcd_grid(wg_node));
                precond(it, wg_vec);
            }
            rs_check_hip(hipGetLastError(), "kernel launch");
        }
        if (hs->status == RS_FAILED)
            return decline(msg, "reduced PCG: a diagonal block is not positive definite or p.Ap is not positive");
        if (hs->status == RS_ZERO) {
            std::memset(step, 0, n6 * sizeof(double));
            *iterations = 0;
            return VBA_OK;
        }
        rs_check_hip(hipMemcpyAsync(step, D.x, n6 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        rs_check_hip(hipStreamSynchronize(s), "sync");
        *iterations = (uint32_t)hs->iter;
        return VBA_OK;
    } catch (const RsFail& f) {
        *msg = f.msg;
        return f.code;
    }
}

int32_t reduced_eliminate(Context& ctx, const vba_point_elimination& el, double* run_blocks, double* coarse_sums,
                          std::string* msg) {
    const uint32_t n = el.num_nodes, E = el.num_edges, NP = el.num_points;
    if (n == 0 || n > (1u << 24) || E >= (1u << 31) || el.structure_id == 0)
        return decline(msg, "point elimination: no nodes, too large, or structure_id 0");
    if ((E && (!el.edges || !el.row_off)) || !el.slot_off || (NP && (!el.point_off || !el.has_inverse)) ||
        (el.num_extra && (!el.extra_edge || !el.extra_blocks)) || (el.num_run_edges && (!el.run_edges || !run_blocks)) ||
        (el.coarse_group && !coarse_sums))
        return decline(msg, "point elimination: missing arrays");
    const uint64_t H = NP ? el.point_off[NP] : 0;
    if (H && (!el.point_nodes || !el.hcp || !el.point_inverse))
        return decline(msg, "point elimination: missing point blocks");
    if (!ctx.reduced) ctx.reduced = new ReducedState();
    ReducedState& R = *ctx.reduced;
    hipStream_t s = ctx.stream;
    const int32_t m = el.coarse_group ? (int32_t)((el.coarse_nodes + el.coarse_group - 1) / el.coarse_group) : 0;
    if (el.coarse_group && (el.coarse_nodes == 0 || el.coarse_nodes > n || m > 4096))
        return decline(msg, "point elimination: invalid coarse aggregates");
    for (uint32_t i = 0; i < el.num_run_edges; ++i)
        if (el.run_edges[i] >= E) return decline(msg, "point elimination: a run edge is out of range");
    {
        std::vector<uint8_t> seen(E, 0);
        for (uint32_t x = 0; x < el.num_extra; ++x) {
            if (el.extra_edge[x] >= E || seen[el.extra_edge[x]]) return decline(msg, "point elimination: extra edges must be distinct edges");
            seen[el.extra_edge[x]] = 1;
        }
    }
    try {
        ElDev D{};
        D.num_edges = (int32_t)E;
        D.num_slots = (int32_t)H;
        D.num_extra = (int32_t)el.num_extra;
        D.num_run_edges = (int32_t)el.num_run_edges;
        D.coarse_group = (int32_t)el.coarse_group;
        D.coarse_nodes = (int32_t)el.coarse_nodes;
        D.coarse_m = m;
        if (el.structure_id != R.elim_structure || R.structure_id != el.structure_id || R.n != (int32_t)n ||
            R.num_edges != (int32_t)E) {
            // validate the structure and build the pair index
            if (el.row_off[0] != 0 || el.row_off[n] != E) return decline(msg, "point elimination: row_off");
            for (uint32_t a = 0; a < n; ++a) {
                if (el.row_off[a] > el.row_off[a + 1]) return decline(msg, "point elimination: row_off");
                for (uint64_t e = el.row_off[a]; e < el.row_off[a + 1]; ++e) {
                    const uint32_t x = el.edges[2 * e], y = el.edges[2 * e + 1];
                    if (x != a || y <= a || y >= n || (e > el.row_off[a] && el.edges[2 * e - 1] >= y))
                        return decline(msg, "point elimination: edges are not sorted (s, t), s < t < n, by row");
                }
            }
            if (el.point_off[0] != 0 || H >= (1ull << 32)) return decline(msg, "point elimination: point_off");
            std::vector<int64_t> term_ptr((size_t)E + 1, 0);
            std::vector<uint32_t> pair_edge;
            std::vector<uint32_t> point_of(H);
            for (uint32_t p = 0; p < NP; ++p) {
                const uint64_t po = el.point_off[p], pe = el.point_off[p + 1];
                if (pe < po || pe > H) return decline(msg, "point elimination: point_off");
                for (uint64_t k = po; k < pe; ++k) {
                    point_of[k] = p;
                    if (el.point_nodes[k] >= n || (k > po && el.point_nodes[k - 1] >= el.point_nodes[k]))
                        return decline(msg, "point elimination: a point's nodes are not ascending below num_nodes");
                }
                for (uint64_t k = po; k < pe; ++k) {
                    const uint32_t a = el.point_nodes[k];
                    uint64_t lo = el.row_off[a];
                    const uint64_t hi = el.row_off[a + 1];
                    for (uint64_t j = k + 1; j < pe; ++j) {
                        const uint32_t b = el.point_nodes[j];
                        // partners ascend, so the search continues from the last edge
                        uint64_t l2 = lo, h2 = hi;
                        while (l2 < h2) {
                            const uint64_t mid = (l2 + h2) / 2;
                            if (el.edges[2 * mid + 1] < b) l2 = mid + 1;
                            else h2 = mid;
                        }
                        if (l2 >= hi || el.edges[2 * l2 + 1] != b)
                            return decline(msg, "point elimination: a pair of a point's nodes is not an edge");
                        lo = l2;
                        pair_edge.push_back((uint32_t)l2);
                        ++term_ptr[l2 + 1];
                    }
                }
            }
            for (uint32_t e = 0; e < E; ++e) term_ptr[e + 1] += term_ptr[e];
            std::vector<uint32_t> terms(3 * pair_edge.size());
            {
                std::vector<int64_t> cur(term_ptr.begin(), term_ptr.end() - 1);
                size_t q = 0;
                for (uint32_t p = 0; p < NP; ++p) {
                    const uint64_t po = el.point_off[p], pe = el.point_off[p + 1];
                    for (uint64_t k = po; k < pe; ++k)
                        for (uint64_t j = k + 1; j < pe; ++j) {
                            const int64_t t = cur[pair_edge[q++]]++;
                            terms[3 * t] = (uint32_t)k;
                            terms[3 * t + 1] = (uint32_t)j;
                            terms[3 * t + 2] = p;
                        }
                }
            }
            // the CPU elimination's point shards (elimination_shard_count,
            // balanced_shard_boundaries, rig_ba_solver.rs:1903-1928, 2902-2925):
            // a pure function of the edge count and the track lengths
            const uint64_t per_shard = (uint64_t)std::max<uint32_t>(E, 1) * 288;
            const uint64_t shards = std::min<uint64_t>(8, std::max<uint64_t>(1, (1536ull << 20) / per_shard));
            std::vector<uint64_t> prefix((size_t)NP + 1, 0);
            for (uint32_t p = 0; p < NP; ++p) {
                const uint64_t k = el.point_off[p + 1] - el.point_off[p];
                prefix[p + 1] = prefix[p] + k * (k > 0 ? k - 1 : 0) / 2;
            }
            std::vector<uint8_t> shard(std::max<uint32_t>(NP, 1), 0);
            size_t prev = 0;
            for (uint64_t sh = 1; sh <= shards; ++sh) {
                size_t idx = NP;
                if (sh < shards) {
                    const uint64_t target = prefix[NP] * sh / shards;
                    idx = (size_t)(std::partition_point(prefix.begin(), prefix.end(),
                                                        [&](uint64_t v) { return v < target; }) -
                                   prefix.begin());
                    idx = std::min<size_t>(std::max(idx, prev), NP);
                }
                for (size_t p = prev; p < idx; ++p) shard[p] = (uint8_t)(sh - 1);
                prev = idx;
            }
            ensure_structure(R, el.structure_id, n, E, el.edges, s);
            rs_upload(R, R_EL_TERM_PTR, term_ptr.data(), term_ptr.size(), s);
            rs_upload(R, R_EL_TERMS, terms.data(), terms.size(), s);
            rs_upload(R, R_EL_SHARD, shard.data(), shard.size(), s);
            rs_upload(R, R_EL_POINT_OF, point_of.data(), point_of.size(), s);
            rs_check_hip(hipStreamSynchronize(s), "sync");  // the host vectors go out of scope
            R.elim_structure = el.structure_id;
        }
        D.term_ptr = (const int64_t*)R.bufs[R_EL_TERM_PTR];
        D.terms = (const uint32_t*)R.bufs[R_EL_TERMS];
        D.shard = (const uint8_t*)R.bufs[R_EL_SHARD];
        D.point_of = (const uint32_t*)R.bufs[R_EL_POINT_OF];
        D.edges = (const uint32_t*)R.bufs[R_EDGES];
        D.out_ptr = (const int32_t*)R.bufs[R_OUT_PTR];
        D.hcp = rs_upload(R, R_EL_HCP, el.hcp, 18 * (size_t)H, s);
        D.vinv = rs_upload(R, R_EL_VINV, el.point_inverse, 9 * (size_t)NP, s);
        D.has = rs_upload(R, R_EL_HAS, el.has_inverse, (size_t)NP, s);
        D.scaled = (double*)rs_buf(R, R_EL_SCALED, 18 * (size_t)H * sizeof(double));
        D.b64 = (double*)rs_buf(R, R_EL_B64, 36 * (size_t)E * sizeof(double));
        D.off = (float*)rs_buf(R, R_OFF, 36 * (size_t)E * sizeof(float));
        D.extra_edge = rs_upload(R, R_EL_EXTRA_E, el.extra_edge, el.num_extra, s);
        D.extra_blocks = rs_upload(R, R_EL_EXTRA_B, el.extra_blocks, 36 * (size_t)el.num_extra, s);
        D.run_edges = rs_upload(R, R_EL_RUN_E, el.run_edges, el.num_run_edges, s);
        D.run_out = (double*)rs_buf(R, R_EL_RUN_OUT, 36 * (size_t)el.num_run_edges * sizeof(double));
        D.coarse_out = (double*)rs_buf(R, R_EL_COARSE, 36 * (size_t)m * m * sizeof(double));
        R.blocks_of = 0;
        auto grid = [](size_t work) { return dim3((unsigned)std::max<size_t>(1, (work + kRsThreads - 1) / kRsThreads)); };
        if (H) hipLaunchKernelGGL(k_el_scaled, grid(H), dim3(kRsThreads), 0, s, D);
        if (E) {
            const int wgs = (int)((E + kRsThreads / 64 - 1) / (kRsThreads / 64));
            hipLaunchKernelGGL(k_el_blocks, dim3(xcd_grid(wgs)), dim3(kRsThreads), 0, s, D);
        }
        if (el.num_extra) hipLaunchKernelGGL(k_el_extra, grid(36 * (size_t)el.num_extra), dim3(kRsThreads), 0, s, D);
        hipLaunchKernelGGL(k_el_store, grid(36 * std::max<size_t>(E, el.num_run_edges)), dim3(kRsThreads), 0, s, D);
        if (m) hipLaunchKernelGGL(k_el_coarse, grid((size_t)m * m), dim3(kRsThreads), 0, s, D);
        rs_check_hip(hipGetLastError(), "kernel launch");
        if (el.num_run_edges)
            rs_check_hip(hipMemcpyAsync(run_blocks, D.run_out, 36 * (size_t)el.num_run_edges * sizeof(double),
                                        hipMemcpyDeviceToHost, s), "D2H");
        if (m)
            rs_check_hip(hipMemcpyAsync(coarse_sums, D.coarse_out, 36 * (size_t)m * m * sizeof(double),
                                        hipMemcpyDeviceToHost, s), "D2H");
        rs_check_hip(hipStreamSynchronize(s), "sync");
        R.blocks_of = el.structure_id;
        return VBA_OK;
    } catch (const RsFail& f) {
        *msg = f.msg;
        return f.code;
    }
}

}  // namespace vba
