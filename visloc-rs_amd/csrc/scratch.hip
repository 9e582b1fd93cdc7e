// scratch.hip — the dense normal-equation solve without device allocation
// (vba_dense_solve_noalloc).
//
// Every device buffer of the solve (the row-major copy of H, the LU
// fallback's copy with its right-hand-side workspace, b, x and the status
// word) lives in one file-scope device array of fixed size. Each GPU holds its
// own copy of the array from the library's load until the process ends, and
// every call on that GPU shares it: calls that overlap on one GPU (from two
// contexts, or two threads) race on its contents.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/vba.h"
#include "kernels.hpp"
#include "solver.hpp"

namespace vba {
namespace {

constexpr uint32_t kMaxN = VBA_DENSE_NOALLOC_MAX_N;
constexpr size_t kAlign = 32;  // region starts on 256-byte boundaries (32 doubles)
constexpr size_t round_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Regions for a system of n unknowns, in doubles from the array's start.
struct Regions {
    size_t s, work, b, x, status, end;
};
constexpr Regions regions(size_t n) {
    const size_t s = 0;
    const size_t work = round_up(s + n * n);
    const size_t b = round_up(work + n * n + n);
    const size_t x = round_up(b + n);
    const size_t status = round_up(x + n);
    return Regions{s, work, b, x, status, status + 1};
}
constexpr size_t kScratchDoubles = regions(kMaxN).end;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

// Column-major H (the caller's layout) to the row-major copy the dense
// kernels factor.
__global__ void k_transpose_square(const double* __restrict__ col, double* __restrict__ row, int n) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nn = (size_t)n * n;
    if (k >= nn) return;
    const size_t r = k / n, c = k % n;
    row[k] = col[c * n + r];
}

}  // namespace

// The device address of this device's array, looked up once per device with
// the device current. hipGetSymbolAddress's result is kept with its error.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_done[device]) {
        void* p = nullptr;
        g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[device] = g_err[device] == hipSuccess ? static_cast<double*>(p) : nullptr;
        g_done[device] = true;
    }
    return g_err[device];
}

int32_t dense_solve_noalloc(Context& ctx, uint32_t n, const double* h, const double* b, double* x,
                            std::string* msg) {
    if (n == 0) return VBA_OK;
    if (n > kMaxN) {
        *msg = "vba_dense_solve_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
               std::to_string(kMaxN) + " unknowns";
        return VBA_ERR_INVALID_ARGUMENT;
    }
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (ctx.device >= 0 && ctx.device < kMaxDevices && g_done[ctx.device]) base = g_ptr[ctx.device];
    }
    if (!base) {
        *msg = "vba_dense_solve_noalloc: the device scratch address was not found when the context was created";
        return VBA_ERR_DEVICE;
    }
    if (!ctx.host_status) {
        *msg = "vba_dense_solve_noalloc: the context has no pinned status word";
        return VBA_ERR_DEVICE;
    }
    const Regions rg = regions(n);
    double* dS = base + rg.s;
    double* dwork = base + rg.work;
    double* db = base + rg.b;
    double* dx = base + rg.x;
    int32_t* dstatus = reinterpret_cast<int32_t*>(base + rg.status);
    hipStream_t s = ctx.stream;
    const size_t nn = (size_t)n * n;
    auto fail = [&](hipError_t e, const char* what) {
        *msg = std::string("vba_dense_solve_noalloc: ") + what + ": " + hipGetErrorString(e);
        (void)hipStreamSynchronize(s);
        return VBA_ERR_DEVICE;
    };
    hipError_t e;
    // H arrives column-major in the work region; the transpose writes the
    // row-major S, and the solve launcher then copies S back over the work
    // region as the LU fallback's saved copy.
    if ((e = hipMemcpyAsync(dwork, h, nn * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess)
        return fail(e, "upload of H");
    if ((e = hipMemcpyAsync(db, b, n * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess)
        return fail(e, "upload of b");
    const int threads = 256;
    hipLaunchKernelGGL(k_transpose_square, dim3((unsigned)((nn + threads - 1) / threads)), dim3(threads), 0, s,
                       dwork, dS, (int)n);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "transpose");
    if ((e = launch_dense_solve_plain(dS, dwork, db, dx, (int32_t)n, dstatus, s)) != hipSuccess)
        return fail(e, "dense solve");
    if ((e = hipMemcpyAsync(ctx.host_status, dstatus, sizeof(int32_t), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return fail(e, "status read");
    if ((e = hipMemcpyAsync(x, dx, n * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return fail(e, "download of x");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e, "sync");
    if (*ctx.host_status != 0) {
        *msg = n <= (uint32_t)kLuFallbackMax
                   ? "singular system: neither Cholesky nor LU with partial pivoting solves it"
                   : "singular system: Cholesky fails (no LU fallback beyond 3072 unknowns on the device)";
        return VBA_ERR_SINGULAR_SYSTEM;
    }
    return VBA_OK;
}

}  // namespace vba
