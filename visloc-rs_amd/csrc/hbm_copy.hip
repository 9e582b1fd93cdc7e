// hbm_copy.hip — a plain device-to-device copy of known size.
//
// Calibrates hardware counters: under `rocprofv3 --pmc FETCH_SIZE` (then
// WRITE_SIZE) the counter values of this kernel for a copy of B bytes give
// the factor that turns the counters of the solver kernels into HBM bytes.
// Also prints the copy bandwidth measured with HIP events.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void hbm_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                       size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

#define CHECK(x)                                                                        \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));         \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 10;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *src = nullptr, *dst = nullptr;
    CHECK(hipMalloc(&src, bytes));
    CHECK(hipMalloc(&dst, bytes));
    CHECK(hipMemset(src, 1, bytes));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    const int grid = 256 * 32;
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(grid), dim3(256), 0, 0, src, dst, n);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a, 0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(hbm_copy_kernel, dim3(grid), dim3(256), 0, 0, src, dst, n);
    CHECK(hipEventRecord(b, 0));
    CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, a, b));
    const double sec = ms * 1e-3 / reps;
    std::printf("hbm_copy: %zu MiB read + %zu MiB written per launch, %.1f us, %.1f GB/s (read+write)\n", mib, mib,
                sec * 1e6, 2.0 * bytes / sec / 1e9);
    CHECK(hipFree(src));
    CHECK(hipFree(dst));
    return 0;
}
