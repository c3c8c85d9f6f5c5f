// HBM copy probe: a plain streaming copy of known size, used to measure the
// attainable HBM bandwidth and to calibrate hardware byte counters.
#include "common.hpp"

namespace {
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" int wcsph_copy_probe(int32_t device, int64_t bytes, int32_t reps, double* gbps) {
    return wcsph::guarded([&] {
        WCSPH_REQUIRE(bytes >= 4096 && reps > 0 && gbps, WCSPH_EINVAL, "bytes >= 4096, reps > 0, gbps non-null");
        WCSPH_HIP(hipSetDevice(device));
        const int64_t n = bytes / 16;
        float4 *a = nullptr, *b = nullptr;
        WCSPH_HIP(hipMalloc(&a, n * 16));
        WCSPH_HIP(hipMalloc(&b, n * 16));
        hipEvent_t e0, e1;
        WCSPH_HIP(hipEventCreate(&e0));
        WCSPH_HIP(hipEventCreate(&e1));
        WCSPH_HIP(hipMemset(a, 0, n * 16));
        const int blocks = 256 * 8 * 4;
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);  // warm-up
        WCSPH_HIP(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
        WCSPH_HIP(hipEventRecord(e1, 0));
        WCSPH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        WCSPH_HIP(hipEventElapsedTime(&ms, e0, e1));
        *gbps = 2.0 * n * 16 * reps / (ms * 1e-3) / 1e9;  // read + write
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(a);
        (void)hipFree(b);
    });
}
