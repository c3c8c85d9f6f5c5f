// Hardware probes.
// - HBM copy probe: a plain streaming copy of known size, used to measure the
//   attainable HBM bandwidth and to calibrate hardware byte counters.
// - VALU probe: independent FMA chains in every lane of a full grid, timed
//   with HIP events. Kind 0: FP32 FMA (v_fma_f32) with VGPR and constant
//   operands only; kind 1: packed FP16 FMA (v_pk_fma_f16, two FMAs per
//   instruction); kind 2: FP32 FMA with one wave-uniform (SGPR) operand.
//   On gfx950 kind 0 issues a wave64 instruction every 2 cycles and kinds 1
//   and 2 every 4, so the FP32 rate depends on the operand form, and packed
//   FP16 reaches the same flop rate as FP32 FMA with VGPR operands.
#include "common.hpp"

namespace {
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Device buffers and events of one probe call, released on every exit path.
struct ProbeResources {
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~ProbeResources() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (void* p : buf)
            if (p) (void)hipFree(p);
    }
};

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
constexpr int kValuChains = 12;  // independent chains per lane, enough to cover the FMA latency

// KIND 0: x = x a + 1/2 in FP32 with a in a VGPR (a per-lane copy), so the
// FMA reads only VGPRs and an inline constant; KIND 2: the same with a
// read from its SGPR; KIND 1: the packed FP16 form (two FMAs per op). The
// chains converge to 0.5 / (1 - a), so no value overflows or goes
// subnormal.
template <int KIND>
__global__ __launch_bounds__(256) void k_valu(float* __restrict__ out, int32_t iters, float a, float b) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // a per-lane value the compiler keeps in a VGPR (it differs from a by
    // less than one rounding)
    const float av = a + b * 1.0e-12f * (float)(threadIdx.x & 7);
    float sum = 0.f;
    if (KIND != 1) {
        const float m = KIND == 0 ? av : a;
        float x[kValuChains];
#pragma unroll
        for (int k = 0; k < kValuChains; ++k) x[k] = (float)(threadIdx.x + k) * 1e-3f;
        for (int32_t it = 0; it < iters; ++it) {
#pragma unroll
            for (int k = 0; k < kValuChains; ++k) x[k] = __builtin_fmaf(x[k], m, 0.5f);
        }
#pragma unroll
        for (int k = 0; k < kValuChains; ++k) sum += x[k];
    } else {
        const half2_t a2 = {(_Float16)av, (_Float16)av}, b2 = {(_Float16)0.5f, (_Float16)0.5f};
        half2_t x[kValuChains];
#pragma unroll
        for (int k = 0; k < kValuChains; ++k) x[k] = half2_t{(_Float16)(threadIdx.x * 1e-3f), (_Float16)(k * 1e-3f)};
        for (int32_t it = 0; it < iters; ++it) {
#pragma unroll
            for (int k = 0; k < kValuChains; ++k) x[k] = __builtin_elementwise_fma(x[k], a2, b2);
        }
#pragma unroll
        for (int k = 0; k < kValuChains; ++k) sum += (float)x[k].x + (float)x[k].y;
    }
    out[gid] = sum;
}
}  // namespace

extern "C" int wcsph_copy_probe(int32_t device, int64_t bytes, int32_t reps, double* gbps) {
    return wcsph::guarded([&] {
        WCSPH_REQUIRE(bytes >= 4096 && reps > 0 && gbps, WCSPH_EINVAL, "bytes >= 4096, reps > 0, gbps non-null");
        WCSPH_HIP(hipSetDevice(device));
        const int64_t n = bytes / 16;
        ProbeResources r;
        WCSPH_HIP(hipMalloc(&r.buf[0], n * 16));
        WCSPH_HIP(hipMalloc(&r.buf[1], n * 16));
        WCSPH_HIP(hipEventCreate(&r.ev[0]));
        WCSPH_HIP(hipEventCreate(&r.ev[1]));
        float4* a = static_cast<float4*>(r.buf[0]);
        float4* b = static_cast<float4*>(r.buf[1]);
        hipEvent_t e0 = r.ev[0], e1 = r.ev[1];
        WCSPH_HIP(hipMemset(a, 0, n * 16));
        const int blocks = 256 * 8 * 4;
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);  // warm-up
        WCSPH_HIP(hipEventRecord(e0, 0));
        for (int k = 0; k < reps; ++k) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
        WCSPH_HIP(hipEventRecord(e1, 0));
        WCSPH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        WCSPH_HIP(hipEventElapsedTime(&ms, e0, e1));
        *gbps = 2.0 * n * 16 * reps / (ms * 1e-3) / 1e9;  // read + write
    });
}

extern "C" int wcsph_valu_probe(int32_t device, int32_t kind, int32_t reps, double* tflops) {
    return wcsph::guarded([&] {
        WCSPH_REQUIRE(kind >= 0 && kind <= 2 && reps > 0 && tflops, WCSPH_EINVAL,
                      "kind 0 (FP32 FMA), 1 (packed FP16 FMA) or 2 (FP32 FMA, SGPR operand), reps > 0, "
                      "tflops non-null");
        WCSPH_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        WCSPH_HIP(hipGetDeviceProperties(&prop, device));
        const int blocks = prop.multiProcessorCount * 8;  // 32 waves per CU
        const int32_t iters = 1 << 15;
        ProbeResources r;
        WCSPH_HIP(hipMalloc(&r.buf[0], sizeof(float) * blocks * 256));
        WCSPH_HIP(hipEventCreate(&r.ev[0]));
        WCSPH_HIP(hipEventCreate(&r.ev[1]));
        float* out = static_cast<float*>(r.buf[0]);
        hipEvent_t e0 = r.ev[0], e1 = r.ev[1];
        auto launch = [&] {
            if (kind == 0)
                hipLaunchKernelGGL(k_valu<0>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1.0f);
            else if (kind == 1)
                hipLaunchKernelGGL(k_valu<1>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1.0f);
            else
                hipLaunchKernelGGL(k_valu<2>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1.0f);
        };
        launch();  // warm-up
        float best = 1e30f;
        for (int k = 0; k < reps; ++k) {
            WCSPH_HIP(hipEventRecord(e0, 0));
            launch();
            WCSPH_HIP(hipEventRecord(e1, 0));
            WCSPH_HIP(hipEventSynchronize(e1));
            float ms = 0;
            WCSPH_HIP(hipEventElapsedTime(&ms, e0, e1));
            best = ms < best ? ms : best;
        }
        const double fmas = (double)blocks * 256 * iters * kValuChains * (kind == 1 ? 2 : 1);
        *tflops = 2.0 * fmas / (best * 1e-3) / 1e12;
    });
}
