// Issue cost of single VALU instruction forms on the GPU it runs on.
//
// Each kernel runs NC independent in-place chains of one instruction form per
// lane (one asm statement per instruction, chains interleaved so no
// instruction reads the result of the one before it), over a grid of 32
// waves per CU. The time per wave64 instruction per SIMD is printed beside
// its ratio to the fastest FP32 FMA form. Only register arithmetic: no loads,
// and every store goes to the kernel's own output element.
//
// Usage: valu_classes [iters]   (prints one line per form; exit 0)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {
constexpr int NC = 12;

#define VALU_FORM(NAME, ASM)                                                                   \
    __global__ __launch_bounds__(256) void NAME(float* out, int iters, float s, unsigned long long m) { \
        float v[NC];                                                                           \
        for (int c = 0; c < NC; ++c) v[c] = 1.0f + 1.0e-3f * (float)(threadIdx.x + c);         \
        const float w = s + 1.0e-9f * (float)(threadIdx.x & 7);                                \
        for (int it = 0; it < iters; ++it) {                                                   \
            _Pragma("unroll") for (int c = 0; c < NC; ++c) asm volatile(ASM : "+v"(v[c]) : "v"(w), "s"(s), "s"(m)); \
        }                                                                                      \
        float acc = 0.f;                                                                       \
        for (int c = 0; c < NC; ++c) acc += v[c];                                              \
        out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;                              \
    }

// %0: the chain, %1: a per-lane VGPR, %2: a wave-uniform SGPR, %3: an SGPR-pair lane mask
VALU_FORM(f_fma_vv, "v_fma_f32 %0, %0, %1, 0.5")
VALU_FORM(f_fma_vs, "v_fma_f32 %0, %0, %2, 0.5")
VALU_FORM(f_add_vv, "v_add_f32 %0, %1, %0")
VALU_FORM(f_add_vs, "v_add_f32 %0, %2, %0")
VALU_FORM(f_mul_vv, "v_mul_f32 %0, %1, %0")
VALU_FORM(f_addu_vv, "v_add_u32 %0, %1, %0")
VALU_FORM(f_addu_vs, "v_add_u32 %0, %2, %0")
VALU_FORM(f_and_vv, "v_and_b32 %0, %1, %0")
VALU_FORM(f_lshr, "v_lshrrev_b32 %0, 4, %0")
VALU_FORM(f_lshl, "v_lshlrev_b32 %0, 4, %0")
VALU_FORM(f_max, "v_max_f32 %0, %1, %0")
VALU_FORM(f_cnd, "v_cndmask_b32_e64 %0, %1, %0, %3")
VALU_FORM(f_cvt, "v_cvt_f32_f16 %0, %0")
VALU_FORM(f_pkfma, "v_pk_fma_f16 %0, %0, %1, %0")
VALU_FORM(f_pkmul, "v_pk_mul_f16 %0, %1, %0")
VALU_FORM(f_pkadd, "v_pk_add_f16 %0, %1, %0")
VALU_FORM(f_addf16, "v_add_f16 %0, %1, %0")
VALU_FORM(f_mix, "v_fma_mix_f32 %0, %0, %1, %0 op_sel_hi:[1,0,0]")
VALU_FORM(f_dot2, "v_dot2c_f32_f16 %0, %1, %0")
VALU_FORM(f_rcp32, "v_rcp_f32 %0, %0")
VALU_FORM(f_rcp16, "v_rcp_f16 %0, %0")
VALU_FORM(f_rsq16, "v_rsq_f16 %0, %0")

typedef void (*Form)(float*, int, float, unsigned long long);
struct Entry {
    const char* name;
    Form f;
};
}  // namespace

int main(int argc, char** argv) {
    const int iters = argc > 1 ? atoi(argv[1]) : 1000;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 2;
    const int blocks = prop.multiProcessorCount * 8;  // 32 waves per CU
    float* out = nullptr;
    if (hipMalloc(&out, sizeof(float) * (size_t)blocks * 256) != hipSuccess) return 2;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 2;
    const Entry forms[] = {
        {"v_fma_f32 (VGPR operands)", f_fma_vv}, {"v_fma_f32 (one SGPR operand)", f_fma_vs},
        {"v_add_f32 (VGPR)", f_add_vv}, {"v_add_f32 (SGPR)", f_add_vs}, {"v_mul_f32 (VGPR)", f_mul_vv},
        {"v_add_u32 (VGPR)", f_addu_vv}, {"v_add_u32 (SGPR)", f_addu_vs}, {"v_and_b32 (VGPR)", f_and_vv},
        {"v_lshrrev_b32", f_lshr}, {"v_lshlrev_b32", f_lshl}, {"v_max_f32", f_max},
        {"v_cndmask_b32 (SGPR mask)", f_cnd}, {"v_cvt_f32_f16", f_cvt}, {"v_pk_fma_f16", f_pkfma},
        {"v_pk_mul_f16", f_pkmul}, {"v_pk_add_f16", f_pkadd}, {"v_add_f16", f_addf16},
        {"v_fma_mix_f32", f_mix}, {"v_dot2c_f32_f16", f_dot2}, {"v_rcp_f32", f_rcp32},
        {"v_rcp_f16", f_rcp16}, {"v_rsq_f16", f_rsq16}};
    const int nf = sizeof(forms) / sizeof(forms[0]);
    const double simds = 4.0 * prop.multiProcessorCount;
    double base = 0.0;
    printf("%-30s %10s %8s\n", "instruction form", "ns/instr", "x fma");
    for (int i = 0; i < nf; ++i) {
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {  // the first launch warms the clocks
            if (hipEventRecord(e0, 0) != hipSuccess) return 2;
            hipLaunchKernelGGL(forms[i].f, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 0xf0f0f0f0f0f0f0f0ull);
            if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 2;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 2;
            if (rep > 0 && ms < best) best = ms;
        }
        // wave64 instructions per SIMD = waves x iters x NC / SIMDs
        const double ns = best * 1.0e6 / ((double)blocks * 4 * iters * NC / simds);
        if (i == 0) base = ns;
        printf("%-30s %10.3f %8.2f\n", forms[i].name, ns, ns / base);
    }
    (void)hipFree(out);
    return 0;
}
