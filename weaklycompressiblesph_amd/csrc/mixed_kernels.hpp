// Mixed-precision fused neighbour + force pass (uniform smoothing length).
//
// Units: lengths in h, velocities in c0, densities in rho0, P/rho^2 in
// c0^2/rho0, masses in rho0 h^D (times 3 C_D/2^(D+1): the kernel constant
// and the 3 of the spline's derivative).
// In these units every quantity of the pair body is O(1e-4..50) and fits
// FP16 (in SI units 1/H^4 alone would overflow it).
//
// Pair records: particles 2m and 2m+1 (cell order) share one record of
// each kind, the FP16 ones as the two halves of packed half2 values:
//   PosRec16   {x, y, z, x-cell}   position relative to the cell centre / h
//                                  (|x| <= L/2h, where FP16 has its finest
//                                  steps) and the cell's x index
//   VelRec16   {vx, vy, vz, m}     v / c0 and the scaled mass, negated for a
//                                  wall particle (the pair body masks the
//                                  sign off with the hit bits)
//   StateRec16 {rho[2], pr[2]}     FP32 rho / rho0 and (P / rho^2)(rho0 / c0^2)
//   ch[m]                          half2 -alpha c / c0, an array of its own:
//                                  the artificial viscosity's -2 alpha times
//                                  c / (2 c0), so ch_i + ch_j is -2 alpha c_bar / c0
// The three 16-byte records share one offset (m * 16) per candidate.
// One 16-byte load therefore brings two candidates' positions, and the
// geometric part of the pair body -- differences, r^2, dr.dv, the spline
// factor -- is one packed FP16 VALU op (v_pk_fma_f16, v_pk_mul_f16 ...) for
// two pairs. Pair differences are taken in FP16 from the h-scaled cell-local
// coordinates, the small difference first and the whole-cell offset last:
// dx = (x_i - x_j) + (cx_i - cx_j) L/h, dy = (y_i - y_j) - dy_row L/h, so only
// the last rounding is at the scale of the result. The scan's predicate uses
// the cheaper per-row form (y_i - dy_row L/h) - y_j. Density and pressure
// stay FP32: 1/(rho_i + rho_j) and P_i/rho_i^2 + P_j/rho_j^2 + Pi are formed
// per pair in FP32 from the FP32 records, and so are the two pair
// coefficients they enter (pressure and viscous). The per-particle sums
// accumulate those FP32 coefficients times the FP16 differences in FP32
// (v_fma_mix_f32), and div v through v_dot2_f32_f16.
//
// Scan: per row of cells each lane walks the records of its chord-clipped
// window; a record with at least one neighbour half goes onto the lane's
// LDS stack as (record | row << 25 | hit bits << 30). The top K entries of
// every lane are evaluated together, so every lane of the wave has work for
// most of a drain. The neighbour set is this path's FP16 predicate
// 0 < r^2 < 4 h^2: it differs from the FP32 path's only for pairs within
// ~1e-3 h of the support edge, whose kernel weight is ~1e-6.
#pragma once
#include "sim_kernels.hpp"

namespace wcsph {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));

struct alignas(16) PosRec16 {
    h2 x, y, z, cx;
};
struct alignas(16) VelRec16 {
    h2 vx, vy, vz, m;
};
struct alignas(16) StateRec16 {
    float rho[2], pr[2];
};
// the three records of a pair share one byte offset
static_assert(sizeof(PosRec16) == 16 && sizeof(VelRec16) == 16 && sizeof(StateRec16) == 16, "record size");

struct Mixed {
    PosRec16* pos;  // [(n + 1) / 2]
    VelRec16* vel;
    StateRec16* st;
    h2* ch;
};

__host__ __device__ inline int64_t mixed_records(int64_t n) { return (n + 1) / 2; }

__device__ __forceinline__ h2 as_h2(uint32_t u) { return __builtin_bit_cast(h2, u); }
__device__ __forceinline__ uint32_t as_u32(h2 v) { return __builtin_bit_cast(uint32_t, v); }

struct MixedConsts {
    float inv_h, ch_scale, inv_rho0, pr_scale, m_scale;  // SI -> scaled units (ch_scale = -alpha / c0)
    float half_Ls;  // half the cell edge in h: positions are stored relative to the cell centre
};

// The state of particle p in its records (after the Shepard filter changes
// the density): thread p owns half p & 1 of record p >> 1.
__device__ __forceinline__ void write_mixed_state(const Mixed& mx, const MixedConsts& mc, int64_t p, double rho,
                                                  double pr, double c) {
    const int o = (int)(p & 1);
    StateRec16& S = mx.st[p >> 1];
    S.rho[o] = (float)(rho * mc.inv_rho0);
    S.pr[o] = (float)(pr * mc.pr_scale);
    reinterpret_cast<_Float16*>(&mx.ch[p >> 1])[o] = (_Float16)(float)(c * mc.ch_scale);
}

// The pair records written whole by k_reorder: lanes p and p ^ 1 (one wave)
// swap their values and the even lane stores the records. A lane past the
// end supplies zeros. Every lane of the wave must call it.
template <int D>
__device__ __forceinline__ void write_mixed_pair(const Mixed& mx, const MixedConsts& mc, int64_t p, bool live,
                                                 const float* xl, int cx, const float* vf, double rho, double pr,
                                                 double c, double m, bool wall) {
    uint32_t mine[4] = {0u, 0u, 0u, 0u};
    uint32_t fs[3] = {0u, 0u, 0u};  // rho, pr (FP32 bits) and c/2 (FP16 bits)
    if (live) {
        const h2 v[4] = {
            h2{(_Float16)(xl[0] * mc.inv_h - mc.half_Ls), (_Float16)(xl[1] * mc.inv_h - mc.half_Ls)},
            h2{(_Float16)(D == 3 ? xl[2] * mc.inv_h - mc.half_Ls : 0.f), (_Float16)(float)cx},
            h2{(_Float16)vf[0], (_Float16)vf[1]},
            h2{(_Float16)(D == 3 ? vf[2] : 0.f), (_Float16)(float)(wall ? -(m * mc.m_scale) : m * mc.m_scale)}};
#pragma unroll
        for (int k = 0; k < 4; ++k) mine[k] = as_u32(v[k]);
        fs[0] = __builtin_bit_cast(uint32_t, (float)(rho * mc.inv_rho0));
        fs[1] = __builtin_bit_cast(uint32_t, (float)(pr * mc.pr_scale));
        fs[2] = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(float)(c * mc.ch_scale));
    }
    uint32_t other[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) other[k] = (uint32_t)__shfl_xor((int)mine[k], 1, 64);
    uint32_t ofs[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) ofs[k] = (uint32_t)__shfl_xor((int)fs[k], 1, 64);
    if ((p & 1) == 0 && live) {
        // record half 0 = this lane (even p), half 1 = the partner
        auto lo = [](uint32_t a, uint32_t b) { return h2{as_h2(a).x, as_h2(b).x}; };
        auto hi = [](uint32_t a, uint32_t b) { return h2{as_h2(a).y, as_h2(b).y}; };
        mx.pos[p >> 1] = PosRec16{lo(mine[0], other[0]), hi(mine[0], other[0]), lo(mine[1], other[1]),
                                  hi(mine[1], other[1])};
        mx.vel[p >> 1] = VelRec16{lo(mine[2], other[2]), hi(mine[2], other[2]), lo(mine[3], other[3]),
                                  hi(mine[3], other[3])};
        StateRec16 S;
        S.rho[0] = __builtin_bit_cast(float, fs[0]);
        S.rho[1] = __builtin_bit_cast(float, ofs[0]);
        S.pr[0] = __builtin_bit_cast(float, fs[1]);
        S.pr[1] = __builtin_bit_cast(float, ofs[1]);
        mx.st[p >> 1] = S;
        mx.ch[p >> 1] = as_h2(fs[2] | (ofs[2] << 16));
    }
}

template <int D>
struct MixedExtra {  // k_reorder / k_apply_density hook writing the pair records
    static constexpr bool kFp32Records = false;  // the mixed pass reads only its pair records and the master state
    Mixed mx;
    MixedConsts mc;
    __device__ void store_pair(int64_t p, bool live, const float* xl, int cx, const float* vf, double rho,
                               double pr, double c, double m, bool wall) const {
        write_mixed_pair<D>(mx, mc, p, live, xl, cx, vf, rho, pr, c, m, wall);
    }
    __device__ void state(int64_t p, double rho, double pr, double c) const {
        write_mixed_state(mx, mc, p, rho, pr, c);
    }
};

__device__ __forceinline__ h2 hsplat(float v) {
    const _Float16 h = (_Float16)v;
    return h2{h, h};
}
// FP16 transcendentals of both halves into one register: the low half by
// the plain op, the high half by its SDWA form writing word 1 in place (no
// repacking); each is followed by the one wait state a read of a
// transcendental's result needs.
__device__ __forceinline__ h2 hrcp(h2 v) {
    uint32_t r;
    asm("v_rcp_f16_e32 %0, %1\n\ts_nop 0\n\tv_rcp_f16_sdwa %0, %1 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:WORD_1\n\ts_nop 0"
        : "=&v"(r) : "v"(as_u32(v)));
    return as_h2(r);
}
__device__ __forceinline__ h2 hrsq(h2 v) {
    uint32_t r;
    asm("v_rsq_f16_e32 %0, %1\n\ts_nop 0\n\tv_rsq_f16_sdwa %0, %1 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:WORD_1\n\ts_nop 0"
        : "=&v"(r) : "v"(as_u32(v)));
    return as_h2(r);
}
__device__ __forceinline__ h2 hmax(h2 a, h2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ h2 hmin(h2 a, h2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ h2 hfma(h2 a, h2 b, h2 c) { return __builtin_elementwise_fma(a, b, c); }

struct NFMixedArgs {
    Mixed mx;
    const double* pos;   // master FP64 positions, cell order (row windows of i)
    const double* dens;  // master FP64 densities, cell order (epilogue of i)
    const uint32_t* start;
    Grid g;
    int64_t n;
    int64_t nrec;  // pair records
    float h;
    // wave-uniform constants (kept in scalar registers)
    h2 k_eps, k_Ls;
    float beta;  // 2 beta (the -2 alpha is in the ch records)
    double inv_rho0;
    float gm1;           // gamma - 1 of the Tait law (sound speed of i in the epilogue)
    // epilogue constants, formed once on the host: the force and viscous
    // scales c0^2 / h and 4 nu c0 / h^2 (FP64), -c0 / h, c0 and C_CFL h
    // (FP32), the viscous time-step bound 0.125 h^2 / nu (FP64), and whether
    // gamma = 7, where the sound speed is c0 (rho / rho0)^3
    double ks, kv, t2;
    float kdivv, c0f, cfl_h, hf;
    int g7;
    double* acc;
    double* divv;
    double* dti;
    Partials P;
    Scalars* s;
    int fused_final;  // run the final stage (FinalArgs) in the epilogue
    FinalArgs fin;
    int wall_mode;  // frozen walls: a wave of walls alone takes fluid neighbours only
};

// Tuning constants (build-time overridable): stack entries per lane and
// entries drained per lane per drain (3D, and 2D after the 2), scan
// records in flight per lane, and the waves per SIMD the register budget
// is fitted to. In 3D, 20 entries keep a 256-thread block at 22.5 KB of
// LDS, so the 7 blocks per CU its registers allow still fit in 160 KB; the
// 2D kernel's registers allow 8, which 16 entries (18.5 KB) keep.
#ifndef WCSPH_MIX_Q
#define WCSPH_MIX_Q 20
#endif
#ifndef WCSPH_MIX_K
#define WCSPH_MIX_K 10
#endif
#ifndef WCSPH_MIX_Q2
#define WCSPH_MIX_Q2 16
#endif
#ifndef WCSPH_MIX_K2
#define WCSPH_MIX_K2 8
#endif
#ifndef WCSPH_MIX_U
#define WCSPH_MIX_U 2
#endif
#ifndef WCSPH_MIX_WAVES
#define WCSPH_MIX_WAVES 4
#endif
constexpr int kMixQ = WCSPH_MIX_Q;
constexpr int kMixK = WCSPH_MIX_K;
constexpr int kMixQ2 = WCSPH_MIX_Q2;
constexpr int kMixK2 = WCSPH_MIX_K2;
constexpr int kMixU = WCSPH_MIX_U;
constexpr int kMixRecBits = 25;  // record index bits of a stack entry
// one drain of K entries makes room for the next U pushes
static_assert(kMixU <= kMixK && kMixK <= kMixQ && kMixU <= kMixK2 && kMixK2 <= kMixQ2, "stack shape");

// Buffer resource over n bytes at p for range-checked loads: a raw buffer
// load whose offset lies beyond n returns zeros instead of touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, int32_t n) {
    constexpr int32_t kGfx9RawBuffer = 0x00020000;  // dword 3 of a raw (untyped) buffer on gfx9-family parts
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, n, kGfx9RawBuffer);
}

template <int D, int R, int VS, int BT>
__device__ __forceinline__ void nforce_mixed_body(VBlock vb, const NFMixedArgs& a) {
    constexpr int S = 2 * R + 1;
    constexpr int NR = D == 3 ? S * S : S;
    constexpr int Q = D == 3 ? kMixQ : kMixQ2, K = D == 3 ? kMixK : kMixK2, U = kMixU;
    constexpr uint32_t kRecMask = (1u << kMixRecBits) - 1u;
    static_assert(NR <= 32, "row index of a stack entry has 5 bits");
    // per (row, hit bits) of a stack entry: the row's (dy L/h, dz L/h) as
    // half2 splats, the half2 mask keeping the hit halves (sign bits off),
    // the hit count
    __shared__ uint4 etab[128];
    // slot-major stacks: slot t of lane l at [t][l], so lanes reading
    // different slots still hit different banks
    __shared__ uint32_t qbuf[kBlock / 64][Q * 64];
    if (threadIdx.x < 128) {
        const int row = (int)threadIdx.x & 31, bits = (int)threadIdx.x >> 5;
        const h2 dy = hsplat((float)(row % S - R)) * a.k_Ls;
        const h2 dz = hsplat(D == 3 ? (float)(row / S - R) : 0.f) * a.k_Ls;
        // hit halves, sign bits off: the mass's sign is the wall flag
        const uint32_t sel = ((bits & 1) ? 0x00007fffu : 0u) | ((bits & 2) ? 0x7fff0000u : 0u);
        etab[threadIdx.x] = make_uint4(as_u32(dy), as_u32(dz), sel, (uint32_t)__builtin_popcount(bits));
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // 32-bit LDS pointers, so stack moves and compares are single ops
    using lds_u32 = __attribute__((address_space(3))) uint32_t;
    lds_u32* const q = (lds_u32*)&qbuf[threadIdx.x >> 6][lane];
    // every slot at or above a lane's stack top holds zero or a hit-free
    // entry (the pushes write the top slot always, the drains clear what
    // they pop), so a drain evaluates slots above the top with no select
#pragma unroll
    for (int s = 0; s < Q; ++s) q[s * 64] = 0u;
    const int64_t i = (int64_t)xcd_block(vb) * blockDim.x + threadIdx.x;
    const bool failed = sim_failed(a.s);
    const bool live = i < a.n && !failed;
    const int64_t ii = live ? i : 0;
    // the cell and FP32 cell-local coordinates the binning pass gave i (the
    // same FP64 arithmetic on the same position)
    int c[3] = {0, 0, 0};
    float xi[3] = {0.f, 0.f, 0.f};
    {
        double x[D];
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = a.pos[(int64_t)d * a.n + ii];
        bool esc = false;
        cell_of<D>(a.g, x, c, esc);
#pragma unroll
        for (int d = 0; d < D; ++d) xi[d] = cell_local(x[d], a.g.lo[d], c[d], a.g.L);
    }
    const bool odd = (ii & 1) != 0;
    const PosRec16 Pi = a.mx.pos[ii >> 1];
    const VelRec16 Vi = a.mx.vel[ii >> 1];
    const StateRec16 Si = a.mx.st[ii >> 1];
    auto pick = [&](h2 v) {
        const _Float16 s = odd ? v.y : v.x;
        return h2{s, s};
    };
    const h2 xi2 = pick(Pi.x), yi2 = pick(Pi.y), zi2 = pick(Pi.z), cxi2 = pick(Pi.cx);
    const h2 vxi = pick(Vi.vx), vyi = pick(Vi.vy), vzi = pick(Vi.vz);
    const float rhoi = odd ? Si.rho[1] : Si.rho[0], pri = odd ? Si.pr[1] : Si.pr[0];
    const h2 chi2 = pick(a.mx.ch[ii >> 1]);
    const float beta = a.beta;
    const h2 Ls2 = a.k_Ls, eps2 = a.k_eps;
    const h2 zero2 = hsplat(0.f), tiny2 = hsplat(1.0e-6f), sixteen2 = hsplat(16.f);
    const _Float16 four = (_Float16)4.f;
    const float Lf = a.g.Lf, invLf = 1.0f / Lf;
    const float hc = 2.f * a.h * 1.0001f, hc2 = hc * hc;
    float accp[3] = {0.f, 0.f, 0.f}, accv[3] = {0.f, 0.f, 0.f}, divv = 0.f;
    h2 vsig2 = zero2;
    // the lane's stack top: the next free slot, 64 words per slot; pushes
    // write through it and move it by one slot per kept record, so a push
    // costs no address arithmetic
    lds_u32* qp = q;
    lds_u32* const qlim = q + (Q - U) * 64;  // more than this and U pushes could overflow
    uint32_t total = 0;
    // range-checked record loads: a hit-free entry at the stack top may name
    // a record past the last one, which then reads as zeros
    const int32_t rec_bytes = (int32_t)(a.nrec * (int64_t)sizeof(PosRec16));
    const __amdgpu_buffer_rsrc_t pos_rsrc = buffer_rsrc(a.mx.pos, rec_bytes);
    const __amdgpu_buffer_rsrc_t vel_rsrc = buffer_rsrc(a.mx.vel, rec_bytes);
    const __amdgpu_buffer_rsrc_t st_rsrc = buffer_rsrc(a.mx.st, rec_bytes);
    const __amdgpu_buffer_rsrc_t ch_rsrc = buffer_rsrc(a.mx.ch, (int32_t)(a.nrec * (int64_t)sizeof(h2)));

    // Two pair interactions (the halves of record m), each kept only if its
    // hit bit is set: the other half is then outside the window, beyond the
    // support or the particle itself. The geometry is packed FP16; the
    // state terms and the coefficients are FP32 per half. Every lane runs
    // this for every drain step; an entry with no hit bits (zero, or a
    // hit-free entry of any record) adds exactly zero. A record index past
    // 2^25 - 1 can only carry into the row bits (row 24 -> 25 at most),
    // never into the hit bits.
    auto body = [&](uint32_t e) {
        const uint32_t m = e & kRecMask;
        const uint4 T = etab[e >> kMixRecBits];
        const int32_t off = (int32_t)(m * (uint32_t)sizeof(PosRec16));
        const PosRec16 P = __builtin_bit_cast(PosRec16, __builtin_amdgcn_raw_buffer_load_b128(pos_rsrc, off, 0, 0));
        const VelRec16 V = __builtin_bit_cast(VelRec16, __builtin_amdgcn_raw_buffer_load_b128(vel_rsrc, off, 0, 0));
        const StateRec16 St = __builtin_bit_cast(StateRec16, __builtin_amdgcn_raw_buffer_load_b128(st_rsrc, off, 0, 0));
        const h2 chj = as_h2(__builtin_amdgcn_raw_buffer_load_b32(ch_rsrc, (int32_t)(m * (uint32_t)sizeof(h2)), 0, 0));
        total += T.w;
        const h2 dx = hfma(cxi2 - P.cx, Ls2, xi2 - P.x);
        const h2 dy = (yi2 - P.y) - as_h2(T.x);
        const h2 dz = (zi2 - P.z) - as_h2(T.y);
        const h2 dvx = vxi - V.vx, dvy = vyi - V.vy, dvz = vzi - V.vz;
        h2 r2 = dx * dx;
        r2 = hfma(dy, dy, r2);
        h2 drdv = dx * dvx;
        drdv = hfma(dy, dvy, drdv);
        if (D == 3) {
            r2 = hfma(dz, dz, r2);
            drdv = hfma(dz, dvz, drdv);
        }
        // keep r^2 in [1e-6, 16]: a masked half may be the particle itself
        // (r = 0) or one far outside the support (r^2 may overflow FP16);
        // the kernel weight is zero beyond r^2 = 4 either way
        r2 = hmin(hmax(r2, tiny2), sixteen2);
        if (VS == 1) drdv = as_h2(as_u32(drdv) & (T.z | ((T.z << 1) & 0x80008000u)));  // whole hit halves
        const h2 mj = as_h2(as_u32(V.m) & T.z);
        const h2 inv_den = hrcp(r2 + eps2);     // 1 / (r^2 + eps h_i h_j)
        const h2 mu = hmin(drdv, zero2) * inv_den;  // (h_i + h_j)/2 = 1
        if (VS == 1) vsig2 = hmax(vsig2, -mu);
        const h2 cbar = chi2 + chj;  // -2 alpha c_bar / c0
        const h2 amu = BT ? hfma(hsplat(beta), mu, cbar) * mu : cbar * mu;  // Pi (rho_i + rho_j)
        const h2 invr = hrsq(r2);
        const h2 rr = r2 * invr;               // r / h
        // the spline's derivative over q, 12 bb^2 - 3 aa^2 with aa = (1 - q/2)+
        // and bb = (1/2 - q/2)+, taken as b2^2 - aa^2 with b2 = 2 bb = (1 - q)+
        // (its 3 is in the scaled mass)
        const h2 aa = hmax(hfma(rr, hsplat(-0.5f), hsplat(1.f)), zero2);
        const h2 b2 = hmax(hsplat(1.f) - rr, zero2);
        const h2 g = hfma(b2, b2, -(aa * aa));
        const h2 mF = mj * g * invr;           // m_j F (scaled)
        const h2 mFv = mF * (r2 * inv_den);    // m_j F r^2 / den, the viscous factor before 1/(rho_i + rho_j)
        // density and pressure terms in FP32, one half at a time
        const float amus[2] = {(float)amu.x, (float)amu.y};
        const _Float16 mFh[2] = {mF.x, mF.y};
        const _Float16 mFvh[2] = {mFv.x, mFv.y};
        const float dxs[2][3] = {{(float)dx.x, (float)dy.x, (float)dz.x}, {(float)dx.y, (float)dy.y, (float)dz.y}};
        const float dvs[2][3] = {{(float)dvx.x, (float)dvy.x, (float)dvz.x}, {(float)dvx.y, (float)dvy.y, (float)dvz.y}};
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const float inv_rs = __builtin_amdgcn_rcpf(rhoi + St.rho[hh]);  // 1 / (rho_i + rho_j)
            // FP16 x FP32 products written as fma(., ., 0): one v_fma_mix_f32
            // each, equal to the product up to the sign of an exact zero
            const float cp = fmaf((float)mFh[hh], -fmaf(amus[hh], inv_rs, pri + St.pr[hh]), 0.0f);
            const float vc = fmaf((float)mFvh[hh], inv_rs, 0.0f);
#pragma unroll
            for (int d = 0; d < D; ++d) {
                accp[d] = fmaf(cp, dxs[hh][d], accp[d]);
                accv[d] = fmaf(vc, dvs[hh][d], accv[d]);
            }
        }
        divv = __builtin_amdgcn_fdot2(mF, drdv, divv, false);
    };
    auto drain = [&]() {
        // pop the top k entries and evaluate them bottom-up, clearing each
        // slot read; a lane with k < kmax reads on into slots above its top
        // (all below Q: its stack is then empty and kmax <= K <= Q)
        const uint32_t k = min((uint32_t)(qp - q) >> 6, (uint32_t)K);
        const uint32_t kmax = wave_max(k);
        qp -= k * 64;
        for (uint32_t t = 0; t < kmax; ++t) {
            const uint32_t e = qp[t * 64];
            qp[t * 64] = 0u;
            body(e);
        }
    };

    // Row r of the stencil (offsets dy, dz): the index range [b0, b1) of
    // the cells of that row inside lane i's support chord; b0 = b1 = 0 when
    // the row is outside the grid or the support. The window is widened by
    // the 1e-4 margin of hc, which covers the approximate square root.
    const uint32_t ncx = (uint32_t)a.g.nc[0], ncy = (uint32_t)a.g.nc[1], ncz = (uint32_t)a.g.nc[2];
    const uint32_t own = (uint32_t)((c[2] * (int)ncy + c[1]) * (int)ncx + c[0]);  // own cell index
    const float xi0L = xi[0] * invLf;
    const int xa_min = -c[0], xb_max = (int)ncx - 1 - c[0];
    float gz2 = 0.f;  // squared z gap of the current plane of rows
    auto row_bounds = [&](int r, uint32_t& b0, uint32_t& b1) {
        const int dy = r % S - R, dz = D == 3 ? r / S - R : 0;
        const int srow = (dz * (int)ncy + dy) * (int)ncx;  // wave-uniform
        if (D == 3 && dy == -R) {  // first row of a plane (a wave-uniform branch)
            const float gz = axis_gap(dz, xi[2], Lf, c[2] + dz, (int)ncz);
            gz2 = gz * gz;
        }
        const float gy = axis_gap(dy, xi[1], Lf, c[1] + dy, (int)ncy);
        const float g2 = fmaf(gy, gy, gz2);
        const float w = __builtin_amdgcn_sqrtf(fmaxf(hc2 - g2, 0.f)) * invLf;
        // the row's end cells are open outward (axis_gap): a chord that
        // reaches past one keeps it
        const int xa = min(max(max((int)floorf(xi0L - w), -R), xa_min), xb_max);
        const int xb = max(min(min((int)floorf(xi0L + w), R), xb_max), xa_min);
        const bool ok = live && (uint32_t)(c[1] + dy) < ncy && (D == 2 || (uint32_t)(c[2] + dz) < ncz) &&
                        g2 < hc2 && xa <= xb;
        b0 = ok ? own + srow + xa : 0u;
        b1 = ok ? own + srow + xb + 1 : 0u;
    };
    // The rows' records; with walls_only (a wave of frozen walls, below)
    // only fluid halves are kept: a wall-wall pair adds exactly zero to
    // div v (dv = 0) and to v_sig (mu = 0), and a wall's acceleration is
    // not used. The stack then holds fewer entries, so the kept terms may
    // be added in another order.
    auto scan_rows = [&](auto walls_only) {
        constexpr bool WO = decltype(walls_only)::value;
        constexpr int UW = WO ? 1 : U;  // records per trip (one in a wall wave: fewer registers)
        uint32_t nb0, nb1;
        row_bounds(0, nb0, nb1);
        uint32_t np0 = a.start[nb0], npe = a.start[nb1];
        for (int r = 0; r < NR; ++r) {
            const uint32_t p0 = np0, pend = npe;
            if (r + 1 < NR) {  // prefetch the next row's range behind this row's scan
                row_bounds(r + 1, nb0, nb1);
                np0 = a.start[nb0];
                npe = a.start[nb1];
            }
            const uint4 T = etab[r];
            const h2 ry = yi2 - as_h2(T.x), rz = zi2 - as_h2(T.y);
            // record k of the window holds particles 2 (m0 + k) and 2 (m0 + k) + 1;
            // its low half lies in [p0, pend) for k < npair except k = 0 when p0
            // is odd, its high half for k < hi_lim (an empty window starting at
            // an odd index scans one record with neither half valid)
            const uint32_t m0 = p0 >> 1;
            const uint32_t npair = ((pend + 1) >> 1) - m0;
            const uint32_t hi_lim = (pend >> 1) - m0;
            const bool even_start = (p0 & 1) == 0;
            const uint32_t ebase = m0 | ((uint32_t)r << kMixRecBits);
            const int32_t voff = (int32_t)(m0 * (uint32_t)sizeof(PosRec16));
            const uint32_t npmax = wave_max(npair);
            for (uint32_t t = 0; t < npmax; t += UW) {
                PosRec16 rec[UW];
                uint32_t tw[UW];  // the mass halves, whose signs flag walls (walls_only)
#pragma unroll
                for (int u = 0; u < UW; ++u) {
                    rec[u] = __builtin_bit_cast(
                        PosRec16, __builtin_amdgcn_raw_buffer_load_b128(
                                      pos_rsrc, voff + (int32_t)((t + u) * sizeof(PosRec16)), 0, 0));
                    tw[u] = WO ? __builtin_amdgcn_raw_buffer_load_b32(
                                     vel_rsrc, voff + (int32_t)((t + u) * sizeof(VelRec16) + 12), 0, 0)
                               : 0u;
                }
#pragma unroll
                for (int u = 0; u < UW; ++u) {
                    const uint32_t k = t + u;
                    const h2 dx = hfma(cxi2 - rec[u].cx, Ls2, xi2) - rec[u].x;
                    const h2 dy = ry - rec[u].y;
                    const h2 dz = rz - rec[u].z;
                    h2 r2 = dx * dx;
                    r2 = hfma(dy, dy, r2);
                    if (D == 3) r2 = hfma(dz, dz, r2);
                    const bool lo = k < npair && (k != 0 || even_start) && r2.x < four &&
                                    (!WO || (tw[u] & 0x8000u) == 0u);
                    const bool hi = k < hi_lim && r2.y < four && (!WO || (tw[u] >> 31) == 0u);
                    // the top slot is free (at most Q - U entries before this
                    // trip), so write it always and keep it only for a record
                    // with a hit
                    *qp = (ebase + k) | (lo ? 1u << 30 : 0u) | (hi ? 1u << 31 : 0u);
                    qp += (lo || hi) ? 64 : 0;
                }
                if (__any(qp > qlim)) drain();
            }
        }
    };
    // frozen walls (wall_mode 1): a wave whose particles are all walls
    const bool wall_i = ((as_u32(Vi.m) >> (odd ? 31 : 15)) & 1u) != 0u;
    if (a.wall_mode && __all(!live || wall_i)) {
        scan_rows(BoolTag<true>{});
    } else {
        // the particle itself passes the predicate (r = 0) and adds zero:
        // it is not a pair
        total = live ? ~0u : 0u;
        scan_rows(BoolTag<false>{});
    }
    while (__any(qp != q)) drain();
    double cand = failed ? 0.0 : 1.0e30;  // a diverged state stops the clock
    if (live) {
        const int64_t n = a.n;
        // FP32 density ratio and its Tait sound speed,
        // c = sqrt(max(c0^2 (rho/rho0)^(gamma-1), 1e-16)) (src/EoS.h:86)
        const float r = (float)(a.dens[i] * a.inv_rho0);
        const float cs = a.g7 ? fmaxf(a.c0f * ((r * r) * r), 1.0e-8f)
                              : sqrtf(fmaxf(a.c0f * a.c0f * powf(r, a.gm1), 1.0e-16f));
        double ad[D];
        const bool frozen = a.wall_mode && a.fin.m.type[i] != 0;  // a frozen wall's acceleration reads as zero
#pragma unroll
        for (int d = 0; d < D; ++d) ad[d] = frozen ? 0.0 : a.ks * accp[d] + a.kv * accv[d];
        // div v = -(c0 / h) sum / (rho / rho0), and dt = C_CFL h / (c_i + v_sig c0), in FP32
        a.divv[i] = (double)(a.kdivv * divv * __builtin_amdgcn_rcpf(r));
        const float vs = fmaxf((float)vsig2.x, (float)vsig2.y);
        const double dti = (double)(a.cfl_h * __builtin_amdgcn_rcpf(fmaf(vs, a.c0f, cs)));
        a.dti[i] = dti;
        if (a.fused_final) cand = final_particle_h<D>(a.fin, a.s, i, n, ad, dti, a.hf, a.t2);
#pragma unroll
        for (int d = 0; d < D; ++d) a.acc[d * n + i] = ad[d];
    }
    block_count_store(total, a.P, vb.id, a.fused_final != 0, cand);
}

template <int D, int R, int VS, int BT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WCSPH_MIX_WAVES)))
void k_nforce_mixed(NFMixedArgs a) {
    nforce_mixed_body<D, R, VS, BT>(hw_block(), a);
}

}  // namespace wcsph
