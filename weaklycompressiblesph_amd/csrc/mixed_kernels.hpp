// Mixed-precision fused neighbour + force pass (uniform smoothing length).
//
// Units: lengths in h, velocities in c0, densities in rho0, P/rho^2 in
// c0^2/rho0, masses in rho0 h^D. In these units every quantity of the pair
// body is O(1e-4..50) and fits FP16 (in SI units 1/H^4 alone would
// overflow it).
//
//  * scan: particles are stored in pairs (2m, 2m+1) as half2 lanes of one
//    16-byte record {x, y, z, x-cell}, cell-local and h-scaled, so one load
//    and one packed FP16 predicate test two candidates;
//  * body: two queued neighbours per lane step; pair differences are formed
//    in FP32 from the FP32 cell-local records, scaled, rounded once to FP16
//    (v_cvt_pk_f16_f32), and the whole pair body runs on packed FP16 VALU
//    ops (v_pk_fma_f16 ...); per-particle sums accumulate in FP32 through
//    v_dot2_f32_f16.
// The neighbour set of this path is its own FP16 predicate's: it differs from
// the FP32 path's only for pairs within ~1e-3 h of the support edge, whose
// kernel weight is ~1e-6.
#pragma once
#include "sim_kernels.hpp"

namespace wcsph {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));

struct alignas(16) ScanRec16 {  // particles 2m (.x) and 2m+1 (.y)
    h2 x, y, z, cx;
};

struct alignas(8) Quad16 {  // per particle
    _Float16 rho, pr, c, m;
};

struct Mixed {
    ScanRec16* sc16;  // [(n+1)/2]
    Quad16* q16;      // [n]
};

struct MixedConsts {
    float inv_h, inv_c0, inv_rho0, pr_scale, m_scale;  // SI -> scaled units
};

// Extra stores of k_reorder in mixed mode (thread p owns half p&1 of record p>>1).
template <int D>
__device__ __forceinline__ void write_mixed(const Mixed& mx, const MixedConsts& mc, int64_t p, const float* xl,
                                            int cx, double rho, double pr, double c, double m) {
    _Float16* r = reinterpret_cast<_Float16*>(&mx.sc16[p >> 1]);
    const int o = (int)(p & 1);
    r[0 + o] = (_Float16)(xl[0] * mc.inv_h);
    r[2 + o] = (_Float16)(xl[1] * mc.inv_h);
    r[4 + o] = (_Float16)(D == 3 ? xl[2] * mc.inv_h : 0.f);
    r[6 + o] = (_Float16)(float)cx;
    Quad16 q;
    q.rho = (_Float16)(float)(rho * mc.inv_rho0);
    q.pr = (_Float16)(float)(pr * mc.pr_scale);
    q.c = (_Float16)(float)(c * mc.inv_c0);
    q.m = (_Float16)(float)(m * mc.m_scale);
    mx.q16[p] = q;
}

template <int D>
struct MixedExtra {  // k_reorder hook writing the FP16 records
    Mixed mx;
    MixedConsts mc;
    __device__ void operator()(int64_t p, const float* xl, int cx, double rho, double pr, double c, double m) const {
        write_mixed<D>(mx, mc, p, xl, cx, rho, pr, c, m);
    }
};

__device__ __forceinline__ h2 hsplat(float v) {
    const _Float16 h = (_Float16)v;
    return h2{h, h};
}
__device__ __forceinline__ h2 hpack(float a, float b) { return h2{(_Float16)a, (_Float16)b}; }
__device__ __forceinline__ h2 hrcp(h2 v) {
    return h2{(_Float16)__builtin_amdgcn_rcph(v.x), (_Float16)__builtin_amdgcn_rcph(v.y)};
}
__device__ __forceinline__ h2 hrsq(h2 v) {
    return h2{(_Float16)__builtin_amdgcn_rsqh(v.x), (_Float16)__builtin_amdgcn_rsqh(v.y)};
}
__device__ __forceinline__ h2 hmax(h2 a, h2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ h2 hmin(h2 a, h2 b) { return __builtin_elementwise_min(a, b); }

struct NFMixedArgs {
    Packed pk;  // px: FP32 cell-local (SI) + x cell; pv: v / c0 (FP32) + m
    Mixed mx;
    const uint32_t* start;
    Grid g;
    int64_t n;
    float h, inv_h, Ls;  // uniform smoothing length, its inverse (as in the records); cell edge / h
    float alpha, beta, eps, cnorm_s;  // cnorm_s = C_D / 2^(D+1)
    int vsig_mode;
    // wave-uniform packed constants (kept in scalar registers)
    h2 k_eps, k_cn, k_malpha, k_beta, k_invh, k_Ls;
    double c0, nu, c_cfl, inv_rho0;
    double* acc;
    double* divv;
    double* dti;
    Partials P;
};

template <int D, int R, int Q>
__global__ __launch_bounds__(kBlock) void k_nforce_mixed(NFMixedArgs a) {
    constexpr int S = 2 * R + 1;
    constexpr int NR = D == 3 ? S * S : S;
    constexpr int kShift = 27;
    constexpr uint32_t kMask = (1u << kShift) - 1u;
    __shared__ uint32_t qbuf[kBlock / 64][64 * (Q + 1)];
    __shared__ float2 rowdc[NR];
    if ((int)threadIdx.x < NR)
        rowdc[threadIdx.x] = make_float2((float)(R - (int)threadIdx.x % S), D == 3 ? (float)(R - (int)threadIdx.x / S) : 0.f);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    uint32_t* q = &qbuf[threadIdx.x >> 6][lane * (Q + 1)];
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    const bool live = i < a.n;
    const float4 xi4 = live ? a.pk.px[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vi4 = live ? a.pk.pv[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    Quad16 qi;
    if (live) qi = a.mx.q16[i];
    else qi = Quad16{(_Float16)1.f, (_Float16)0.f, (_Float16)1.f, (_Float16)0.f};
    const float xi[3] = {xi4.x, xi4.y, xi4.z};
    const float vi[3] = {vi4.x, vi4.y, vi4.z};
    const float Lf = a.g.Lf, invLf = 1.0f / Lf, inv_h = a.inv_h;
    const float xs[3] = {xi[0] * inv_h, xi[1] * inv_h, xi[2] * inv_h};
    int c[3] = {0, 0, 0};
    if (live) key_cell<D>(a.g, a.pk.pcell[i], c);
    const float hc = 2.f * a.h * 1.0001f, hc2 = hc * hc;
    const _Float16 H2s = (_Float16)4.0f;  // (2h)^2 in units of h^2
    const h2 cxi2 = hsplat((float)c[0]), xi2 = hsplat(xs[0]), Ls2 = a.k_Ls;
    const h2 rhoi2 = h2{qi.rho, qi.rho}, pri2 = h2{qi.pr, qi.pr}, ci2 = h2{qi.c, qi.c};
    const h2 eps2 = a.k_eps, half2c = h2{(_Float16)0.5f, (_Float16)0.5f}, one2 = h2{(_Float16)1.f, (_Float16)1.f};
    const h2 three2 = h2{(_Float16)-3.f, (_Float16)-3.f}, twelve2 = h2{(_Float16)12.f, (_Float16)12.f};
    const h2 cn2 = a.k_cn, malpha2 = a.k_malpha, beta2 = a.k_beta, inv_h2 = a.k_invh;
    float accp[3] = {0.f, 0.f, 0.f}, accv[3] = {0.f, 0.f, 0.f}, divv = 0.f;
    h2 vsig2 = hsplat(0.f);
    uint32_t qn = 0, total = 0;

    auto flush = [&]() {
        for (uint32_t t = 0; __any(t < qn); t += 2) {
            if (t < qn) {
                const bool two = t + 1 < qn;
                const uint32_t e0 = q[t], e1 = two ? q[t + 1] : e0;
                const uint32_t j0 = e0 & kMask, j1 = e1 & kMask;
                const float4 p0 = a.pk.px[j0], p1 = a.pk.px[j1];
                const float4 v0 = a.pk.pv[j0], v1 = a.pk.pv[j1];
                const Quad16 q0 = a.mx.q16[j0], q1 = a.mx.q16[j1];
                const float2 d0 = rowdc[e0 >> kShift], d1 = rowdc[e1 >> kShift];
                const float dc0[3] = {xi4.w - p0.w, d0.x, d0.y}, dc1[3] = {xi4.w - p1.w, d1.x, d1.y};
                const float pj0[3] = {p0.x, p0.y, p0.z}, pj1[3] = {p1.x, p1.y, p1.z};
                const float vj0[3] = {v0.x, v0.y, v0.z}, vj1[3] = {v1.x, v1.y, v1.z};
                h2 dr[D], dv[D];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    dr[d] = hpack(fmaf(dc0[d], Lf, xi[d]) - pj0[d], fmaf(dc1[d], Lf, xi[d]) - pj1[d]) * inv_h2;
                    dv[d] = hpack(vi[d] - vj0[d], vi[d] - vj1[d]);
                }
                h2 r2 = dr[0] * dr[0], drdv = dr[0] * dv[0];
#pragma unroll
                for (int d = 1; d < D; ++d) {
                    r2 = dr[d] * dr[d] + r2;
                    drdv = dr[d] * dv[d] + drdv;
                }
                r2 = hmax(r2, hsplat(1.0e-6f));  // FP16 underflow guard (kernel weight ~0 there)
                const h2 rhoj = h2{q0.rho, q1.rho}, prj = h2{q0.pr, q1.pr};
                const h2 cj = h2{q0.c, q1.c}, mj = h2{q0.m, q1.m * (_Float16)(two ? 1.f : 0.f)};
                const h2 den = r2 + eps2;                     // r^2 + eps h_i h_j
                const h2 rho_s = rhoi2 + rhoj;
                const h2 X = hrcp(den * rho_s);               // 1 / (den (rho_i + rho_j))
                const h2 inv_den = X * rho_s, inv_rs = X * den;
                const h2 mu = hmin(drdv, hsplat(0.f)) * inv_den;  // (h_i + h_j)/2 = 1
                if (a.vsig_mode) vsig2 = hmax(vsig2, -mu);
                const h2 cbar = (ci2 + cj) * half2c;
                const h2 Pi = (malpha2 * cbar + beta2 * mu) * mu * inv_rs;
                const h2 invr = hrsq(r2);
                const h2 sx = r2 * invr * half2c;             // r / (2h)
                const h2 aa = hmax(one2 - sx, hsplat(0.f)), bb = hmax(half2c - sx, hsplat(0.f));
                const h2 g = three2 * aa * aa + twelve2 * bb * bb;
                const h2 mF = mj * cn2 * g * invr;            // m_j F (scaled)
                const h2 cp = -mF * (pri2 + prj + Pi);
                const h2 vc = mF * r2 * X;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    accp[d] = __builtin_amdgcn_fdot2(cp, dr[d], accp[d], false);
                    accv[d] = __builtin_amdgcn_fdot2(vc, dv[d], accv[d], false);
                }
                divv = __builtin_amdgcn_fdot2(mF, drdv, divv, false);
            }
        }
        total += qn;
        qn = 0;
    };

    auto row_bounds = [&](int r, uint32_t& b0, uint32_t& b1) {
        const int dy = r % S - R, dz = D == 3 ? r / S - R : 0;
        const int cy = c[1] + dy, cz = c[2] + dz;
        b0 = b1 = 0;
        if (live && cy >= 0 && cy < a.g.nc[1] && (D == 2 || (cz >= 0 && cz < a.g.nc[2]))) {
            const float gy = axis_gap(dy, xi[1], Lf);
            const float gz = D == 3 ? axis_gap(dz, xi[2], Lf) : 0.f;
            const float g2 = gy * gy + gz * gz;
            if (g2 < hc2) {
                const float w = sqrtf(hc2 - g2);
                int xa = (int)floorf((xi[0] - w) * invLf), xb = (int)floorf((xi[0] + w) * invLf);
                xa = max(max(xa, -R), -c[0]);
                xb = min(min(xb, R), a.g.nc[0] - 1 - c[0]);
                if (xa <= xb) {
                    const uint32_t row = D == 3 ? (uint32_t)((cz * a.g.nc[1] + cy) * a.g.nc[0])
                                                : (uint32_t)(cy * a.g.nc[0]);
                    b0 = row + c[0] + xa;
                    b1 = row + c[0] + xb + 1;
                }
            }
        }
    };
    uint32_t nb0, nb1;
    row_bounds(0, nb0, nb1);
    uint32_t np0 = nb1 ? a.start[nb0] : 0u, npe = nb1 ? a.start[nb1] : 0u;
    for (int r = 0; r < NR; ++r) {
        const int dy = r % S - R, dz = D == 3 ? r / S - R : 0;
        const uint32_t p0 = np0, len = npe - np0;
        if (r + 1 < NR) {
            row_bounds(r + 1, nb0, nb1);
            np0 = nb1 ? a.start[nb0] : 0u;
            npe = nb1 ? a.start[nb1] : 0u;
        }
        const uint32_t m0 = p0 >> 1;
        const uint32_t npair = len ? ((p0 + len + 1) >> 1) - m0 : 0;
        const h2 yi2 = hsplat(fmaf((float)(-dy), a.Ls, xs[1]));
        const h2 zi2 = hsplat(D == 3 ? fmaf((float)(-dz), a.Ls, xs[2]) : 0.f);
        constexpr int U = kScanUnroll / 2;  // records in flight (two candidates each)
        for (uint32_t t = 0; __any(t < npair); t += U) {
            ScanRec16 rec[U];
#pragma unroll
            for (int u = 0; u < U; ++u) rec[u] = a.mx.sc16[m0 + min(t + u, npair ? npair - 1 : 0u)];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const h2 dx = (cxi2 - rec[u].cx) * Ls2 + xi2 - rec[u].x;
                const h2 dyy = yi2 - rec[u].y;
                h2 r2 = dx * dx + dyy * dyy;
                if (D == 3) {
                    const h2 dzz = zi2 - rec[u].z;
                    r2 = dzz * dzz + r2;
                }
                const uint32_t j0 = 2 * (m0 + t + u), j1 = j0 + 1;
                const bool ok = t + u < npair;
                if (ok && j0 >= p0 && r2.x > (_Float16)0.f && r2.x < H2s) q[qn++] = j0 | ((uint32_t)r << kShift);
                if (ok && j1 < p0 + len && r2.y > (_Float16)0.f && r2.y < H2s) q[qn++] = j1 | ((uint32_t)r << kShift);
            }
            if (__any(qn + 2 * U > (uint32_t)Q)) flush();
        }
    }
    flush();
    if (live) {
        const int64_t n = a.n;
        const float4 qi4 = a.pk.pq[i];  // FP32 rho, P/rho^2, c, h
        const double c0 = a.c0, h = a.h;
        const double ks = c0 * c0 / h, kv = 4.0 * a.nu * c0 / (h * h);
#pragma unroll
        for (int d = 0; d < D; ++d) a.acc[d * n + i] = ks * accp[d] + kv * accv[d];
        a.divv[i] = -(c0 / h) * divv * (1.0 / ((double)qi4.x * (double)a.inv_rho0));
        const float vs = fmaxf((float)vsig2.x, (float)vsig2.y);
        a.dti[i] = a.c_cfl * h / ((double)qi4.z + (double)vs * c0);
    }
    block_count_partial(total, 0u, a.P, blockIdx.x);
}

}  // namespace wcsph
