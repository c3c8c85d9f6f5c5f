// Device pair interaction of the hydro force (pressure gradient, Monaghan
// artificial viscosity, laminar viscosity, velocity divergence), written for
// FP32 and FP64 from the reference functor's formula (src/force.h:38-57):
//
//   dr = r_i - r_j, dv = v_i - v_j, r2 = dr.dr            (skip if r2 <= 0)
//   den = r2 + eps h_i h_j
//   mu  = dr.dv < 0 ? (h_i + h_j)/2 dr.dv / den : 0
//   Pi  = (-alpha cbar mu + beta mu^2) / rhobar            (beta = 0 at :46)
//   gradW = F dr,  F = C_D g(s) / (H^(D+1) r),  H = h_i + h_j, s = r / H
//           g(s) = -3 (1-s)+^2 + 12 (1/2-s)+^2             (src/kernel.h:49-55
//           evaluated at the mean smoothing length, src/force.h:49)
//           grad_mode 1: F = (F(2 h_i) + F(2 h_j)) / 2, the reference GPU
//           kernel's mean of the gradients (src/kernel.cu:265), nonzero
//           within 2 max(h_i, h_j)
//   acc  -= m_j (P_i/rho_i^2 + P_j/rho_j^2 + Pi) gradW
//   acc  += m_j 2 (nu_i + nu_j) (dr.gradW) / (rho_i + rho_j) / den dv
//   divv -= m_j dv.gradW
//
// Outside the support F is exactly zero (both (.)+ terms vanish), so the
// caller may skip the second half there; mu still feeds the signal velocity.
#pragma once
#include <hip/hip_runtime.h>

namespace wcsph {

template <typename T>
struct PairConsts {
    T alpha, beta, eps, cnorm;  // cnorm = C_D: 80/(7 pi) in 2D, 16/pi in 3D (src/kernel.h:35-41)
    int vsig_mode;
    int grad_mode;  // 0: gradient at the mean h; 1: mean of the gradients at h_i and h_j
};

// Squared support radius of a pair: (h_i + h_j)^2, or (2 max(h_i, h_j))^2
// when the gradient is the mean of the two particles' gradients.
template <typename T>
__device__ __forceinline__ T pair_support2(T hi, T hj, int grad_mode) {
    const T H = grad_mode ? T(2) * (hi > hj ? hi : hj) : hi + hj;
    return H * H;
}

// g(r / H) / H^(D+1) of the cubic spline (src/kernel.h:49-55) for 1/H = invH.
template <int D, typename T>
__device__ __forceinline__ T spline_gH(T r, T invH) {
    const T s = r * invH;
    const T a = s < T(1) ? T(1) - s : T(0);
    const T b = s < T(0.5) ? T(0.5) - s : T(0);
    const T g = T(-3) * a * a + T(12) * b * b;
    const T invH2 = invH * invH;
    return g * (D == 3 ? invH2 * invH2 : invH2 * invH);
}

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double fast_rsqrt(double x) { return 1.0 / sqrt(x); }

template <typename T>
struct IState {  // per-i constants of the inner loop
    T h, rho, pr, c, nu;  // pr = P / rho^2
};

template <typename T>
struct JState {
    T h, rho, pr, c, nu, m;
};

// First half: regulariser, mu, and the signal-velocity update. Returns den.
template <typename T>
__device__ __forceinline__ T pair_mu(T r2, T drdv, const IState<T>& I, T hj, const PairConsts<T>& k,
                                     T& mu, T& vsig) {
    const T den = r2 + k.eps * (I.h * hj);
    const T hs = I.h + hj;
    mu = drdv < T(0) ? T(0.5) * hs * drdv * fast_rcp(den) : T(0);
    const T vs = k.vsig_mode ? -mu : mu;
    vsig = vs > vsig ? vs : vsig;
    return den;
}

// Second half, valid for 0 < r2 < (h_i+h_j)^2.
template <int D, typename T>
__device__ __forceinline__ void pair_force(const T* dr, const T* dv, T r2, T drdv, T den, T mu,
                                           const IState<T>& I, const JState<T>& J,
                                           const PairConsts<T>& k, T* acc, T& divv) {
    const T rho_s = I.rho + J.rho;
    const T inv_rs = fast_rcp(rho_s);
    const T cbar = T(0.5) * (I.c + J.c);
    const T Pi = (-k.alpha * cbar * mu + k.beta * mu * mu) * (T(2) * inv_rs);  // 1/rhobar = 2/(rho_i+rho_j)
    const T invr = fast_rsqrt(r2);
    const T r = r2 * invr;
    const T gH = k.grad_mode ? T(0.5) * (spline_gH<D>(r, fast_rcp(T(2) * I.h)) + spline_gH<D>(r, fast_rcp(T(2) * J.h)))
                             : spline_gH<D>(r, fast_rcp(I.h + J.h));
    const T mF = J.m * k.cnorm * gH * invr;                   // m_j F
    const T pt = I.pr + J.pr + Pi;
    const T vc = mF * T(2) * (I.nu + J.nu) * r2 * inv_rs * fast_rcp(den);
    const T cp = -mF * pt;
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] += cp * dr[d] + vc * dv[d];
    divv -= mF * drdv;
}

// Uniform-h form of the same pair body (both particles have smoothing
// length h): H = 2h and 1/H^(D+1) are constants (hc.invH, hc.invHD1), the
// regulariser is eps h^2 (hc.eps_h2), and one reciprocal of
// den (rho_i + rho_j) serves both 1/den and 1/(rho_i + rho_j). VS selects the
// signal-velocity rule at compile time (0: max(v_sig, mu), which stays 0;
// 1: max(v_sig, -mu)).
struct UniformH {
    float h, eps_h2, invH, invHD1;
};

template <int D, int VS>
__device__ __forceinline__ void pair_uniform(const float* dr, const float* dv, float r2, float drdv,
                                             const IState<float>& I, const JState<float>& J,
                                             const PairConsts<float>& k, const UniformH& hc, float* acc,
                                             float& divv, float& vsig) {
    const float den = r2 + hc.eps_h2;
    const float rho_s = I.rho + J.rho;
    const float X = fast_rcp(den * rho_s);              // 1 / (den (rho_i + rho_j))
    const float mneg = fminf(drdv, 0.f);
    const float mu = hc.h * mneg * (X * rho_s);          // (h_i + h_j)/2 = h
    if (VS == 1) vsig = fmaxf(vsig, -mu);
    const float cbar = 0.5f * (I.c + J.c);
    const float Pi = (-k.alpha * cbar * mu + k.beta * mu * mu) * (2.f * X * den);  // / rhobar
    const float invr = fast_rsqrt(r2);
    const float s = r2 * invr * hc.invH;
    const float a = fmaxf(1.f - s, 0.f);
    const float b = fmaxf(0.5f - s, 0.f);
    const float g = -3.f * a * a + 12.f * b * b;
    const float mF = J.m * k.cnorm * hc.invHD1 * g * invr;  // m_j F
    const float pt = I.pr + J.pr + Pi;
    const float vc = mF * 2.f * (I.nu + J.nu) * r2 * X;
    const float cp = -mF * pt;
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] += cp * dr[d] + vc * dv[d];
    divv -= mF * drdv;
}

}  // namespace wcsph
