// Device kernels of the device-resident simulation step.
//
// Layout in HBM (n particles, all SoA with stride n, kept in cell order and
// re-permuted every step):
//   pos, vel, vh : FP64 [D][n]       master state (kick-drift in FP64)
//   dens, mass, h: FP64 [n]
//   type: u8 [n] (0 fluid, 1 wall), id: u32 [n] (index at upload)
//   hf: FP32 [n] smoothing length, q4: float4 [n] {rho, P/rho^2, c, m}
//   acc: FP64 [D][n], divv, dti: FP64 [n]   force outputs
//   cell_start: u32 [ncell+1], nbr: u32 [K][n] (column per list slot), cnt: u32 [n]
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pair.hpp"

namespace wcsph {

constexpr int kBlock = 256;

struct Grid {
    double lo[3];
    double invL;
    int nc[3];
};

struct Master {
    double* pos;
    double* vel;
    double* vh;
    double* dens;
    double* mass;
    double* h;
    uint8_t* type;
    uint32_t* id;
};

struct Scalars {  // device-resident step scalars
    double dt, time;
    unsigned long long dtmin_bits;
    long long step;
    unsigned long long pairs_last, pairs_total;
    unsigned int escaped, maxcnt, overflow, pad;
};

// Map blockIdx so consecutive logical blocks share an XCD (and its L2):
// hardware places block b on XCD b % 8.
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int q = nb / 8, r = nb % 8, x = b % 8;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + b / 8;
}

template <int D>
__device__ __forceinline__ void cell_of(const Grid& g, const double* x, int* c, bool& esc) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double u = (x[d] - g.lo[d]) * g.invL;
        int k = (int)floor(u);
        if (!(u >= 0.0)) { k = 0; esc = true; }  // also catches NaN
        if (k > g.nc[d] - 1) { k = g.nc[d] - 1; esc = true; }
        c[d] = k;
    }
}

template <int D>
__device__ __forceinline__ uint32_t cell_key(const Grid& g, const int* c) {
    return D == 3 ? (uint32_t)((c[2] * g.nc[1] + c[1]) * g.nc[0] + c[0]) : (uint32_t)(c[1] * g.nc[0] + c[0]);
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// mode 0: full reset; 1: commit dt = dtmin, time += dt, step += 1;
// 2: reset the per-pass counters only. Modes 0 and 1 also reset dtmin.
__global__ void k_scalars(Scalars* s, int mode) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (mode == 0) {
        s->dt = 0.0;
        s->time = 0.0;
        s->step = 0;
        s->pairs_total = 0;
        s->overflow = 0;
    }
    if (mode == 1) {
        s->dt = __longlong_as_double((long long)s->dtmin_bits);
        s->time += s->dt;
        s->step += 1;
    }
    if (mode != 2) s->dtmin_bits = (unsigned long long)__double_as_longlong(1.0e30);
    s->pairs_last = 0;
    s->escaped = 0;
    s->maxcnt = 0;
}

// src/class.h:146-159 (initialKick, fullDrift, predict), all in FP64.
template <int D>
__global__ void k_kick_drift(Master m, const double* __restrict__ acc, const double* __restrict__ divv,
                             const Scalars* s, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = s->dt;
    m.dens[i] += -m.dens[i] * dt * divv[i];
    if (m.type[i] != 0) return;  // FREEZE: no kick, no drift
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double a = acc[d * n + i];
        const double v = m.vel[d * n + i];
        const double vh = v + 0.5 * dt * a;
        m.vh[d * n + i] = vh;
        m.pos[d * n + i] += dt * vh;
        m.vel[d * n + i] = v + dt * a;
    }
}

// Counting sort, pass 1: cell key and slot within the cell.
template <int D>
__global__ void k_bin(const double* __restrict__ pos, int64_t n, Grid g, uint32_t* __restrict__ key,
                      uint32_t* __restrict__ slot, uint32_t* __restrict__ count, Scalars* s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = pos[d * n + i];
    int c[3] = {0, 0, 0};
    bool esc = false;
    cell_of<D>(g, x, c, esc);
    const uint32_t k = cell_key<D>(g, c);
    key[i] = k;
    slot[i] = atomicAdd(&count[k], 1u);
    if (esc) atomicAdd(&s->escaped, 1u);
}

// Exclusive scan of count[0..m) into start[0..m], start[m] = total.
// 3 phases over tiles of 1024 elements (256 threads x 4).
constexpr int kScanTile = 1024;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[kBlock - 1];
    const uint32_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t m,
                                                         uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[kBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) v += (base + k < m) ? in[base + k] : 0u;
    uint32_t total;
    block_exclusive_scan(v, sh, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_partials(uint32_t* part, int64_t np) {
    __shared__ uint32_t sh[kBlock];
    uint32_t carry = 0;
    for (int64_t b = 0; b < np; b += kBlock) {
        const int64_t i = b + threadIdx.x;
        const uint32_t v = i < np ? part[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, sh, total);
        if (i < np) part[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_final(const uint32_t* __restrict__ in, int64_t m,
                                                        const uint32_t* __restrict__ part,
                                                        uint32_t* __restrict__ out) {
    __shared__ uint32_t sh[kBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < m) ? in[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = part[blockIdx.x] + block_exclusive_scan(sum, sh, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < m) out[base + k] = run;
        run += v[k];
        if (base + k == m - 1) out[m] = run;
    }
}

// Counting sort, pass 2: scatter old indices to their cell slots.
__global__ void k_scatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ slot,
                          const uint32_t* __restrict__ start, uint32_t* __restrict__ perm, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[start[key[i]] + slot[i]] = (uint32_t)i;
}

// The slot order within a cell comes from atomics; sorting each cell's
// segment by old index makes the permutation a function of the input only.
__global__ void k_cell_sort(const uint32_t* __restrict__ start, uint32_t* __restrict__ perm, int64_t ncell) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const uint32_t b = start[c], e = start[c + 1];
    for (uint32_t i = b + 1; i < e; ++i) {
        const uint32_t v = perm[i];
        uint32_t j = i;
        while (j > b && perm[j - 1] > v) {
            perm[j] = perm[j - 1];
            --j;
        }
        perm[j] = v;
    }
}

struct EosConsts {
    double rho0, B, gamma;
};

// Gather the master state into cell order and derive the force inputs
// (CalcPressure, src/force.h:3-9 with Murnaghan, src/EoS.h:78-86).
// pr64/c64 (P/rho^2 and c in FP64) are written only for the FP64 path.
template <int D>
__global__ void k_reorder(const uint32_t* __restrict__ perm, Master src, Master dst, float* __restrict__ hf,
                          float4* __restrict__ q4, double* __restrict__ pr64, double* __restrict__ c64,
                          EosConsts e, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t o = perm[p];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        dst.pos[d * n + p] = src.pos[d * n + o];
        dst.vel[d * n + p] = src.vel[d * n + o];
        dst.vh[d * n + p] = src.vh[d * n + o];
    }
    const double rho = src.dens[o], m = src.mass[o], h = src.h[o];
    dst.dens[p] = rho;
    dst.mass[p] = m;
    dst.h[p] = h;
    dst.type[p] = src.type[o];
    dst.id[p] = src.id[o];
    const double r = rho / e.rho0;
    double P = e.B * (pow(r, e.gamma) - 1.0);
    P = P > 0.0 ? P : 0.0;
    double c2 = e.B * e.gamma / e.rho0 * pow(r, e.gamma - 1.0);
    c2 = c2 > 1.0e-16 ? c2 : 1.0e-16;
    const double pr = P / (rho * rho), c = sqrt(c2);
    hf[p] = (float)h;
    q4[p] = make_float4((float)rho, (float)pr, (float)c, (float)m);
    if (pr64) {
        pr64[p] = pr;
        c64[p] = c;
    }
}

// Wave-then-block reduction of (sum, max) into the step scalars.
__device__ __forceinline__ void block_count_reduce(uint32_t cnt, uint32_t over, Scalars* s) {
    __shared__ unsigned long long sh_sum[kBlock / 64];
    __shared__ uint32_t sh_max[kBlock / 64];
    __shared__ uint32_t sh_over[kBlock / 64];
    unsigned long long v = cnt;
    uint32_t mx = cnt, ov = over;
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_down(v, off, 64);
        const uint32_t o = __shfl_down(mx, off, 64);
        mx = o > mx ? o : mx;
        ov |= __shfl_down(ov, off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        sh_sum[wv] = v;
        sh_max[wv] = mx;
        sh_over[wv] = ov;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        uint32_t m = 0, o = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            t += sh_sum[k];
            m = sh_max[k] > m ? sh_max[k] : m;
            o |= sh_over[k];
        }
        atomicAdd(&s->pairs_last, t);
        atomicAdd(&s->pairs_total, t);
        atomicMax(&s->maxcnt, m);
        if (o) atomicOr(&s->overflow, 1u);
    }
}

// Neighbour pass: for each i, every j != i in the 3^D cells around i's cell
// with r2 > 0 and r2 < (h_i + h_j)^2, where the pair distance is formed as
// (float)(x_i - x_j) per component and summed in FP32 without contraction.
template <int D>
__global__ __launch_bounds__(kBlock) void k_neigh(const double* __restrict__ pos, const float* __restrict__ hf,
                                                  const uint32_t* __restrict__ start, Grid g, int64_t n,
                                                  int K, uint32_t* __restrict__ nbr, uint32_t* __restrict__ cnt,
                                                  Scalars* s) {
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    uint32_t c_out = 0;
    if (i < n) {
        double xi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xi[d] = pos[d * n + i];
        const float hi = hf[i];
        int c[3] = {0, 0, 0};
        bool esc = false;
        cell_of<D>(g, xi, c, esc);
        const int zr = D == 3 ? 1 : 0;
        const int xa = c[0] > 0 ? c[0] - 1 : 0;
        const int xb = c[0] < g.nc[0] - 1 ? c[0] + 1 : g.nc[0] - 1;
        for (int dz = -zr; dz <= zr; ++dz) {
            const int zz = D == 3 ? c[2] + dz : 0;
            if (D == 3 && (zz < 0 || zz >= g.nc[2])) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int yy = c[1] + dy;
                if (yy < 0 || yy >= g.nc[1]) continue;
                const int64_t row = D == 3 ? ((int64_t)zz * g.nc[1] + yy) * g.nc[0] : (int64_t)yy * g.nc[0];
                const uint32_t jb = start[row + xa], je = start[row + xb + 1];
                for (uint32_t j = jb; j < je; ++j) {
                    if ((int64_t)j == i) continue;
                    float r2 = 0.f;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const float dd = (float)(xi[d] - pos[d * n + j]);
                        r2 = __fadd_rn(r2, __fmul_rn(dd, dd));
                    }
                    const float H = __fadd_rn(hi, hf[j]);
                    if (r2 > 0.f && r2 < __fmul_rn(H, H)) {
                        if ((int)c_out < K) nbr[(int64_t)c_out * n + i] = j;
                        ++c_out;
                    }
                }
            }
        }
        cnt[i] = c_out;
    }
    block_count_reduce(c_out, c_out > (uint32_t)K ? 1u : 0u, s);
}

template <typename T>
struct ForceArgs {
    const double* pos;
    const double* vel;
    const float* hf;
    const float4* q4;
    // FP64 path inputs (master arrays and FP64 EoS outputs)
    const double* h64;
    const double* rho64;
    const double* m64;
    const double* pr64;
    const double* c64;
    const uint32_t* nbr;
    const uint32_t* cnt;
    double* acc;
    double* divv;
    double* dti;
    int64_t n;
    int K;
    double nu;
    double c_cfl;
    PairConsts<T> k;
};

// Force pass over the neighbour lists; T is the pair arithmetic type.
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void k_force(ForceArgs<T> a) {
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int64_t n = a.n;
    double xi[D], vi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xi[d] = a.pos[d * n + i];
        vi[d] = a.vel[d * n + i];
    }
    IState<T> I;
    if constexpr (sizeof(T) == 8) {
        I = IState<T>{a.h64[i], a.rho64[i], a.pr64[i], a.c64[i], (T)a.nu};
    } else {
        const float4 qi = a.q4[i];
        I = IState<T>{(T)a.hf[i], (T)qi.x, (T)qi.y, (T)qi.z, (T)a.nu};
    }
    T acc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] = T(0);
    T divv = T(0), vsig = T(0);
    const int c = min((int)a.cnt[i], a.K);
    for (int t = 0; t < c; ++t) {
        const uint32_t j = a.nbr[(int64_t)t * n + i];
        T dr[D], dv[D];
        T r2 = T(0), drdv = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            dr[d] = (T)(xi[d] - a.pos[d * n + j]);
            dv[d] = (T)(vi[d] - a.vel[d * n + j]);
            r2 += dr[d] * dr[d];
            drdv += dr[d] * dv[d];
        }
        JState<T> J;
        if constexpr (sizeof(T) == 8) {
            J = JState<T>{a.h64[j], a.rho64[j], a.pr64[j], a.c64[j], (T)a.nu, a.m64[j]};
        } else {
            const float4 qj = a.q4[j];
            J = JState<T>{(T)a.hf[j], (T)qj.x, (T)qj.y, (T)qj.z, (T)a.nu, (T)qj.w};
        }
        T mu;
        const T den = pair_mu(r2, drdv, I, J.h, a.k, mu, vsig);
        pair_force<D>(dr, dv, r2, drdv, den, mu, I, J, a.k, acc, divv);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) a.acc[d * n + i] = (double)acc[d];
    a.divv[i] = (double)divv / (double)I.rho;                                   // src/force.h:57
    a.dti[i] = a.c_cfl * (double)I.h / ((double)I.c + (double)vsig);           // src/force.h:56
}

// Gravity (src/init/Dam2.h:104-108), finalKick (src/class.h:160-163),
// density floor (src/init/Dam2.h:111) and the per-particle time-step bound
// of getTimeStepGlobal (src/integral.h:2-12) reduced with an atomic min.
template <int D>
__global__ __launch_bounds__(kBlock) void k_final(Master m, double* __restrict__ acc,
                                                  const double* __restrict__ dti, double g, double nu,
                                                  double rho_floor, int kick, Scalars* s, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double cand = 1.0e30;
    if (i < n) {
        acc[(D - 1) * n + i] -= g;
        const bool fluid = m.type[i] == 0;
        double a2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double a = acc[d * n + i];
            a2 += a * a;
            if (kick && fluid) m.vel[d * n + i] = m.vh[d * n + i] + 0.5 * s->dt * a;
        }
        if (m.dens[i] < rho_floor) m.dens[i] = rho_floor;
        if (fluid) {
            const double h = m.h[i];
            const double t1 = 0.25 * sqrt(h / sqrt(a2));
            const double t2 = 0.125 * h * h / nu;
            cand = dti[i];
            cand = t1 < cand ? t1 : cand;
            cand = t2 < cand ? t2 : cand;
        }
    }
    // positive doubles order like their bit patterns
    unsigned long long b = (unsigned long long)__double_as_longlong(cand);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(b, off, 64);
        b = o < b ? o : b;
    }
    __shared__ unsigned long long sh[kBlock / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = sh[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) t = sh[k] < t ? sh[k] : t;
        atomicMin(&s->dtmin_bits, t);
    }
}

}  // namespace wcsph
