// Device kernels of the device-resident simulation step.
//
// Layout in HBM (n particles, all SoA with stride n, kept in cell order and
// re-permuted every step):
//   pos, vel, vh : FP64 [D][n]       master state (kick-drift in FP64)
//   dens, mass, h: FP64 [n]
//   type: u8 [n] (0 fluid, 1 wall), id: u32 [n] (index at upload)
//   px: float4 [n] {x, y, z cell-local (x - cell origin), x cell index}
//   pv: float4 [n] {vx, vy, vz, m}, pq: float4 [n] {rho, P/rho^2, c, h}
//   pcell: u32 [n] cell key of each particle
//   acc: FP64 [D][n], divv, dti: FP64 [n]   force outputs
//   cell_start: u32 [ncell+1]
//   FP64 path only: nbr u32 [K][n] (slot t of all i contiguous), cnt u32 [n],
//   pr64/c64 FP64 [n]
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pair.hpp"

namespace wcsph {

constexpr int kBlock = 256;

struct Grid {
    double lo[3];
    double invL;  // 1 / L
    double L;     // cell edge (>= largest support / R)
    float Lf;     // (float)L
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
    // the step clock, read by every block of the step kernels
    double dt, time;
    unsigned long long dtmin_bits;
    long long step;
    unsigned long long pad_[28];
    // per-pass counters, updated by atomics from every block of the force
    // and binning passes: kept 256 bytes away from the clock so that reads
    // of dt do not queue behind them in L2
    unsigned long long pairs_last, pairs_total;
    unsigned int escaped, maxcnt, overflow;
    // sticky failure flags (cleared only by a full reset): particles whose
    // position is not finite (binned into the dead bucket past the last
    // cell), and cells holding more than kRankCap particles
    unsigned int nonfinite, overfull, pad_flags[3];
};
static_assert(__builtin_offsetof(Scalars, pairs_last) == 256, "counters on their own cache lines");
static_assert(__builtin_offsetof(Scalars, nonfinite) == 256 + 28, "failure flags are one 8-byte word");

// A state with a non-finite position or an overfull cell has diverged: the
// passes that scan cells then do no work (a diverged state could pile every
// particle into one cell and make each scan quadratic in n), and the host
// API reports the failure.
__device__ __forceinline__ bool sim_failed(const Scalars* s) { return (s->nonfinite | s->overfull) != 0u; }

// Particles a cell may hold for the ranked gather of k_reorder (beyond it
// the cell is flagged overfull).
constexpr uint32_t kRankCap = 256;

// The dead bucket: the cell key past the grid's last cell, for particles
// whose position is not finite. No stencil visits it.
template <int D>
__device__ __forceinline__ uint32_t dead_key(const Grid& g) {
    return (uint32_t)g.nc[0] * (uint32_t)g.nc[1] * (uint32_t)(D == 3 ? g.nc[2] : 1);
}

// Per-block partial results (one slot per block of the particle kernels),
// reduced by a one-block pass: no same-address atomics, which serialise
// far from the cores across the eight XCDs (one per block on one address
// cost a 2D force pass of 2,770 blocks more than half its time). The
// neighbour-list pass's counts are reduced at once by k_reduce_partials;
// a force pass's counts stay pending (*pending = 1) until the next
// k_scalars or k_fold_counts folds them into the step scalars.
struct Partials {
    unsigned long long* pairs;  // pair count per block
    uint32_t* maxcnt;           // largest per-particle count per block
    uint32_t* over;             // list overflow flag per block
    double* dtmin;              // time-step bound per block
    int64_t nblk;
    uint32_t* pending;          // 1: pairs/maxcnt hold a force pass not yet folded
};

// Fold a pending force pass's per-block counts into the step scalars:
// pairs_last and maxcnt of that pass, pairs_total += its pairs. Every thread
// of the block calls it (blockDim.x a multiple of 64, at most 1024).
__device__ __forceinline__ void fold_counts(const Partials& P, Scalars* s) {
    __shared__ unsigned long long fsum[16];
    __shared__ uint32_t fmax[16];
    if (P.pending == nullptr || *P.pending == 0u) return;  // the same value for every thread
    unsigned long long v = 0;
    uint32_t m = 0;
    for (int64_t b = threadIdx.x; b < P.nblk; b += blockDim.x) {
        v += P.pairs[b];
        m = P.maxcnt[b] > m ? P.maxcnt[b] : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_down(v, off, 64);
        const uint32_t o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) {
        fsum[threadIdx.x >> 6] = v;
        fmax[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            v += fsum[k];
            m = fmax[k] > m ? fmax[k] : m;
        }
        s->pairs_last = v;
        s->maxcnt = m;
        s->pairs_total += v;
        *P.pending = 0u;
    }
    __syncthreads();
}

// Map blockIdx onto particle blocks for the 8 XCDs (hardware places block b
// on XCD b % 8, each XCD with its own L2). WCSPH_XCD_CHUNK > 0: each XCD
// takes runs of that many consecutive particle blocks in turn, so all XCDs
// sweep the domain together (the cost per particle varies along the cell
// order) while neighbouring blocks still share an L2; blocks past the last
// whole cycle of 8 runs keep their index. WCSPH_XCD_CHUNK == 0: each XCD takes one
// contiguous eighth.
#ifndef WCSPH_XCD_CHUNK
#define WCSPH_XCD_CHUNK 64
#endif
// A block's index and the block count of the launch it belongs to: the
// hardware's for an ordinary launch; for a grouped launch (one grid over
// several simulations) the block's index among its simulation's blocks.
struct VBlock {
    int id, n;
};

__device__ __forceinline__ VBlock hw_block() { return VBlock{(int)blockIdx.x, (int)gridDim.x}; }

// A grouped launch starts each simulation's blocks at a multiple of 8, so a
// virtual block lands on the XCD its index names.
__device__ __forceinline__ int xcd_block(VBlock vb) {
    const int b = vb.id, nb = vb.n;
    constexpr int C = WCSPH_XCD_CHUNK;
    if (C > 0) {
        const int full = nb / (8 * C) * (8 * C);
        if (b >= full) return b;
        const int x = b % 8, k = b / 8;
        return ((k / C) * 8 + x) * C + k % C;
    }
    const int q = nb / 8, r = nb % 8, x = b % 8;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + b / 8;
}

__device__ __forceinline__ int xcd_block() { return xcd_block(hw_block()); }

template <int D>
__device__ __forceinline__ void cell_of(const Grid& g, const double* x, int* c, bool& esc) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double u = (x[d] - g.lo[d]) * g.invL;
        int k = (int)floor(fmin(fmax(u, -1.0), 2.0e9));  // in int range before the clamps
        if (!(u >= 0.0)) { k = 0; esc = true; }
        if (k > g.nc[d] - 1) { k = g.nc[d] - 1; esc = true; }
        c[d] = k;
    }
}

template <int D>
__device__ __forceinline__ uint32_t cell_key(const Grid& g, const int* c) {
    return D == 3 ? (uint32_t)((c[2] * g.nc[1] + c[1]) * g.nc[0] + c[0]) : (uint32_t)(c[1] * g.nc[0] + c[0]);
}

// mode 0: full reset; 1: commit dt = the minimum of the last final stage's
// per-block bounds, time += dt, step += 1; 2: reset the per-pass counters
// only. Modes 1 and 2 first fold a pending force pass's counts; mode 0
// drops them. One block of kScalarsThreads threads, four loads in flight
// each.
constexpr int kScalarsThreads = 1024;
__device__ __forceinline__ void scalars_body(Scalars* s, const Partials& P, int mode) {
    __shared__ double sh[kScalarsThreads / 64];
    __shared__ unsigned long long shp[kScalarsThreads / 64];
    __shared__ uint32_t shm[kScalarsThreads / 64];
    const int64_t nblk = P.nblk;
    // a pending force pass's pair counts are folded in the same pass over
    // the per-block slots as the dt bounds (the flag reads the same in every
    // thread: only thread 0 clears it, after the reduction)
    const bool fold = mode != 0 && P.pending != nullptr && *P.pending != 0u;
    const bool dt = mode == 1;
    double d = 1.0e30;
    unsigned long long v = 0;
    uint32_t m = 0;
    if (fold || dt) {
        const int64_t T = kScalarsThreads;
        double d0 = 1.0e30, d1 = 1.0e30, d2 = 1.0e30, d3 = 1.0e30;
        int64_t b = threadIdx.x;
        for (; b + 3 * T < nblk; b += 4 * T) {
            if (dt) {
                const double x0 = P.dtmin[b], x1 = P.dtmin[b + T], x2 = P.dtmin[b + 2 * T], x3 = P.dtmin[b + 3 * T];
                d0 = x0 < d0 ? x0 : d0;
                d1 = x1 < d1 ? x1 : d1;
                d2 = x2 < d2 ? x2 : d2;
                d3 = x3 < d3 ? x3 : d3;
            }
            if (fold) {
                v += (P.pairs[b] + P.pairs[b + T]) + (P.pairs[b + 2 * T] + P.pairs[b + 3 * T]);
                const uint32_t m0 = max(P.maxcnt[b], P.maxcnt[b + T]), m1 = max(P.maxcnt[b + 2 * T], P.maxcnt[b + 3 * T]);
                m = max(m, max(m0, m1));
            }
        }
        for (; b < nblk; b += T) {
            if (dt) d0 = P.dtmin[b] < d0 ? P.dtmin[b] : d0;
            if (fold) {
                v += P.pairs[b];
                m = max(m, P.maxcnt[b]);
            }
        }
        d0 = d1 < d0 ? d1 : d0;
        d2 = d3 < d2 ? d3 : d2;
        d = d2 < d0 ? d2 : d0;
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_down(d, off, 64);
            d = o < d ? o : d;
            v += __shfl_down(v, off, 64);
            const uint32_t om = __shfl_down(m, off, 64);
            m = om > m ? om : m;
        }
        if ((threadIdx.x & 63) == 0) {
            sh[threadIdx.x >> 6] = d;
            shp[threadIdx.x >> 6] = v;
            shm[threadIdx.x >> 6] = m;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 1; k < kScalarsThreads / 64; ++k) {
                d = sh[k] < d ? sh[k] : d;
                v += shp[k];
                m = shm[k] > m ? shm[k] : m;
            }
    }
    if (threadIdx.x != 0) return;
    if (fold) {  // the folded pass's pairs_last and maxcnt are replaced by the new pass's
        s->pairs_total += v;
        *P.pending = 0u;
    }
    if (mode == 0) {
        if (P.pending) *P.pending = 0u;
        s->dt = 0.0;
        s->time = 0.0;
        s->step = 0;
        s->pairs_total = 0;
        s->overflow = 0;
        s->nonfinite = 0;
        s->overfull = 0;
    }
    if (mode == 1) {
        s->dt = d;
        s->dtmin_bits = (unsigned long long)__double_as_longlong(d);
        s->time += s->dt;
        s->step += 1;
    }
    s->pairs_last = 0;
    s->escaped = 0;
    s->maxcnt = 0;
}

static __global__ __launch_bounds__(kScalarsThreads) void k_scalars(Scalars* s, Partials P, int mode) {
    scalars_body(s, P, mode);
}

// The fold alone (before the host reads the step scalars).
static __global__ __launch_bounds__(kBlock) void k_fold_counts(Partials P, Scalars* s) { fold_counts(P, s); }

// Counting sort, pass 1: cell key and slot within the cell, after the
// first half of a step when kick != 0: initialKick and fullDrift
// (src/class.h:146-155) in FP64 -- v_half = v + dt/2 a, rho += -rho dt div v,
// x += dt v_half -- walls (FREEZE) keep x and v. predict (v + dt a,
// src/class.h:156-159) is formed by k_reorder from v_half, so v itself is
// not written here. The state is in the previous pass's cell order, so
// neighbouring lanes mostly share a key: each run of equal keys within a
// wave takes its slots with one atomic, made by the run's first lane and
// broadcast to the rest.
template <int D>
__device__ __forceinline__ void bin_body(int64_t blk, Master m, const double* __restrict__ acc,
                                         const double* __restrict__ divv, int kick, int64_t n, Grid g,
                                         uint32_t* __restrict__ key, uint32_t* __restrict__ slot,
                                         uint32_t* __restrict__ count, Scalars* s) {
    const int64_t i = blk * blockDim.x + threadIdx.x;
    const bool live = i < n;
    uint32_t k = 0xffffffffu;
    if (live) {
        double x[D];
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = m.pos[d * n + i];
        if (kick) {
            const double dt = s->dt;
            m.dens[i] += -m.dens[i] * dt * divv[i];
            if (m.type[i] == 0) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double vh = m.vel[d * n + i] + 0.5 * dt * acc[d * n + i];
                    m.vh[d * n + i] = vh;
                    x[d] += dt * vh;
                    m.pos[d * n + i] = x[d];
                }
            }
        }
        bool finite = true;
#pragma unroll
        for (int d = 0; d < D; ++d) finite = finite && isfinite(x[d]);
        if (finite) {
            int c[3] = {0, 0, 0};
            bool esc = false;
            cell_of<D>(g, x, c, esc);
            k = cell_key<D>(g, c);
            if (esc) atomicAdd(&s->escaped, 1u);
        } else {
            k = dead_key<D>(g);
            atomicAdd(&s->nonfinite, 1u);
        }
        key[i] = k;
    }
    const int lane = threadIdx.x & 63;
    const uint32_t kp = (uint32_t)__shfl_up((int)k, 1, 64);
    const bool lead = lane == 0 || kp != k;  // first lane of a run (dead lanes form their own run)
    const unsigned long long leads = __ballot(lead);
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;
    const int first = 63 - __clzll(leads & upto);                    // first lane of my run
    const unsigned long long after = leads & ~upto;
    const int next = after ? __ffsll((long long)after) - 1 : 64;      // first lane of the next run
    uint32_t base = 0;
    if (lead && live) base = atomicAdd(&count[k], (uint32_t)(next - lane));
    base = (uint32_t)__shfl((int)base, first, 64);
    if (live) slot[i] = base + (uint32_t)(lane - first);
}

template <int D>
__global__ void k_bin(Master m, const double* __restrict__ acc, const double* __restrict__ divv, int kick,
                      int64_t n, Grid g, uint32_t* __restrict__ key, uint32_t* __restrict__ slot,
                      uint32_t* __restrict__ count, Scalars* s) {
    bin_body<D>(blockIdx.x, m, acc, divv, kick, n, g, key, slot, count, s);
}

// Exclusive scan of count[0..m) into start[0..m], start[m] = total, over
// tiles of 1024 elements (256 threads x 4).
constexpr int kScanTile = 1024;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Exclusive scan over the block (a multiple of 64 threads, at most 1024);
// total = the block's sum. sh holds at least blockDim.x / 64 words.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t incl = wave_inclusive_scan(v);
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
        const uint32_t t = sh[k];
        before += k < w ? t : 0u;
        tot += t;
    }
    __syncthreads();  // sh may be reused after the call
    total = tot;
    return before + incl - v;
}

// Two-pass exclusive scan of the cell counts, over tiles of kScanTile
// counts gathered into groups of `gsize` tiles (gsize chosen by the host
// so there are at most kScanMaxGroups groups). k_scan_reduce (one block
// per group, one wave per tile in turn) writes every tile's total and the
// group's total; k_scan_apply gives each tile its offset as the sum of the
// groups before it and of the tiles before it in its own group (at most
// kScanMaxGroups + gsize values, read by the block's threads together),
// then scans the tile. No atomics: same-address atomics from every XCD
// serialise far from the cores.
constexpr int kScanMaxGroups = 1024;

__host__ __device__ inline int64_t scan_group_size(int64_t ntiles) {
    const int64_t g = (ntiles + kScanMaxGroups - 1) / kScanMaxGroups;
    return g < 16 ? 16 : g;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* sh) {
    uint32_t total;
    block_exclusive_scan(v, sh, total);
    return total;
}

__device__ __forceinline__ void scan_reduce_body(int64_t blk, const uint32_t* __restrict__ in, int64_t m,
                                                 uint32_t* __restrict__ tile_total,
                                                 uint32_t* __restrict__ group_total, int64_t gsize) {
    __shared__ uint32_t sh[kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t ntiles = (m + kScanTile - 1) / kScanTile;
    const int64_t t0 = blk * gsize, t1 = t0 + gsize < ntiles ? t0 + gsize : ntiles;
    uint32_t mine = 0;  // this wave's share of the group total (lane 0)
    for (int64_t t = t0 + wv; t < t1; t += kBlock / 64) {
        // lane l sums counts [t * kScanTile + 16 l, + 16)
        const int64_t base = t * kScanTile + lane * 16;
        uint32_t v = 0;
        if (base + 16 <= m) {
            const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint4 q = p[k];
                v += q.x + q.y + q.z + q.w;
            }
        } else {
            for (int k = 0; k < 16; ++k) v += base + k < m ? in[base + k] : 0u;
        }
        for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
        if (lane == 0) {
            tile_total[t] = v;
            mine += v;
        }
    }
    const uint32_t g = block_sum(lane == 0 ? mine : 0u, sh);
    if (threadIdx.x == 0) group_total[blk] = g;
}

static __global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t m,
                                                                uint32_t* __restrict__ tile_total,
                                                                uint32_t* __restrict__ group_total, int64_t gsize) {
    scan_reduce_body(blockIdx.x, in, m, tile_total, group_total, gsize);
}

__device__ __forceinline__ void scan_apply_body(int64_t blk, const uint32_t* __restrict__ count, int64_t m,
                                                const uint32_t* __restrict__ tile_total,
                                                const uint32_t* __restrict__ group_total, int64_t gsize,
                                                uint32_t* __restrict__ start) {
    __shared__ uint32_t sh[kBlock / 64];
    const int64_t tile = blk, g = tile / gsize;
    uint32_t pre = 0;
    for (int64_t k = threadIdx.x; k < g; k += kBlock) pre += group_total[k];
    for (int64_t k = g * gsize + threadIdx.x; k < tile; k += kBlock) pre += tile_total[k];
    const uint32_t offset = block_sum(pre, sh);
    const int64_t base = tile * kScanTile + threadIdx.x * 4;
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < m) ? count[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = offset + block_exclusive_scan(sum, sh, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < m) start[base + k] = run;
        run += v[k];
        if (base + k == m - 1) start[m] = run;
    }
}

static __global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* __restrict__ count, int64_t m,
                                                               const uint32_t* __restrict__ tile_total,
                                                               const uint32_t* __restrict__ group_total,
                                                               int64_t gsize, uint32_t* __restrict__ start) {
    scan_apply_body(blockIdx.x, count, m, tile_total, group_total, gsize, start);
}

// Counting sort, pass 2: scatter old indices to their cell slots, and
// clear the counts the scan has consumed for the next binning pass (only
// occupied cells were counted, so only they need clearing).
__device__ __forceinline__ void scatter_body(int64_t blk, const uint32_t* __restrict__ key,
                                             const uint32_t* __restrict__ slot, const uint32_t* __restrict__ start,
                                             uint32_t* __restrict__ perm, uint32_t* __restrict__ count, int64_t n) {
    const int64_t i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = key[i];
    perm[start[k] + slot[i]] = (uint32_t)i;
    count[k] = 0u;
}

static __global__ void k_scatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ slot,
                                 const uint32_t* __restrict__ start, uint32_t* __restrict__ perm,
                                 uint32_t* __restrict__ count, int64_t n) {
    scatter_body(blockIdx.x, key, slot, start, perm, count, n);
}

struct EosConsts {
    double rho0, B, gamma;
    double inv_rho0, c0;  // 1 / rho0 and c0 = sqrt(B gamma / rho0)
};

// The Tait constants of rest density rho0, sound speed c0 and exponent
// gamma: B = c0^2 rho0 / gamma (src/init/Dam2.h:38-43).
inline EosConsts eos_consts(double rho0, double c0, double gamma) {
    return EosConsts{rho0, c0 * c0 * rho0 / gamma, gamma, 1.0 / rho0, c0};
}

// Murnaghan/Tait (src/EoS.h:78-86, CalcPressure src/force.h:3-9): P/rho^2
// and the sound speed of density rho, in FP64. For the water law (gamma =
// 7) the powers are multiplies and c = c0 (rho/rho0)^3, the square root of
// c^2 = B gamma / rho0 (rho/rho0)^6 taken in closed form.
__device__ __forceinline__ void tait(const EosConsts& e, double rho, double& pr, double& cs) {
    if (e.gamma == 7.0) {
        const double r = rho * e.inv_rho0;
        const double r3 = (r * r) * r;
        double P = e.B * ((r3 * r3) * r - 1.0);
        P = P > 0.0 ? P : 0.0;
        pr = P / (rho * rho);
        const double c = e.c0 * r3;
        cs = c > 1.0e-8 ? c : 1.0e-8;  // c^2 >= 1e-16
        return;
    }
    const double r = rho / e.rho0;
    double P = e.B * (pow(r, e.gamma) - 1.0);
    P = P > 0.0 ? P : 0.0;
    double c2 = e.B * e.gamma / e.rho0 * pow(r, e.gamma - 1.0);
    c2 = c2 > 1.0e-16 ? c2 : 1.0e-16;
    pr = P / (rho * rho);
    cs = sqrt(c2);
}

// Cell-local coordinate along one axis: x - (lo + c L), both steps in FP64
// without contraction, rounded once to FP32. The oracle evaluates the same
// expression (orc_neighbours_brute_cells).
__device__ __forceinline__ float cell_local(double x, double lo, int c, double L) {
    const double origin = __dadd_rn(lo, __dmul_rn((double)c, L));
    return (float)__dsub_rn(x, origin);
}

template <int D>
__device__ __forceinline__ void key_cell(const Grid& g, uint32_t key, int* c) {
    c[0] = (int)(key % (uint32_t)g.nc[0]);
    c[1] = (int)((key / (uint32_t)g.nc[0]) % (uint32_t)g.nc[1]);
    c[2] = D == 3 ? (int)(key / ((uint32_t)g.nc[0] * (uint32_t)g.nc[1])) : 0;
}

// The member of slot range [s0, s1) that slot p takes in cell order: the
// one with the (p - s0)-th smallest previous index (the scatter fills a
// cell's slots in atomic order; this makes the order a function of the
// state alone). Cells of up to 8 are ranked in registers; larger ones
// (callers cap them at kRankCap) by a search quadratic in the count.
__host__ __device__ inline uint32_t ranked_member(const uint32_t* perm, uint32_t s0, uint32_t s1, uint32_t p) {
    const uint32_t cnt = s1 - s0, want = p - s0;
    uint32_t o = perm[p];
    constexpr uint32_t kC = 8;
    if (cnt > 1 && cnt <= kC) {
        uint32_t mem[kC];
#pragma unroll
        for (uint32_t t = 0; t < kC; ++t) mem[t] = t < cnt ? perm[s0 + t] : 0xffffffffu;
#pragma unroll
        for (uint32_t t = 0; t < kC; ++t) {
            uint32_t r = 0;
#pragma unroll
            for (uint32_t u = 0; u < kC; ++u) r += mem[u] < mem[t] ? 1u : 0u;
            if (t < cnt && r == want) o = mem[t];
        }
    } else if (cnt > kC) {
        for (uint32_t q = s0; q < s1; ++q) {
            const uint32_t cand = perm[q];
            uint32_t r = 0;
            for (uint32_t q2 = s0; q2 < s1; ++q2) r += perm[q2] < cand ? 1u : 0u;
            if (r == want) {
                o = cand;
                break;
            }
        }
    }
    return o;
}

struct Packed {  // FP32 force records in cell order
    float4* px;
    float4* pv;
    float4* pq;
    uint32_t* pcell;
};

// Counting sort, last pass: gather the master state into cell order and
// derive the force inputs (CalcPressure, src/force.h:3-9 with Murnaghan,
// src/EoS.h:78-86). k_scatter put each cell's particles into the cell's slot
// range in atomic order; slot s0 + r takes the member with the r-th smallest
// previous index, so the order is a function of the state alone (cells hold
// a handful of particles).
// step != 0 (inside a step, after k_bin with kick): the force records take
// the predicted velocity v + dt a (predict, src/class.h:156-159) of fluid
// particles as v_half + dt/2 a_old, and v_half moves with them; the master
// velocity of fluid particles is left for the final kick to write, except
// with keep_vel (the FP64 force pass reads it). Walls copy v.
// step == 0: velocities are copied as they are.
// pr64/c64 (P/rho^2 and c in FP64) are written only for the FP64 path.
template <int D, class Extra>
__device__ __forceinline__ void reorder_body(int64_t blk, const uint32_t* __restrict__ perm,
                                             const uint32_t* __restrict__ key, const uint32_t* __restrict__ start,
                                             Master src, Master dst, const double* __restrict__ acc_old,
                                             Scalars* s, int step, int keep_vel, Packed pk, Grid g,
                                             double* __restrict__ pr64, double* __restrict__ c64, EosConsts e,
                                             float v_scale, double h_const, double m_const, Extra extra,
                                             int64_t n) {
    const int64_t p = blk * blockDim.x + threadIdx.x;
    const bool live = p < n;
    float xl[3] = {0.f, 0.f, 0.f}, vf[3] = {0.f, 0.f, 0.f};
    int c[3] = {0, 0, 0};
    double rho = 0.0, m = 0.0, pr = 0.0, cs = 0.0;
    bool wall = false;
    if (live) {
        const uint32_t k = key[perm[p]];
        const uint2 cr = make_uint2(start[k], start[k + 1]);  // the cell's slot range
        // the dead bucket and overfull cells keep the scatter's order (the
        // state has diverged; the flag makes the host API report it)
        const bool ranked = k != dead_key<D>(g) && cr.y - cr.x <= kRankCap;
        if (k != dead_key<D>(g) && cr.y - cr.x > kRankCap && (uint32_t)p == cr.x) atomicOr(&s->overfull, 1u);
        const uint32_t o = ranked ? ranked_member(perm, cr.x, cr.y, (uint32_t)p) : perm[p];
        key_cell<D>(g, k, c);
        const bool moving = step && src.type[o] == 0;
        const double hdt = step ? 0.5 * s->dt : 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double x = src.pos[d * n + o];
            dst.pos[d * n + p] = x;
            xl[d] = cell_local(x, g.lo[d], c[d], g.L);
            double v;
            if (moving) {
                const double vh = src.vh[d * n + o];
                v = vh + hdt * acc_old[d * n + o];
                dst.vh[d * n + p] = vh;
                if (keep_vel) dst.vel[d * n + p] = v;
            } else {
                v = src.vel[d * n + o];
                dst.vel[d * n + p] = v;
            }
            vf[d] = (float)(v * v_scale);
        }
        rho = src.dens[o];
        // h_const > 0 (m_const > 0): every smoothing length (mass) is that
        // constant, in both buffers, so the permutation leaves it as it is
        m = m_const > 0.0 ? m_const : src.mass[o];
        const double h = h_const > 0.0 ? h_const : src.h[o];
        dst.dens[p] = rho;
        if (!(m_const > 0.0)) dst.mass[p] = m;
        if (!(h_const > 0.0)) dst.h[p] = h;
        dst.type[p] = src.type[o];
        wall = src.type[o] != 0;
        dst.id[p] = src.id[o];
        tait(e, rho, pr, cs);
        if (Extra::kFp32Records) {
            pk.px[p] = make_float4(xl[0], xl[1], xl[2], (float)c[0]);
            pk.pv[p] = make_float4(vf[0], vf[1], vf[2], (float)m);
            pk.pq[p] = make_float4((float)rho, (float)pr, (float)cs, (float)h);
            pk.pcell[p] = k;
        }
        if (pr64) {
            pr64[p] = pr;
            c64[p] = cs;
        }
    }
    extra.store_pair(p, live, xl, c[0], vf, rho, pr, cs, m, wall);  // every lane takes part
}

template <int D, class Extra>
__global__ __launch_bounds__(kBlock) void k_reorder(const uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ key,
                                                    const uint32_t* __restrict__ start, Master src, Master dst,
                                                    const double* __restrict__ acc_old, Scalars* s, int step,
                                                    int keep_vel, Packed pk, Grid g, double* __restrict__ pr64,
                                                    double* __restrict__ c64, EosConsts e, float v_scale,
                                                    double h_const, double m_const, Extra extra, int64_t n) {
    reorder_body<D, Extra>(blockIdx.x, perm, key, start, src, dst, acc_old, s, step, keep_vel, pk, g, pr64, c64, e,
                           v_scale, h_const, m_const, extra, n);
}

struct NoExtra {  // FP32 and FP64 paths: the FP32 records are all there is
    static constexpr bool kFp32Records = true;
    __device__ void store_pair(int64_t, bool, const float*, int, const float*, double, double, double,
                               double, bool) const {}
    __device__ void state(int64_t, double, double, double) const {}
};

// Walks the (2R+1)^D cells around particle i's cell (R = cells per
// support length), yielding the candidate index range of each non-empty
// cell in turn, with the FP32 offset (c_i - c_j) L of that cell.
// With R = 2 cells whose box is farther than hcut from i are skipped.
// Gap along one axis between a cell-local coordinate x and the cell d cells
// away (index t = c + d of nc cells). Cells span [d L, (d + 1) L) in the
// local frame, except that the grid's first and last cells also hold every
// particle clamped into them from outside the grid (cell_of), so they are
// open below (t == 0) or above (t == nc - 1).
__device__ __forceinline__ float axis_gap(int d, float x, float Lf, int t, int nc) {
    const float lo = t <= 0 ? -3.0e38f : (float)d * Lf;
    const float hi = t >= nc - 1 ? 3.0e38f : (float)(d + 1) * Lf;
    return fmaxf(fmaxf(lo - x, x - hi), 0.f);
}


// The neighbour predicate on cell-local coordinates: per axis
// d = fma(c_i - c_j, L, xl_i) - xl_j, r2 = fma(dz, dz, fma(dy, dy, dx dx)),
// neighbour iff 0 < r2 < (h_i + h_j)^2; every operation is written out so
// nothing else is contracted (the oracle evaluates the same sequence).
// dc: (c_i - c_j) per axis as exact small floats.
template <int D>
__device__ __forceinline__ bool nbr_pred(const float* xi, float hi, const float4& pj, const float* dc, float Lf,
                                         float hj) {
    const float pjv[3] = {pj.x, pj.y, pj.z};
    float d[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < D; ++k) d[k] = __fsub_rn(fmaf(dc[k], Lf, xi[k]), pjv[k]);
    float r2 = __fmul_rn(d[0], d[0]);
    r2 = fmaf(d[1], d[1], r2);
    if (D == 3) r2 = fmaf(d[2], d[2], r2);
    const float H = __fadd_rn(hi, hj);
    return r2 > 0.f && r2 < __fmul_rn(H, H);
}

// Wave-then-block reduction of (sum, max, or) into this block's partial slot.
__device__ __forceinline__ void block_count_partial(uint32_t cnt, uint32_t over, const Partials& P, int64_t blk) {
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
        P.pairs[blk] = t;
        P.maxcnt[blk] = m;
        P.over[blk] = o;
    }
}

// Reduce the per-block partials in index order into the step scalars.
// what: bit 0 pair counts (added to pairs_last, and to pairs_total if
// bit 2), bit 1 the dt bound (into dtmin_bits). One block; sh is LDS
// scratch (a caller may lend memory it no longer uses).
struct ReduceScratch {
    unsigned long long p[kBlock];
    double d[kBlock];
    uint32_t m[kBlock], o[kBlock];
};

__device__ __forceinline__ void reduce_partials_block(const Partials& P, Scalars* s, int what, ReduceScratch& sh) {
    unsigned long long p = 0;
    uint32_t m = 0, o = 0;
    double d = 1.0e30;
    for (int64_t b = threadIdx.x; b < P.nblk; b += kBlock) {
        if (what & 1) {
            p += P.pairs[b];
            m = P.maxcnt[b] > m ? P.maxcnt[b] : m;
            o |= P.over[b];
        }
        if (what & 2) d = P.dtmin[b] < d ? P.dtmin[b] : d;
    }
    const int t = threadIdx.x;
    sh.p[t] = p;
    sh.m[t] = m;
    sh.o[t] = o;
    sh.d[t] = d;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (t < off) {
            sh.p[t] += sh.p[t + off];
            sh.m[t] = sh.m[t + off] > sh.m[t] ? sh.m[t + off] : sh.m[t];
            sh.o[t] |= sh.o[t + off];
            sh.d[t] = sh.d[t + off] < sh.d[t] ? sh.d[t + off] : sh.d[t];
        }
        __syncthreads();
    }
    if (t == 0) {
        if (what & 1) {
            s->pairs_last += sh.p[0];
            if (what & 4) s->pairs_total += sh.p[0];
            s->maxcnt = sh.m[0] > s->maxcnt ? sh.m[0] : s->maxcnt;
            s->overflow |= sh.o[0];
        }
        if (what & 2) s->dtmin_bits = (unsigned long long)__double_as_longlong(sh.d[0]);
    }
}

static __global__ __launch_bounds__(kBlock) void k_reduce_partials(Partials P, Scalars* s, int what) {
    __shared__ ReduceScratch sh;
    reduce_partials_block(P, s, what, sh);
}

// One barrier for the block's reductions of a force pass: the pair count
// (sum and max) into the block's partial slots, and, with with_dt, the
// minimum of the positive dt bounds (they order like their bit patterns)
// into its dt slot. Block 0 marks the pass pending (k_scalars or
// k_fold_counts folds it). Every thread of the block calls it.
__device__ __forceinline__ void block_count_store(uint32_t cnt, const Partials& P, int64_t blk, bool with_dt,
                                                  double dt_cand) {
    __shared__ unsigned long long sh_sum[kBlock / 64];
    __shared__ uint32_t sh_max[kBlock / 64];
    __shared__ unsigned long long sh_dt[kBlock / 64];
    unsigned long long v = cnt;
    uint32_t mx = cnt;
    unsigned long long b = (unsigned long long)__double_as_longlong(dt_cand);
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_down(v, off, 64);
        const uint32_t o = __shfl_down(mx, off, 64);
        mx = o > mx ? o : mx;
        if (with_dt) {
            const unsigned long long ob = __shfl_down(b, off, 64);
            b = ob < b ? ob : b;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        sh_sum[wv] = v;
        sh_max[wv] = mx;
        sh_dt[wv] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0, tb = sh_dt[0];
        uint32_t m = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            t += sh_sum[k];
            m = sh_max[k] > m ? sh_max[k] : m;
            tb = sh_dt[k] < tb ? sh_dt[k] : tb;
        }
        if (with_dt) P.dtmin[blk] = __longlong_as_double((long long)tb);
        P.pairs[blk] = t;
        P.maxcnt[blk] = m;
        if (blk == 0) *P.pending = 1u;
    }
}

// Neighbour lists (FP64 path and on-demand list download): every j with
// the cell-local FP32 predicate, in stencil order, slot t of all i stored
// contiguously.
// The candidate windows of one particle in the lockstep row scan: row r of
// the (2R+1)^(D-1) rows of cells around cell c (z-major, then y) gives the
// cells [b0, b1) of that row that the support chord of radius hc reaches
// (b0 = b1 = 0 when none); the rows and the cells within a row come in the
// order of a z, y, x walk over the stencil. The row's end cells are open
// outward (axis_gap), so a chord reaching past one keeps it.
template <int D, int R>
struct RowWindows {
    static constexpr int S = 2 * R + 1;
    static constexpr int NR = D == 3 ? S * S : S;
    const Grid& g;
    const int* c;
    const float* xi;
    float hc2, invLf;
    bool live;

    __device__ __forceinline__ void bounds(int r, uint32_t& b0, uint32_t& b1) const {
        const int dy = r % S - R, dz = D == 3 ? r / S - R : 0;
        const int cy = c[1] + dy, cz = c[2] + dz;
        b0 = b1 = 0;
        if (live && cy >= 0 && cy < g.nc[1] && (D == 2 || (cz >= 0 && cz < g.nc[2]))) {
            const float gy = axis_gap(dy, xi[1], g.Lf, cy, g.nc[1]);
            const float gz = D == 3 ? axis_gap(dz, xi[2], g.Lf, cz, g.nc[2]) : 0.f;
            const float g2 = gy * gy + gz * gz;
            if (g2 < hc2) {
                const float w = sqrtf(hc2 - g2);
                int xa = (int)floorf((xi[0] - w) * invLf), xb = (int)floorf((xi[0] + w) * invLf);
                const int kmin = -c[0], kmax = g.nc[0] - 1 - c[0];
                xa = min(max(max(xa, -R), kmin), kmax);
                xb = max(min(min(xb, R), kmax), kmin);
                if (xa <= xb) {
                    const uint32_t row = D == 3 ? (uint32_t)((cz * g.nc[1] + cy) * g.nc[0]) : (uint32_t)(cy * g.nc[0]);
                    b0 = row + c[0] + xa;
                    b1 = row + c[0] + xb + 1;
                }
            }
        }
    }
};

template <int D, int R>
__global__ __launch_bounds__(kBlock) void k_neigh(Packed pk, const uint32_t* __restrict__ start, Grid g,
                                                  int64_t n, int K, float hmax, uint32_t* __restrict__ nbr,
                                                  uint32_t* __restrict__ cnt, Partials P, const Scalars* s) {
    // the lockstep row scan (RowWindows); lists come out in the order of a
    // z, y, x walk over the stencil's cells, index order within a cell
    constexpr int U = 4;  // candidates loaded per trip
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    const bool failed = sim_failed(s);
    const bool live = i < n && !failed;
    const int64_t ii = live ? i : 0;
    const float4 xi4 = pk.px[ii];
    const float xi[3] = {xi4.x, xi4.y, xi4.z};
    const float cxi = xi4.w;
    const float hi = pk.pq[ii].w;
    int c[3] = {0, 0, 0};
    key_cell<D>(g, pk.pcell[ii], c);
    const float hc = (hi + hmax) * 1.0001f;
    const RowWindows<D, R> rw{g, c, xi, hc * hc, 1.f / g.Lf, live};
    uint32_t c_out = 0;
    uint32_t nb0, nb1;
    rw.bounds(0, nb0, nb1);
    uint32_t np0 = nb1 ? start[nb0] : 0u, npe = nb1 ? start[nb1] : 0u;
    for (int r = 0; r < RowWindows<D, R>::NR; ++r) {
        const int dy = r % RowWindows<D, R>::S - R, dz = D == 3 ? r / RowWindows<D, R>::S - R : 0;
        const uint32_t p0 = np0, len = npe - np0;
        if (r + 1 < RowWindows<D, R>::NR) {  // the next row's range, fetched behind this row's scan
            rw.bounds(r + 1, nb0, nb1);
            np0 = nb1 ? start[nb0] : 0u;
            npe = nb1 ? start[nb1] : 0u;
        }
        for (uint32_t t = 0; __any(t < len); t += U) {
            float4 pj[U];
            float hj[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t j = len ? p0 + min(t + u, len - 1) : 0u;  // in bounds when empty
                pj[u] = pk.px[j];
                hj[u] = pk.pq[j].w;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float dc[3] = {cxi - pj[u].w, (float)(-dy), (float)(-dz)};
                if (t + u < len && nbr_pred<D>(xi, hi, pj[u], dc, g.Lf, hj[u])) {
                    if ((int)c_out < K) nbr[(int64_t)c_out * n + i] = p0 + t + u;
                    ++c_out;
                }
            }
        }
    }
    if (i < n) cnt[i] = c_out;
    block_count_partial(c_out, c_out > (uint32_t)K ? 1u : 0u, P, blockIdx.x);
}

// Fused neighbour + force pass (FP32). The wave walks the (2R+1)^(D-1)
// rows of cells around its particles in lockstep (R = cells per support);
// per row each lane clips its own x window to the chord of its support
// sphere, so with R = 2 about 2.5 candidates are tested per neighbour.
// Hits are queued in LDS (index | row << 27); whenever some lane's queue is
// full the wave evaluates all queued pairs. No neighbour list reaches HBM.
// The summation order per i is the scan order, so results are deterministic.
// The final stage of a step for one particle: gravity along the last axis
// (src/init/Dam2.h:104-108), finalKick v = v_half + dt/2 a of fluid
// particles when kick != 0 (src/class.h:160-163), the density floor
// (src/init/Dam2.h:111) and the particle's bound in getTimeStepGlobal
// (src/integral.h:2-12: min of dt_i, 0.25 sqrt(h/|a|), 0.125 h^2/nu over
// fluid particles). a holds the hydro acceleration on entry and the total
// on return; the bound is returned (1e30 for walls).
struct FinalArgs {
    Master m;
    double g, nu, rho_floor;
    int kick;
    int wall_mode;  // 1: wall accelerations are written as zero (frozen walls)
};

template <int D>
__device__ __forceinline__ double final_particle(const FinalArgs& f, const Scalars* s, int64_t i, int64_t n, double* a,
                                                 double dti) {
    a[D - 1] -= f.g;
    const bool fluid = f.m.type[i] == 0;
    const double hdt = 0.5 * s->dt;
    double a2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        a2 += a[d] * a[d];
        if (f.kick && fluid) f.m.vel[d * n + i] = f.m.vh[d * n + i] + hdt * a[d];
    }
    if (f.m.dens[i] < f.rho_floor) f.m.dens[i] = f.rho_floor;
    double cand = 1.0e30;
    if (fluid) {
        const double h = f.m.h[i];
        const double t1 = 0.25 * sqrt(h / sqrt(a2));
        const double t2 = 0.125 * h * h / f.nu;
        cand = dti;
        cand = t1 < cand ? t1 : cand;
        cand = t2 < cand ? t2 : cand;
    } else if (f.wall_mode) {
#pragma unroll
        for (int d = 0; d < D; ++d) a[d] = 0.0;
    }
    return cand;
}

// final_particle for one smoothing length hf with the viscous bound t2 =
// 0.125 h^2 / nu formed on the host; the acceleration bound 0.25 sqrt(h/|a|)
// is taken in FP32 (the kick and the density floor stay FP64).
template <int D>
__device__ __forceinline__ double final_particle_h(const FinalArgs& f, const Scalars* s, int64_t i, int64_t n,
                                                   double* a, double dti, float hf, double t2) {
    a[D - 1] -= f.g;
    const bool fluid = f.m.type[i] == 0;
    const double hdt = 0.5 * s->dt;
    float a2 = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float af = (float)a[d];
        a2 = fmaf(af, af, a2);
        if (f.kick && fluid) f.m.vel[d * n + i] = f.m.vh[d * n + i] + hdt * a[d];
    }
    if (f.m.dens[i] < f.rho_floor) f.m.dens[i] = f.rho_floor;
    double cand = 1.0e30;
    if (fluid) {
        const double t1 = (double)(0.25f * __builtin_amdgcn_sqrtf(hf * __builtin_amdgcn_rsqf(a2)));
        cand = dti;
        cand = t1 < cand ? t1 : cand;
        cand = t2 < cand ? t2 : cand;
    } else if (f.wall_mode) {
#pragma unroll
        for (int d = 0; d < D; ++d) a[d] = 0.0;
    }
    return cand;
}

// Minimum over the block of positive doubles (they order like their bit
// patterns) into *slot. Every thread of the block calls it.
__device__ __forceinline__ void block_min_store(double cand, double* slot) {
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
        *slot = __longlong_as_double((long long)t);
    }
}

struct NFArgs {
    Packed pk;
    const uint32_t* start;
    Grid g;
    int64_t n;
    float hmax, nu;
    double c_cfl;
    PairConsts<float> k;
    double* acc;
    double* divv;
    double* dti;
    Partials P;
    Scalars* s;
    int fused_final;  // run the final stage (FinalArgs) in the epilogue
    FinalArgs fin;
    // outputs of cell-order particle i go to index out_index[i] (null: i)
    const uint32_t* out_index;
    // frozen walls (wall_mode 1): the particle types in cell order; null: off
    const uint8_t* wtype;
};


// Wave-wide maximum of an unsigned value over a fully active wave, returned
// as a wave-uniform (scalar) value: row shifts 1, 2, 4, 8 leave each 16-lane
// row's maximum in its lane 15, row_bcast:15 / row_bcast:31 carry it across
// rows, and lane 63 ends with the maximum of all 64 lanes.
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
    constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;
    int x = (int)v;
    x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(0, x, kRowShr1, 0xf, 0xf, false));
    x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(0, x, kRowShr2, 0xf, 0xf, false));
    x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(0, x, kRowShr4, 0xf, 0xf, false));
    x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(0, x, kRowShr8, 0xf, 0xf, false));
    x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(0, x, kRowBcast15, 0xa, 0xf, false));
    x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(0, x, kRowBcast31, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane(x, 63);
}

constexpr int kScanUnroll = 4;  // candidate loads in flight per lane

template <bool B>
struct BoolTag {  // a compile-time flag passed to a generic lambda
    static constexpr bool value = B;
};

template <int D, int R, int Q, bool UH, int VS>
__global__ __launch_bounds__(kBlock) void k_nforce(NFArgs a) {
    constexpr int S = 2 * R + 1;
    constexpr int NR = D == 3 ? S * S : S;  // rows of cells along x
    constexpr int kShift = 27;
    constexpr uint32_t kMask = (1u << kShift) - 1u;
    __shared__ uint32_t qbuf[kBlock / 64][64 * (Q + 1)];
    __shared__ float2 rowdc[NR];  // (c_i - c_j) along y and z for each row
    if ((int)threadIdx.x < NR)
        rowdc[threadIdx.x] = make_float2((float)(R - (int)threadIdx.x % S), D == 3 ? (float)(R - (int)threadIdx.x / S) : 0.f);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    uint32_t* q = &qbuf[threadIdx.x >> 6][lane * (Q + 1)];
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    const bool failed = sim_failed(a.s);
    const bool live = i < a.n && !failed;
    const float4 xi4 = live ? a.pk.px[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vi4 = live ? a.pk.pv[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 qi4 = live ? a.pk.pq[i] : make_float4(1.f, 0.f, 1.f, 1.f);
    const float xi[3] = {xi4.x, xi4.y, xi4.z};
    const float vi[3] = {vi4.x, vi4.y, vi4.z};
    const float cxi = xi4.w;
    const IState<float> I{qi4.w, qi4.x, qi4.y, qi4.z, a.nu};
    const float Lf = a.g.Lf, invLf = 1.0f / Lf;
    int c[3] = {0, 0, 0};
    if (live) key_cell<D>(a.g, a.pk.pcell[i], c);
    const float hc = (UH ? 2.f * I.h : I.h + a.hmax) * 1.0001f;
    const float hc2 = hc * hc;
    UniformH uh;
    if (UH) {
        uh.h = I.h;
        uh.eps_h2 = a.k.eps * I.h * I.h;
        uh.invH = 0.5f / I.h;
        const float iH2 = uh.invH * uh.invH;
        uh.invHD1 = D == 3 ? iH2 * iH2 : iH2 * uh.invH;
    }
    float acc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] = 0.f;
    float divv = 0.f, vsig = 0.f;
    uint32_t qn = 0, total = 0;

    auto flush = [&]() {
        for (uint32_t t = 0; __any(t < qn); ++t) {
            if (t < qn) {
                const uint32_t e = q[t];
                const uint32_t j = e & kMask;
                const int r = (int)(e >> kShift);
                const float4 pj = a.pk.px[j];
                const float4 vj = a.pk.pv[j];
                const float4 qj = a.pk.pq[j];
                const float2 dyz = rowdc[r];
                const float dc[3] = {cxi - pj.w, dyz.x, dyz.y};
                const float pjv[3] = {pj.x, pj.y, pj.z};
                const float vjv[3] = {vj.x, vj.y, vj.z};
                float dr[D], dv[D];
                float r2 = 0.f, drdv = 0.f;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    dr[d] = fmaf(dc[d], Lf, xi[d]) - pjv[d];
                    dv[d] = vi[d] - vjv[d];
                    r2 += dr[d] * dr[d];
                    drdv += dr[d] * dv[d];
                }
                const JState<float> J{qj.w, qj.x, qj.y, qj.z, a.nu, vj.w};
                if (UH) {
                    pair_uniform<D, VS>(dr, dv, r2, drdv, I, J, a.k, uh, acc, divv, vsig);
                } else {
                    float mu;
                    const float den = pair_mu(r2, drdv, I, J.h, a.k, mu, vsig);
                    pair_force<D>(dr, dv, r2, drdv, den, mu, I, J, a.k, acc, divv);
                }
            }
        }
        total += qn;
        qn = 0;
    };

    // candidate range of row r for this lane: indices [start[b0], start[b1])
    auto row_bounds = [&](int r, uint32_t& b0, uint32_t& b1) {
        const int dy = r % S - R, dz = D == 3 ? r / S - R : 0;
        const int cy = c[1] + dy, cz = c[2] + dz;
        b0 = b1 = 0;
        if (live && cy >= 0 && cy < a.g.nc[1] && (D == 2 || (cz >= 0 && cz < a.g.nc[2]))) {
            const float gy = axis_gap(dy, xi[1], Lf, cy, a.g.nc[1]);
            const float gz = D == 3 ? axis_gap(dz, xi[2], Lf, cz, a.g.nc[2]) : 0.f;
            const float g2 = gy * gy + gz * gz;
            if (g2 < hc2) {
                const float w = sqrtf(hc2 - g2);
                int xa = (int)floorf((xi[0] - w) * invLf), xb = (int)floorf((xi[0] + w) * invLf);
                // the row's end cells are open outward (axis_gap): a chord
                // that reaches past one keeps it
                const int kmin = -c[0], kmax = a.g.nc[0] - 1 - c[0];
                xa = min(max(max(xa, -R), kmin), kmax);
                xb = max(min(min(xb, R), kmax), kmin);
                if (xa <= xb) {
                    const uint32_t row = D == 3 ? (uint32_t)((cz * a.g.nc[1] + cy) * a.g.nc[0])
                                                : (uint32_t)(cy * a.g.nc[0]);
                    b0 = row + c[0] + xa;
                    b1 = row + c[0] + xb + 1;
                }
            }
        }
    };
    // The rows' candidates; with walls_only (a wave of frozen walls, below)
    // only fluid candidates are kept: a wall-wall pair adds exactly zero to
    // div v (dv = 0) and to v_sig (mu = 0), and a wall's acceleration is
    // not used. The kept candidates keep their order, so every sum is the
    // same.
    auto scan_rows = [&](auto walls_only) {
        constexpr bool WO = decltype(walls_only)::value;
        uint32_t nb0, nb1;
        row_bounds(0, nb0, nb1);
        uint32_t np0 = nb1 ? a.start[nb0] : 0u, npe = nb1 ? a.start[nb1] : 0u;
        for (int r = 0; r < NR; ++r) {
            const int dy = r % S - R, dz = D == 3 ? r / S - R : 0;
            const uint32_t p0 = np0, len = npe - np0;
            if (r + 1 < NR) {  // prefetch the next row's range behind this row's scan
                row_bounds(r + 1, nb0, nb1);
                np0 = nb1 ? a.start[nb0] : 0u;
                npe = nb1 ? a.start[nb1] : 0u;
            }
            // the y/z part of the predicate is fixed for the row
            const float yi = fmaf((float)(-dy), Lf, xi[1]);
            const float zi = D == 3 ? fmaf((float)(-dz), Lf, xi[2]) : 0.f;
            const float HH = UH ? __fmul_rn(__fadd_rn(I.h, I.h), __fadd_rn(I.h, I.h)) : 0.f;
            for (uint32_t t = 0; __any(t < len); t += kScanUnroll) {
                // issue all loads of the group before testing any of them
                float4 pj[kScanUnroll];
                float hj[kScanUnroll];
                uint32_t tj[kScanUnroll];
#pragma unroll
                for (int u = 0; u < kScanUnroll; ++u) {
                    // an empty window may start at n (cells past the last particle): read record 0
                    const uint32_t j = len ? p0 + min(t + u, len - 1) : 0u;
                    pj[u] = a.pk.px[j];
                    hj[u] = UH ? I.h : a.pk.pq[j].w;
                    tj[u] = WO ? (uint32_t)a.wtype[j] : 0u;
                }
#pragma unroll
                for (int u = 0; u < kScanUnroll; ++u) {
                    const float dx = __fsub_rn(fmaf(cxi - pj[u].w, Lf, xi[0]), pj[u].x);
                    const float dy = __fsub_rn(yi, pj[u].y);
                    float r2 = __fmul_rn(dx, dx);
                    r2 = fmaf(dy, dy, r2);
                    if (D == 3) r2 = fmaf(__fsub_rn(zi, pj[u].z), __fsub_rn(zi, pj[u].z), r2);
                    float H2 = HH;
                    if (!UH) {
                        const float H = __fadd_rn(I.h, hj[u]);
                        H2 = __fmul_rn(H, H);
                    }
                    if (t + u < len && r2 > 0.f && r2 < H2 && (!WO || tj[u] == 0u)) {
                        q[qn] = (p0 + t + u) | ((uint32_t)r << kShift);
                        ++qn;
                    }
                }
                if (__any(qn + kScanUnroll > (uint32_t)Q)) flush();
            }
        }
    };
    // frozen walls (wall_mode 1): a wave whose particles are all walls
    if (a.wtype != nullptr && __all(!live || a.wtype[i] != 0)) scan_rows(BoolTag<true>{});
    else scan_rows(BoolTag<false>{});
    flush();
    double cand = failed ? 0.0 : 1.0e30;  // a diverged state stops the clock
    if (live) {
        const int64_t n = a.n;
        double ad[D];
        const bool frozen = a.wtype != nullptr && a.wtype[i] != 0;  // a frozen wall's acceleration reads as zero
#pragma unroll
        for (int d = 0; d < D; ++d) ad[d] = frozen ? 0.0 : (double)acc[d];
        const double dti = a.c_cfl * (double)I.h / ((double)I.c + (double)vsig);
        const int64_t o = a.out_index ? (int64_t)a.out_index[i] : i;
        a.divv[o] = (double)divv / (double)I.rho;
        a.dti[o] = dti;
        if (a.fused_final) cand = final_particle<D>(a.fin, a.s, i, n, ad, dti);
#pragma unroll
        for (int d = 0; d < D; ++d) a.acc[d * n + o] = ad[d];
    }
    block_count_store(total, a.P, blockIdx.x, a.fused_final != 0, cand);
}

// Shepard density filter (ApplyFilter, src/force.h:10-30, with the Gather
// tree of src/main.cpp:65): for every i,
//   rho_i <- sum_j m_j W(r_ij, h_i) / sum_j (m_j / rho_j) W(r_ij, h_i)
// over all j within 2 h_i, i itself included, in FP64 from the master state
// (cell order), gathered over the cell stencil. The result goes to out; it is
// committed by k_apply_density once every particle has read the old values.
template <int D, int R>
__global__ __launch_bounds__(kBlock) void k_shepard(Master m, Packed pk, const uint32_t* __restrict__ start, Grid g,
                                                    int64_t n, double cnorm, double* __restrict__ out,
                                                    const Scalars* sc) {
    // the lockstep row scan (RowWindows) over the cells of i's support;
    // candidates come in the order of a z, y, x walk over the stencil's
    // cells, and the FP64 predicate and sums are the oracle's
    constexpr int U = 4;  // candidates loaded per trip
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    const bool failed = sim_failed(sc);
    const bool live = i < n && !failed;
    const int64_t ii = live ? i : 0;
    double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = m.pos[d * n + ii];
    const double H = 2.0 * m.h[ii], H2 = H * H, invH = 1.0 / H;
    double norm = cnorm;  // C_D / H^D (src/kernel.h:42-48)
#pragma unroll
    for (int d = 0; d < D; ++d) norm *= invH;
    // the cell and cell-local coordinates the binning pass gave i (the same
    // FP64 arithmetic on the same position)
    int c[3] = {0, 0, 0};
    bool esc = false;
    cell_of<D>(g, xi, c, esc);
    float xl[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < D; ++d) xl[d] = cell_local(xi[d], g.lo[d], c[d], g.L);
    const float hc = (float)H * 1.0001f;
    const RowWindows<D, R> rw{g, c, xl, hc * hc, 1.f / g.Lf, live};
    double one = 0.0, sum = 0.0;
    uint32_t nb0, nb1;
    rw.bounds(0, nb0, nb1);
    uint32_t np0 = nb1 ? start[nb0] : 0u, npe = nb1 ? start[nb1] : 0u;
    for (int r = 0; r < RowWindows<D, R>::NR; ++r) {
        const uint32_t p0 = np0, len = npe - np0;
        if (r + 1 < RowWindows<D, R>::NR) {  // the next row's range, fetched behind this row's scan
            rw.bounds(r + 1, nb0, nb1);
            np0 = nb1 ? start[nb0] : 0u;
            npe = nb1 ? start[nb1] : 0u;
        }
        for (uint32_t t = 0; __any(t < len); t += U) {
            double xj[U][3];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t j = len ? p0 + min(t + u, len - 1) : 0u;  // in bounds when empty
#pragma unroll
                for (int d = 0; d < D; ++d) xj[u][d] = m.pos[d * n + j];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                double r2 = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double dd = xi[d] - xj[u][d];
                    r2 += dd * dd;
                }
                if (t + u < len && r2 < H2) {
                    const uint32_t j = p0 + t + u;
                    const double s = sqrt(r2) * invH;
                    const double a = 1.0 - s, b = 0.5 - s;
                    double w = a * a * a;
                    if (b > 0.0) w -= 4.0 * b * b * b;
                    w *= norm;
                    const double mj = m.mass[j];
                    one += w * mj / m.dens[j];
                    sum += mj * w;
                }
            }
        }
    }
    if (i < n) out[i] = failed ? m.dens[i] : sum / one;  // a diverged state keeps its densities
}

// The same filter in FP32 arithmetic for the FP32 and mixed paths, from the
// FP32 force records of the current cell order (cell-local positions, so
// differences lose nothing to the magnitude of the coordinates). The
// density carries its information in rho - rho0 (a few parts in 1e3), so
// the filter is taken in that form: with q_j = rho0 / rho_j - 1 (formed as
// (rho0 - rho_j) / rho_j, the difference exact in FP64),
//   sum_j (m_j / rho_j) W_ij = (S + T) / rho0,  S = sum m_j W_ij,
//   T = sum m_j W_ij q_j,  so  rho_i' = rho0 / (1 + T / S),
// and FP32 rounding of S and T moves T / S, and rho' - rho0, only by parts
// in 1e7. W's normalisation is common to both sums and left out.
//
// The scan is the force kernels': the wave walks the (2R+1)^(D-1) rows of
// cells around its particles in lockstep, each lane clipping its own x
// window to the chord of its support sphere (2 h_i, the filter's support,
// src/force.h:20-26), and every candidate in the window is evaluated in
// place (the body is short, so no queue). Self is included (r = 0).
template <int D, int R>
__global__ __launch_bounds__(kBlock) void k_shepard32(Packed pk, const uint32_t* __restrict__ start, Grid g,
                                                      int64_t n, double rho0, const double* __restrict__ dens,
                                                      double* __restrict__ out, const Scalars* sc) {
    constexpr int S = 2 * R + 1;
    constexpr int NR = D == 3 ? S * S : S;
    constexpr int U = 4;  // candidates loaded per trip
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    const bool failed = sim_failed(sc);
    const bool live = i < n && !failed;
    const int64_t ii = live ? i : 0;
    const float4 xi4 = pk.px[ii];
    const float xi[3] = {xi4.x, xi4.y, xi4.z};
    const float cxi = xi4.w;
    const float H = 2.f * pk.pq[ii].w, H2 = H * H, invH = 1.f / H;
    const float Lf = g.Lf;
    const float hc = H * 1.0001f;
    int c[3] = {0, 0, 0};
    key_cell<D>(g, pk.pcell[ii], c);
    const RowWindows<D, R> rw{g, c, xi, hc * hc, 1.f / Lf, live};
    float Ssum = 0.f, Tsum = 0.f;
    uint32_t nb0, nb1;
    rw.bounds(0, nb0, nb1);
    uint32_t np0 = nb1 ? start[nb0] : 0u, npe = nb1 ? start[nb1] : 0u;
    for (int r = 0; r < NR; ++r) {
        const int dy = r % S - R, dz = D == 3 ? r / S - R : 0;
        const uint32_t p0 = np0, len = npe - np0;
        if (r + 1 < NR) {  // the next row's range, fetched behind this row's scan
            rw.bounds(r + 1, nb0, nb1);
            np0 = nb1 ? start[nb0] : 0u;
            npe = nb1 ? start[nb1] : 0u;
        }
        const float yi = fmaf((float)(-dy), Lf, xi[1]);
        const float zi = D == 3 ? fmaf((float)(-dz), Lf, xi[2]) : 0.f;
        for (uint32_t t = 0; __any(t < len); t += U) {
            float4 pj[U];
#pragma unroll
            for (int u = 0; u < U; ++u) pj[u] = pk.px[len ? p0 + min(t + u, len - 1) : 0u];  // in bounds when empty
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float dx = fmaf(cxi - pj[u].w, Lf, xi[0]) - pj[u].x;
                const float dyv = yi - pj[u].y;
                float r2 = dx * dx + dyv * dyv;
                if (D == 3) {
                    const float dzv = zi - pj[u].z;
                    r2 += dzv * dzv;
                }
                if (t + u < len && r2 < H2) {
                    const uint32_t j = p0 + t + u;
                    const float s = sqrtf(r2) * invH;
                    const float a = 1.f - s, b = 0.5f - s;
                    float w = a * a * a;
                    if (b > 0.f) w -= 4.f * b * b * b;
                    const double rj = dens[j];
                    const float q = (float)(rho0 - rj) / (float)rj;
                    const float mw = pk.pv[j].w * w;
                    Ssum += mw;
                    Tsum += mw * q;
                }
            }
        }
    }
    if (i < n) out[i] = failed ? dens[i] : rho0 / (1.0 + (double)Tsum / (double)Ssum);  // a diverged state keeps its densities
}

// Commit the filtered density (copyFromForce, src/class.h:116-118) and redo
// what k_reorder derived from it: P/rho^2 and c in the force records (and
// the FP64 copies and the mixed state records when present).
template <int D, class Extra>
__global__ void k_apply_density(Master m, Packed pk, const double* __restrict__ rho_new, double* __restrict__ pr64,
                                double* __restrict__ c64, EosConsts e, Extra extra, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double rho = rho_new[p];
    m.dens[p] = rho;
    double pr, cs;
    tait(e, rho, pr, cs);
    if (Extra::kFp32Records) {
        const float4 q = pk.pq[p];
        pk.pq[p] = make_float4((float)rho, (float)pr, (float)cs, q.w);
    }
    if (pr64) {
        pr64[p] = pr;
        c64[p] = cs;
    }
    extra.state(p, rho, pr, cs);
}

// FP64 force pass over neighbour lists (the GPU oracle path): pair
// differences and all arithmetic in FP64 from the master arrays.
struct ForceArgs64 {
    const double* pos;
    const double* vel;
    const double* h;
    const double* rho;
    const double* m;
    const double* pr;
    const double* c;
    const uint32_t* nbr;
    const uint32_t* cnt;
    double* acc;
    double* divv;
    double* dti;
    int64_t n;
    int K;
    double nu;
    double c_cfl;
    PairConsts<double> k;
    const uint8_t* wtype;  // frozen walls (wall_mode 1): wall accelerations written as zero; null: off
};

template <int D>
__global__ __launch_bounds__(kBlock) void k_force64(ForceArgs64 a) {
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int64_t n = a.n;
    double xi[D], vi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xi[d] = a.pos[d * n + i];
        vi[d] = a.vel[d * n + i];
    }
    const IState<double> I{a.h[i], a.rho[i], a.pr[i], a.c[i], a.nu};
    double acc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] = 0.0;
    double divv = 0.0, vsig = 0.0;
    const int cnt = min((int)a.cnt[i], a.K);
    for (int t = 0; t < cnt; ++t) {
        const uint32_t j = a.nbr[(int64_t)t * n + i];
        double dr[D], dv[D];
        double r2 = 0.0, drdv = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            dr[d] = xi[d] - a.pos[d * n + j];
            dv[d] = vi[d] - a.vel[d * n + j];
            r2 += dr[d] * dr[d];
            drdv += dr[d] * dv[d];
        }
        const JState<double> J{a.h[j], a.rho[j], a.pr[j], a.c[j], a.nu, a.m[j]};
        double mu;
        const double den = pair_mu(r2, drdv, I, J.h, a.k, mu, vsig);
        pair_force<D>(dr, dv, r2, drdv, den, mu, I, J, a.k, acc, divv);
    }
    const bool frozen = a.wtype != nullptr && a.wtype[i] != 0;
#pragma unroll
    for (int d = 0; d < D; ++d) a.acc[d * n + i] = frozen ? 0.0 : acc[d];
    a.divv[i] = divv / I.rho;                               // src/force.h:57
    a.dti[i] = a.c_cfl * I.h / (I.c + vsig);               // src/force.h:56
}

// The final stage (final_particle) as a pass of its own, for the FP64 path
// and for init, with the per-block dt bounds in P.dtmin.
template <int D>
__global__ __launch_bounds__(kBlock) void k_final(FinalArgs f, double* __restrict__ acc,
                                                  const double* __restrict__ dti, const Scalars* s, Partials P,
                                                  int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double cand = sim_failed(s) ? 0.0 : 1.0e30;  // a diverged state stops the clock
    if (i < n && cand > 0.0) {
        double a[D];
#pragma unroll
        for (int d = 0; d < D; ++d) a[d] = acc[d * n + i];
        cand = final_particle<D>(f, s, i, n, a, dti[i]);
        acc[(D - 1) * n + i] = a[D - 1];
        if (f.wall_mode && f.m.type[i] != 0)  // a frozen wall's acceleration reads as zero
            for (int d = 0; d < D - 1; ++d) acc[d * n + i] = 0.0;
    }
    block_min_store(cand, &P.dtmin[blockIdx.x]);
}

}  // namespace wcsph
