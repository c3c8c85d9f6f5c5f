// Grouped step kernels: one launch runs a stage of the mixed-precision step
// for several simulations at once. Each simulation keeps its own buffers,
// grid and constants; a launch's grid is the concatenation of the
// simulations' grids (each started at a multiple of 8 blocks, so a block
// runs on the XCD its own index names), and a block finds its simulation
// in the GroupMap and its arguments in a device table of per-simulation
// argument records. Every block runs the same body as the ordinary kernel
// with its index within its simulation, so a simulation's results do not
// depend on which others share the launch.
#pragma once
#include "mixed_kernels.hpp"
#include "sim_kernels.hpp"

namespace wcsph {

constexpr int kGroupMax = 16;  // simulations per grouped launch

struct GroupMap {
    int32_t n;                     // simulations
    int32_t first[kGroupMax + 1];  // first block of each (multiples of 8); first[n] = grid size
    int32_t nblk[kGroupMax];       // blocks each simulation uses (the rest of its span exits)
};

// Argument tables are read through the constant address space, as kernel
// arguments are: the loads are invariant (scalar, and free to repeat near
// their uses), and the pointers they hold are known to address global
// memory (a table read through a generic pointer makes every buffer access
// of the body a flat access).
template <class T>
using ctab = const __attribute__((address_space(4))) T*;

// The record as a generic reference, loads still inferred constant-space
// (as a kernel argument's are).
template <class T>
__device__ __forceinline__ const T& args_at(ctab<T> p) {
    return *(const T*)p;
}

// The simulation of this block (uniform: scalar compares).
__device__ __forceinline__ int group_member(const GroupMap& gm) {
    const int b = (int)blockIdx.x;
    int k = 0;
    for (int j = 1; j < gm.n; ++j) k += b >= gm.first[j] ? 1 : 0;
    return k;
}

// Argument records of one step's launches (the ordinary launches take the
// same fields, so both run with identical arguments).
struct ScalarsArgs {
    Scalars* s;
    Partials P;  // the member's dt bounds and pending pair counts
    int32_t mode;
};

template <int D>
struct BinArgs {
    Master m;
    const double* acc;
    const double* divv;
    int32_t kick;
    int64_t n;
    Grid g;
    uint32_t* key;
    uint32_t* slot;
    uint32_t* count;
    Scalars* s;
};

struct ScanArgs {
    uint32_t* count;
    int64_t m;  // buckets (cells + the dead bucket)
    uint32_t* tile_total;
    uint32_t* group_total;
    int64_t gsize;
    uint32_t* start;
};

struct ScatterArgs {
    const uint32_t* key;
    const uint32_t* slot;
    const uint32_t* start;
    uint32_t* perm;
    uint32_t* count;
    int64_t n;
};

template <int D, class Extra>
struct ReorderArgs {
    const uint32_t* perm;
    const uint32_t* key;
    const uint32_t* start;
    Master src, dst;
    const double* acc_old;
    Scalars* s;
    int32_t step, keep_vel;
    Packed pk;
    Grid g;
    double* pr64;
    double* c64;
    EosConsts e;
    float v_scale;
    double h_const;
    double m_const;
    Extra extra;
    int64_t n;
};

static __global__ __launch_bounds__(kScalarsThreads) void kg_scalars(ctab<ScalarsArgs> t) {
    const ScalarsArgs& a = args_at(t + blockIdx.x);
    scalars_body(a.s, a.P, a.mode);
}

template <int D>
__global__ void kg_bin(ctab<BinArgs<D>> t, GroupMap gm) {
    const int k = group_member(gm);
    const int b = (int)blockIdx.x - gm.first[k];
    if (b >= gm.nblk[k]) return;
    const BinArgs<D>& a = args_at(t + k);
    bin_body<D>(b, a.m, a.acc, a.divv, a.kick, a.n, a.g, a.key, a.slot, a.count, a.s);
}

static __global__ __launch_bounds__(kBlock) void kg_scan_reduce(ctab<ScanArgs> t, GroupMap gm) {
    const int k = group_member(gm);
    const int b = (int)blockIdx.x - gm.first[k];
    if (b >= gm.nblk[k]) return;
    const ScanArgs& a = args_at(t + k);
    scan_reduce_body(b, a.count, a.m, a.tile_total, a.group_total, a.gsize);
}

static __global__ __launch_bounds__(kBlock) void kg_scan_apply(ctab<ScanArgs> t, GroupMap gm) {
    const int k = group_member(gm);
    const int b = (int)blockIdx.x - gm.first[k];
    if (b >= gm.nblk[k]) return;
    const ScanArgs& a = args_at(t + k);
    scan_apply_body(b, a.count, a.m, a.tile_total, a.group_total, a.gsize, a.start);
}

static __global__ void kg_scatter(ctab<ScatterArgs> t, GroupMap gm) {
    const int k = group_member(gm);
    const int b = (int)blockIdx.x - gm.first[k];
    if (b >= gm.nblk[k]) return;
    const ScatterArgs& a = args_at(t + k);
    scatter_body(b, a.key, a.slot, a.start, a.perm, a.count, a.n);
}

template <int D, class Extra>
__global__ __launch_bounds__(kBlock) void kg_reorder(ctab<ReorderArgs<D, Extra>> t, GroupMap gm) {
    const int k = group_member(gm);
    const int b = (int)blockIdx.x - gm.first[k];
    if (b >= gm.nblk[k]) return;
    const ReorderArgs<D, Extra>& a = args_at(t + k);
    reorder_body<D, Extra>(b, a.perm, a.key, a.start, a.src, a.dst, a.acc_old, a.s, a.step, a.keep_vel, a.pk, a.g,
                           a.pr64, a.c64, a.e, a.v_scale, a.h_const, a.m_const, a.extra,
                           a.n);
}

template <int D, int R, int VS, int BT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WCSPH_MIX_WAVES)))
void kg_nforce_mixed(ctab<NFMixedArgs> t, GroupMap gm) {
    const int k = group_member(gm);
    const int b = (int)blockIdx.x - gm.first[k];
    if (b >= gm.nblk[k]) return;
    const NFMixedArgs& a = args_at(t + k);
    nforce_mixed_body<D, R, VS, BT>(VBlock{b, gm.nblk[k]}, a);
}

}  // namespace wcsph
