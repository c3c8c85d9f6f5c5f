// The device-resident simulation object behind the wcsph_sim_* entry
// points, shared by the step orchestration (sim.hip) and the grouped steps
// of the serving loop (group.hip).
#pragma once
#include <vector>

#include "common.hpp"
#include "group_kernels.hpp"
#include "mixed_kernels.hpp"
#include "sim_kernels.hpp"

// Device memory for a simulation's buffers: hipMalloc, or ranges of a
// caller's arena (the serving loop carves every request from one block).
struct wcsph_arena;

using wcsph::Grid;
using wcsph::Master;
using wcsph::Mixed;
using wcsph::MixedConsts;
using wcsph::Packed;
using wcsph::Partials;
using wcsph::Scalars;

struct wcsph_sim {
    wcsph_params p{};
    wcsph_sim_config cfg{};
    int D = 3;
    int64_t n = 0;
    int K = 128;
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    Grid grid{};
    int64_t ncell = 0, npart = 0;
    Master m[2]{};
    int cur = 0;
    double *acc = nullptr, *divv = nullptr, *dti = nullptr;
    Packed pk{};
    int R = 2;      // cells per support length
    int Q = 16;     // LDS queue depth of the fused pass
    float hmax = 0.f;
    bool uniform_h = false;  // all smoothing lengths equal in FP32 (fixed in time)
    double h_const = 0.0;    // > 0: all smoothing lengths are exactly this (kept in both buffers, not permuted)
    double m_const = 0.0;    // > 0: all masses are exactly this (kept in both buffers, not permuted)
    double *pr64 = nullptr, *c64 = nullptr;  // FP64 path only
    Mixed mx{};                               // mixed path only
    MixedConsts mc{};
    uint32_t *key = nullptr, *slot = nullptr, *count = nullptr, *start = nullptr, *perm = nullptr;
    uint32_t *tile_total = nullptr, *group_total = nullptr;  // scan tile and tile-group totals
    int64_t ngroups = 0;
    uint32_t *nbr = nullptr, *cnt = nullptr;
    Scalars* sc = nullptr;
    Partials P{};
    double* diag_part = nullptr;
    double* diag_out = nullptr;
    bool uploaded = false, forced = false;
    bool steppable = false;  // acc, the dt bounds and the clock are those of init or a step
    bool fuse_final = false;  // the next fused force pass also runs the final stage of a step
    bool timing = false;
    bool use_graphs = true;
    int64_t hstep = 0;         // steps queued since init (the Shepard filter schedule)
    double* rho_tmp = nullptr;  // filtered densities before they are committed
    hipGraphExec_t graph2[2] = {nullptr, nullptr};  // two steps starting from cur = 0 / 1
    std::vector<hipEvent_t> tev;  // 5 events per timed step, pooled
    int64_t tsteps = 0;
    std::vector<std::pair<void*, size_t>> allocs;  // device buffers and their sizes
    // pinned copy of the device failure flags (non-finite positions, overfull
    // cells), refreshed after every call that queues steps
    volatile uint32_t* hflags = nullptr;

    wcsph_arena* arena = nullptr;  // buffers from this arena instead of hipMalloc
    bool own_stream = true;        // false: st belongs to the caller (a serving group)
    bool own_flags = true;         // false: hflags belongs to the caller

    void* alloc_bytes(size_t bytes);
    template <class T>
    T* alloc(int64_t count) {
        return (T*)alloc_bytes((size_t)std::max<int64_t>(count, 1) * sizeof(T));
    }
    ~wcsph_sim();
    int64_t nb(int64_t items) const { return wcsph::ceil_div(items, wcsph::kBlock); }
};


// Arena of device memory carved into 256-byte aligned ranges (first fit,
// freed ranges merged with their neighbours). Thread-compatible, not
// thread-safe.
struct wcsph_arena {
    char* base = nullptr;
    size_t cap = 0;
    std::vector<std::pair<size_t, size_t>> free_list;  // (offset, bytes), sorted by offset
    size_t in_use = 0;

    void reset(char* b, size_t c) {
        base = b;
        cap = c;
        free_list.assign(1, {0, c});
        in_use = 0;
    }
    void* take(size_t bytes);           // nullptr when no free range is large enough
    void give(void* p, size_t bytes);   // return a range take() handed out
    size_t largest_free() const;
};

namespace wcsph {

// Device bytes a simulation of n particles on a grid of ncell cells takes
// at the given precision (K-slot neighbour lists: the FP64 path only): the
// sum of its allocations as wcsph_sim_create and the upload make them.
int64_t sim_footprint(int D, int64_t n, int precision, int K, int64_t ncell);

// The cell count of the grid the upload builds for this state (the host
// computation of build_grid, without a simulation); nc (optional) receives
// the cells per axis.
int64_t sim_grid_cells(const wcsph_params* p, const wcsph_sim_config* cfg, int64_t n, const double* pos,
                       const double* h, int32_t* nc = nullptr);

// Create a simulation whose buffers come from `arena` and whose work runs on
// the caller's stream `st`, with its failure-flag mirror at `hflags` (two
// pinned words). The caller keeps all three alive past wcsph_sim_destroy.
wcsph_sim* sim_create_in(const wcsph_params* p, const wcsph_sim_config* cfg, int64_t n, wcsph_arena* arena,
                         hipStream_t st, volatile uint32_t* hflags);

// Kernel arguments of one mixed-precision step of s from its current state
// (s is left as it was): what one_step launches, as records for the
// grouped kernels.
template <int D>
struct MixedStep {
    ScalarsArgs sc;
    BinArgs<D> bin;
    ScanArgs scan;
    ScatterArgs scat;
    ReorderArgs<D, MixedExtra<D>> ro;
    NFMixedArgs nf;
    int32_t nb, ngroups, npart;  // blocks of the particle, scan-group and scan-tile launches
};
template <int D>
MixedStep<D> mixed_step_args(wcsph_sim* s);

// A step of s as one_step queues it, on s's stream (the grouped steps use
// it for an odd step left over).
void sim_one_step(wcsph_sim* s);
// After a grouped launch of `steps` steps: advance s's host-side step
// bookkeeping and queue its failure-flag copy.
void sim_after_group_steps(wcsph_sim* s, int64_t steps);
// Throws the diverged-state error when s's flag mirror reports one.
void sim_check_failed(const wcsph_sim* s);

}  // namespace wcsph
