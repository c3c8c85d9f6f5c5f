// One stateless hydro force pass over caller-owned device arrays: the host
// side shared by wcsph_force_pass (library-owned workspace, grown with
// hipMalloc) and its variants that bring their own workspace memory.
#pragma once
#include <cstddef>
#include <cstdint>

#include "common.hpp"
#include "sim_kernels.hpp"

namespace wcsph {

// Byte offsets of the pass's workspace arrays (256-byte aligned).
struct PassLayout {
    size_t key, slot, perm, id, px, pv, pq, pcell, count, start, tile_total, group_total, bpairs, bmax, pending, sc,
        total;
};

// The cell grid of a pass and its size, from the caller's description
// (throws Error on a bad description).
struct PassGrid {
    Grid g;
    int64_t ncell;
};
PassGrid pass_grid(const wcsph_params* p, const wcsph_pass_desc* d);

constexpr PassLayout pass_layout(int64_t n, int64_t ncell) {
    PassLayout L{};
    size_t off = 0;
    auto take = [&off](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    L.key = take(4 * (size_t)n);
    L.slot = take(4 * (size_t)n);
    L.perm = take(4 * (size_t)n);
    L.id = take(4 * (size_t)n);
    L.px = take(16 * (size_t)n);
    L.pv = take(16 * (size_t)n);
    L.pq = take(16 * (size_t)n);
    L.pcell = take(4 * (size_t)n);
    // one bucket past the last cell: the dead bucket of non-finite positions
    L.count = take(4 * (size_t)(ncell + 1));
    L.start = take(4 * (size_t)(ncell + 2));
    L.tile_total = take(4 * (size_t)((ncell + 1 + kScanTile - 1) / kScanTile));
    L.group_total = take(4 * (size_t)kScanMaxGroups);
    // the force kernel's per-block pair counts and their pending flag
    L.bpairs = take(8 * (size_t)((n + kBlock - 1) / kBlock));
    L.bmax = take(4 * (size_t)((n + kBlock - 1) / kBlock));
    L.pending = take(4);
    L.sc = take(sizeof(Scalars));
    L.total = off;
    return L;
}

// Queue the pass on `st` with the workspace at `ws` (at least
// pass_layout(n, ncell).total bytes, 256-byte aligned, device memory).
// Every array the pass reads before writing is cleared by a kernel first,
// so the workspace may hold anything on entry.
void run_force_pass(const wcsph_params* p, const wcsph_pass_desc* d, const PassGrid& pg, void* ws,
                    const double* pos, const double* vel, const double* dens, const double* mass, const double* h,
                    double* acc, double* div_v, double* dt_i, int64_t* pairs, hipStream_t st);

// The device address of `device`'s copy of the no-allocation variant's
// scratch array (hipGetSymbolAddress once per device; throws Error).
void* noalloc_ptr(int device);

// Argument checks shared by the entry points.
void check_pass_args(const wcsph_params* p, const wcsph_pass_desc* d, const double* pos, const double* vel,
                     const double* dens, const double* mass, const double* h, double* acc, double* div_v,
                     double* dt_i);

}  // namespace wcsph
