// One stateless hydro force pass over caller-owned device arrays
// (wcsph_force_pass). The pass is the device-resident step's cell list and
// FP32 force kernel without the integrator around it:
//   k_scalars(reset)        per-pass counters and failure flags
//   k_fill_u32              cell counts to zero
//   k_bin (no kick)         cell key and slot of each particle
//   k_scan_* / k_scatter    cell starts and cell membership
//   k_pass_records          gather into cell order (ranked as in k_reorder),
//                           Tait EoS, FP32 force records, cell -> caller index
//   k_nforce (general h)    fused neighbour scan + pair body, outputs written
//                           to the caller's index
//   k_pass_pairs            pair count (or -1 for a diverged state) out
// Every array is cleared or written by a kernel before it is read, so the
// whole pass can be captured into a hipGraph and the workspace may hold
// anything on entry.
#include <cmath>
#include <mutex>

#include "pass.hpp"

namespace wcsph {

namespace {

constexpr int kPassR = 2;   // cells per support length
constexpr int kPassQ = 24;  // LDS queue slots per lane of k_nforce

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t m, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = v;
}

// Gather into cell order: slot p takes the member ranked_member names (as
// k_reorder does), its FP32 force records are written with the pressure
// and sound speed of its density (CalcPressure, src/force.h:3-9), and id[p]
// keeps its index in the caller's order.
template <int D>
__global__ __launch_bounds__(kBlock) void k_pass_records(const uint32_t* __restrict__ perm,
                                                         const uint32_t* __restrict__ key,
                                                         const uint32_t* __restrict__ start,
                                                         const double* __restrict__ pos,
                                                         const double* __restrict__ vel,
                                                         const double* __restrict__ dens,
                                                         const double* __restrict__ mass,
                                                         const double* __restrict__ h, int64_t n, Grid g,
                                                         EosConsts e, Packed pk, uint32_t* __restrict__ id,
                                                         Scalars* s) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = key[perm[p]];
    const uint2 cr = make_uint2(start[k], start[k + 1]);  // the cell's slot range
    const bool ranked = k != dead_key<D>(g) && cr.y - cr.x <= kRankCap;
    if (k != dead_key<D>(g) && cr.y - cr.x > kRankCap && (uint32_t)p == cr.x) atomicOr(&s->overfull, 1u);
    const uint32_t o = ranked ? ranked_member(perm, cr.x, cr.y, (uint32_t)p) : perm[p];
    int c[3] = {0, 0, 0};
    key_cell<D>(g, k, c);
    float xl[3] = {0.f, 0.f, 0.f}, vf[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xl[d] = cell_local(pos[d * n + o], g.lo[d], c[d], g.L);
        vf[d] = (float)vel[d * n + o];
    }
    const double rho = dens[o];
    double pr, cs;
    tait(e, rho, pr, cs);
    pk.px[p] = make_float4(xl[0], xl[1], xl[2], (float)c[0]);
    pk.pv[p] = make_float4(vf[0], vf[1], vf[2], (float)mass[o]);
    pk.pq[p] = make_float4((float)rho, (float)pr, (float)cs, (float)h[o]);
    pk.pcell[p] = k;
    id[p] = o;
}

__global__ void k_pass_pairs(const Scalars* s, int64_t* pairs) {
    if (threadIdx.x == 0) *pairs = sim_failed(s) ? -1 : (int64_t)s->pairs_last;
}

template <class T>
T* at(void* ws, size_t off) {
    return reinterpret_cast<T*>(static_cast<char*>(ws) + off);
}

template <int D, int VS>
void launch_pass_force(const NFArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k_nforce<D, kPassR, kPassQ, false, VS>), dim3(ceil_div(a.n, kBlock)), dim3(kBlock), 0, st,
                       a);
}

}  // namespace

void check_pass_args(const wcsph_params* p, const wcsph_pass_desc* d, const double* pos, const double* vel,
                     const double* dens, const double* mass, const double* h, double* acc, double* div_v,
                     double* dt_i) {
    WCSPH_REQUIRE(p && d, WCSPH_EINVAL, "null params or description");
    WCSPH_REQUIRE(p->dim == 2 || p->dim == 3, WCSPH_EINVAL, "dim must be 2 or 3");
    WCSPH_REQUIRE(p->rho0 > 0 && p->c0 > 0 && p->gamma > 0 && p->c_cfl > 0 && p->eps >= 0, WCSPH_EINVAL,
                  "rho0, c0, gamma, c_cfl must be positive and eps non-negative");
    WCSPH_REQUIRE(d->n > 0 && d->n < ((int64_t)1 << 27), WCSPH_EINVAL, "n must be in [1, 2^27)");
    WCSPH_REQUIRE(p->grad_mode == 0, WCSPH_EINVAL,
                  "the force pass takes the gradient at the mean smoothing length only (grad_mode 0)");
    WCSPH_REQUIRE(p->wall_mode == 0, WCSPH_EINVAL,
                  "the force pass evaluates every pair (wall_mode 0); wall_mode 1 is for the device-resident step");
    WCSPH_REQUIRE(pos && vel && dens && mass && h && acc && div_v && dt_i, WCSPH_EINVAL, "null array");
}

PassGrid pass_grid(const wcsph_params* p, const wcsph_pass_desc* d) {
    WCSPH_REQUIRE(std::isfinite(d->h_max) && d->h_max > 0, WCSPH_EINVAL, "h_max must be positive");
    PassGrid pg{};
    Grid& g = pg.g;
    // cell edge slightly above h_max (two cells per support 2 h_max), as the
    // device-resident simulation sizes its grid
    const double L = 2.0 * (double)(float)d->h_max * (1.0 + 1.0e-5) / kPassR;
    g.L = L;
    g.invL = 1.0 / L;
    g.Lf = (float)L;
    int64_t ncell = 1;
    for (int k = 0; k < 3; ++k) {
        if (k < p->dim) {
            WCSPH_REQUIRE(std::isfinite(d->box_lo[k]) && std::isfinite(d->box_hi[k]) && d->box_hi[k] > d->box_lo[k],
                          WCSPH_EINVAL, "box_hi must lie above box_lo on every axis");
            const double cells = std::ceil((d->box_hi[k] - d->box_lo[k]) / L);
            WCSPH_REQUIRE(cells < 2.0e9, WCSPH_EOVERFLOW, "cell grid exceeds 2^31 cells");
            g.lo[k] = d->box_lo[k];
            g.nc[k] = std::max(1, (int)cells);
        } else {
            g.lo[k] = 0.0;
            g.nc[k] = 1;
        }
        ncell *= g.nc[k];
        WCSPH_REQUIRE(ncell < ((int64_t)1 << 31), WCSPH_EOVERFLOW, "cell grid exceeds 2^31 cells");
    }
    pg.ncell = ncell;
    return pg;
}

void run_force_pass(const wcsph_params* p, const wcsph_pass_desc* d, const PassGrid& pg, void* ws,
                    const double* pos, const double* vel, const double* dens, const double* mass, const double* h,
                    double* acc, double* div_v, double* dt_i, int64_t* pairs, hipStream_t st) {
    const int D = p->dim;
    const int64_t n = d->n, ncell = pg.ncell;
    const PassLayout L = pass_layout(n, ncell);
    uint32_t* key = at<uint32_t>(ws, L.key);
    uint32_t* slot = at<uint32_t>(ws, L.slot);
    uint32_t* perm = at<uint32_t>(ws, L.perm);
    uint32_t* id = at<uint32_t>(ws, L.id);
    uint32_t* count = at<uint32_t>(ws, L.count);
    uint32_t* start = at<uint32_t>(ws, L.start);
    uint32_t* tile_total = at<uint32_t>(ws, L.tile_total);
    uint32_t* group_total = at<uint32_t>(ws, L.group_total);
    Scalars* sc = at<Scalars>(ws, L.sc);
    const Partials P{at<unsigned long long>(ws, L.bpairs), at<uint32_t>(ws, L.bmax), nullptr, nullptr,
                     ceil_div(n, kBlock), at<uint32_t>(ws, L.pending)};
    Packed pk{at<float4>(ws, L.px), at<float4>(ws, L.pv), at<float4>(ws, L.pq), at<uint32_t>(ws, L.pcell)};
    const int64_t npart = ceil_div(ncell + 1, kScanTile);
    const dim3 blk(kBlock), grid_n((unsigned)ceil_div(n, kBlock));
    hipLaunchKernelGGL(k_scalars, dim3(1), dim3(kScalarsThreads), 0, st, sc, P, 0);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)ceil_div(ncell + 1, kBlock)), blk, 0, st, count, ncell + 1, 0u);
    const int64_t gsize = scan_group_size(npart), ngroups = ceil_div(npart, gsize);
    Master m{};
    m.pos = const_cast<double*>(pos);
    if (D == 3)
        hipLaunchKernelGGL(k_bin<3>, grid_n, blk, 0, st, m, (const double*)nullptr, (const double*)nullptr, 0, n, pg.g,
                           key, slot, count, sc);
    else
        hipLaunchKernelGGL(k_bin<2>, grid_n, blk, 0, st, m, (const double*)nullptr, (const double*)nullptr, 0, n, pg.g,
                           key, slot, count, sc);
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)ngroups), blk, 0, st, count, ncell + 1, tile_total, group_total,
                       gsize);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)npart), blk, 0, st, count, ncell + 1, tile_total, group_total, gsize,
                       start);
    hipLaunchKernelGGL(k_scatter, grid_n, blk, 0, st, key, slot, start, perm, count, n);
    const EosConsts e = eos_consts(p->rho0, p->c0, p->gamma);
    if (D == 3)
        hipLaunchKernelGGL(k_pass_records<3>, grid_n, blk, 0, st, perm, key, start, pos, vel, dens, mass, h, n, pg.g,
                           e, pk, id, sc);
    else
        hipLaunchKernelGGL(k_pass_records<2>, grid_n, blk, 0, st, perm, key, start, pos, vel, dens, mass, h, n, pg.g,
                           e, pk, id, sc);
    NFArgs a{};
    a.pk = pk;
    a.start = start;
    a.g = pg.g;
    a.n = n;
    a.hmax = (float)d->h_max;
    a.nu = (float)p->nu;
    a.c_cfl = p->c_cfl;
    a.k.alpha = (float)p->alpha;
    a.k.beta = (float)p->beta;
    a.k.eps = (float)p->eps;
    a.k.cnorm = (float)(D == 2 ? 80.0 / (7.0 * M_PI) : 16.0 / M_PI);  // src/kernel.h:35-41
    a.k.vsig_mode = p->vsig_mode;
    a.k.grad_mode = 0;
    a.acc = acc;
    a.divv = div_v;
    a.dti = dt_i;
    a.s = sc;
    a.fused_final = 0;
    a.P = P;
    a.out_index = id;
    if (D == 3) p->vsig_mode ? launch_pass_force<3, 1>(a, st) : launch_pass_force<3, 0>(a, st);
    else p->vsig_mode ? launch_pass_force<2, 1>(a, st) : launch_pass_force<2, 0>(a, st);
    if (pairs) {
        hipLaunchKernelGGL(k_fold_counts, dim3(1), dim3(kBlock), 0, st, P, sc);
        hipLaunchKernelGGL(k_pass_pairs, dim3(1), dim3(64), 0, st, (const Scalars*)sc, pairs);
    }
    WCSPH_HIP(hipGetLastError());
}

// ---------------- the variant that allocates nothing ----------------
// For callers that cannot allocate device memory (code under hipGraph
// capture, tight serving loops, harnesses that forbid hipMalloc):
// wcsph_force_pass_noalloc runs the same pass with its workspace in the
// file-scope array `scratch`, sized for up to kNoAllocMaxParticles particles
// and kNoAllocMaxCells cells (about 56 MB). Each GPU the library runs on
// holds its own copy of the array until the process ends. Host code never
// uses the array's name as an address: the device address of each GPU's
// copy is taken once per device with hipGetSymbolAddress (by wcsph_init,
// or by the first call on that device) and kept in noalloc_ptr. Every call
// on a GPU shares that GPU's array and leaves its contents behind; the pass
// clears what it reads before writing, so no call depends on an earlier
// one, but calls that overlap on one GPU race.
constexpr int64_t kNoAllocMaxParticles = (int64_t)1 << 19;
constexpr int64_t kNoAllocMaxCells = (int64_t)1 << 21;
constexpr size_t kScratchBytes = pass_layout(kNoAllocMaxParticles, kNoAllocMaxCells).total;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct NoAllocState {
    std::mutex mu;
    void* ptr[64] = {};
};

NoAllocState& noalloc_state() {
    static NoAllocState st;
    return st;
}

}  // namespace

// The device address of `device`'s copy of scratch (taken on first use,
// with that device current).
void* noalloc_ptr(int device) {
    WCSPH_REQUIRE(device >= 0 && device < 64, WCSPH_EINVAL, "device ordinal out of range");
    NoAllocState& st = noalloc_state();
    std::lock_guard<std::mutex> lk(st.mu);
    if (!st.ptr[device]) {
        int cur = 0;
        WCSPH_HIP(hipGetDevice(&cur));
        WCSPH_HIP(hipSetDevice(device));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        WCSPH_HIP(e);
        WCSPH_REQUIRE(ptr != nullptr, WCSPH_EHIP, "no device address for the force pass's scratch array");
        st.ptr[device] = ptr;
    }
    return st.ptr[device];
}

namespace {

// The library-owned workspace of wcsph_force_pass, one per device, grown
// on demand.
struct OwnedWorkspace {
    std::mutex mu;
    void* ptr[64] = {};
    size_t bytes[64] = {};
};

OwnedWorkspace& owned() {
    static OwnedWorkspace w;
    return w;
}

}  // namespace

}  // namespace wcsph

using namespace wcsph;

extern "C" int wcsph_force_pass(const wcsph_params* p, const wcsph_pass_desc* d, const double* pos,
                                const double* vel, const double* dens, const double* mass, const double* h,
                                double* acc, double* div_v, double* dt_i, int64_t* pairs, void* stream) {
    return guarded([&] {
        check_pass_args(p, d, pos, vel, dens, mass, h, acc, div_v, dt_i);
        const PassGrid pg = pass_grid(p, d);
        WCSPH_REQUIRE(p->device >= 0 && p->device < 64, WCSPH_EINVAL, "device ordinal out of range");
        WCSPH_HIP(hipSetDevice(p->device));
        const size_t need = pass_layout(d->n, pg.ncell).total;
        OwnedWorkspace& w = owned();
        std::lock_guard<std::mutex> lk(w.mu);
        if (w.bytes[p->device] < need) {
            // a pass queued earlier on any stream may still use the old block
            WCSPH_HIP(hipDeviceSynchronize());
            if (w.ptr[p->device]) WCSPH_HIP(hipFree(w.ptr[p->device]));
            w.ptr[p->device] = nullptr;
            w.bytes[p->device] = 0;
            WCSPH_HIP(hipMalloc(&w.ptr[p->device], need));
            w.bytes[p->device] = need;
        }
        run_force_pass(p, d, pg, w.ptr[p->device], pos, vel, dens, mass, h, acc, div_v, dt_i, pairs,
                       (hipStream_t)stream);
    });
}

extern "C" int wcsph_force_pass_noalloc(const wcsph_params* p, const wcsph_pass_desc* d, const double* pos,
                                        const double* vel, const double* dens, const double* mass, const double* h,
                                        double* acc, double* div_v, double* dt_i, int64_t* pairs, void* stream) {
    return guarded([&] {
        check_pass_args(p, d, pos, vel, dens, mass, h, acc, div_v, dt_i);
        const PassGrid pg = pass_grid(p, d);
        WCSPH_REQUIRE(d->n <= kNoAllocMaxParticles, WCSPH_EOVERFLOW,
                      "wcsph_force_pass_noalloc takes at most 524288 particles (its fixed scratch array); "
                      "use wcsph_force_pass");
        WCSPH_REQUIRE(pg.ncell <= kNoAllocMaxCells, WCSPH_EOVERFLOW,
                      "wcsph_force_pass_noalloc takes a grid of at most 2097152 cells of h_max (its fixed scratch "
                      "array); shrink the box or use wcsph_force_pass");
        void* ws = noalloc_ptr(p->device);
        WCSPH_HIP(hipSetDevice(p->device));
        run_force_pass(p, d, pg, ws, pos, vel, dens, mass, h, acc, div_v, dt_i, pairs, (hipStream_t)stream);
    });
}
