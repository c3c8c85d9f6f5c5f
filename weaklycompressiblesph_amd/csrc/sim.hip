// Device-resident simulation: host orchestration of the step kernels.
//
// One step (src/main.cpp:89-133) is queued as:
//   k_scalars(commit dt)      getTimeStepGlobal result of the previous step
//   k_bin                     initialKick + fullDrift (FP64), then the cell
//                             key and slot of each particle (counting sort)
//   k_scan_reduce/k_scan_apply/k_scatter   cell starts and cell membership
//   k_reorder                 permute the state into cell order (ranked by
//                             previous index within each cell), predict,
//                             EoS, compact FP32 / FP16 force records
//   k_shepard/k_apply_density Shepard filter on steps the schedule names
//   k_nforce / k_nforce_mixed fused neighbour search + hydro force (replaces
//                             the FDPS tree walk and dispatch/retrieve),
//                             ending with gravity + finalKick + density floor
//                             + the block's dt bound; the FP64 path runs
//                             k_neigh (lists) + k_force64 + k_final instead
// The fused force passes add their per-block pair counts into the step
// scalars with one integer atomic per block (order-independent); the
// per-block dt bounds are reduced in index order by the next step's
// k_scalars. Nothing returns to the host between steps; dt lives on the
// device.
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "group_kernels.hpp"
#include "sim.hpp"

using namespace wcsph;

void* wcsph_arena::take(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (size_t k = 0; k < free_list.size(); ++k) {
        auto& f = free_list[k];
        if (f.second < bytes) continue;
        void* p = base + f.first;
        f.first += bytes;
        f.second -= bytes;
        if (f.second == 0) free_list.erase(free_list.begin() + (std::ptrdiff_t)k);
        in_use += bytes;
        return p;
    }
    return nullptr;
}

void wcsph_arena::give(void* p, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    const size_t off = (size_t)((char*)p - base);
    auto it = std::lower_bound(free_list.begin(), free_list.end(), std::make_pair(off, (size_t)0));
    it = free_list.insert(it, {off, bytes});
    in_use -= bytes;
    // merge with the following, then the preceding range
    auto next = it + 1;
    if (next != free_list.end() && it->first + it->second == next->first) {
        it->second += next->second;
        free_list.erase(next);
    }
    if (it != free_list.begin()) {
        auto prev = it - 1;
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            free_list.erase(it);
        }
    }
}

size_t wcsph_arena::largest_free() const {
    size_t m = 0;
    for (const auto& f : free_list) m = std::max(m, f.second);
    return m;
}

void* wcsph_sim::alloc_bytes(size_t bytes) {
    void* ptr = nullptr;
    if (arena) {
        ptr = arena->take(bytes);
        WCSPH_REQUIRE(ptr, WCSPH_ENOMEM, "the serving arena has no free range for a buffer of this request");
    } else {
        WCSPH_HIP(hipMalloc(&ptr, bytes));
    }
    allocs.push_back({ptr, bytes});
    return ptr;
}

wcsph_sim::~wcsph_sim() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& a : allocs) {
        if (arena) arena->give(a.first, a.second);
        else (void)hipFree(a.first);
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : tev) (void)hipEventDestroy(e);
    for (auto& g : graph2)
        if (g) (void)hipGraphExecDestroy(g);
    if (st && own_stream) (void)hipStreamDestroy(st);
    if (hflags && own_flags) (void)hipHostFree((void*)hflags);
}

namespace {

void check_params(const wcsph_params* p) {
    WCSPH_REQUIRE(p != nullptr, WCSPH_EINVAL, "null params");
    WCSPH_REQUIRE(p->dim == 2 || p->dim == 3, WCSPH_EINVAL, "dim must be 2 or 3");
    WCSPH_REQUIRE(p->rho0 > 0 && p->c0 > 0 && p->gamma > 0 && p->c_cfl > 0 && p->eps >= 0, WCSPH_EINVAL,
                  "rho0, c0, gamma, c_cfl must be positive and eps non-negative");
    WCSPH_REQUIRE(p->shepard_interval >= 0, WCSPH_EINVAL, "shepard_interval must be non-negative");
    WCSPH_REQUIRE(p->grad_mode == 0 || p->grad_mode == 1, WCSPH_EINVAL, "grad_mode must be 0 or 1");
    WCSPH_REQUIRE(p->wall_mode == 0 || p->wall_mode == 1, WCSPH_EINVAL, "wall_mode must be 0 or 1");
}

constexpr int kDiag = 10;  // diagnostic sums

// The cell grid of a state: cells slightly wider than the largest support
// 2 h_max / R, over the configured box or the particles' bounding box with
// two cells of margin and half the height again above (room for splashes).
struct GridPlan {
    double lo[3] = {0, 0, 0};
    double L = 0;
    int32_t nc[3] = {1, 1, 1};
    int64_t ncell = 1;
};

GridPlan plan_grid(int D, int R, const wcsph_sim_config& cfg, int64_t n, const double* pos, float hmax) {
    GridPlan G;
    // cell edge slightly above the largest support 2 h_max, so any pair that
    // passes the FP32 predicate lies in adjacent cells
    G.L = 2.0 * (double)hmax * (1.0 + 1.0e-5) / R;
    double hi[3] = {0, 0, 0};
    bool have_box = true;
    for (int d = 0; d < D; ++d) have_box = have_box && cfg.box_hi[d] > cfg.box_lo[d];
    for (int d = 0; d < D; ++d) {
        if (have_box) {
            G.lo[d] = cfg.box_lo[d];
            hi[d] = cfg.box_hi[d];
        } else {
            G.lo[d] = hi[d] = pos[d * n];
            for (int64_t i = 1; i < n; ++i) {
                G.lo[d] = std::min(G.lo[d], pos[d * n + i]);
                hi[d] = std::max(hi[d], pos[d * n + i]);
            }
            const double ext = hi[d] - G.lo[d];
            G.lo[d] -= 2.0 * G.L;
            hi[d] += 2.0 * G.L + (d == D - 1 ? 0.5 * ext : 0.0);  // room above for splashes
        }
    }
    G.ncell = 1;
    for (int d = 0; d < D; ++d) {
        G.nc[d] = (int32_t)std::max<int64_t>(1, (int64_t)std::ceil((hi[d] - G.lo[d]) / G.L));
        G.ncell *= G.nc[d];
    }
    return G;
}

float max_h_fp32(int64_t n, const double* h) {
    double hmax = 0.0;
    for (int64_t i = 0; i < n; ++i) hmax = std::max(hmax, (double)(float)h[i]);
    return (float)hmax;
}

}  // namespace

namespace wcsph {

int64_t sim_footprint(int D, int64_t n, int precision, int K, int64_t ncell) {
    int64_t total = 0;
    auto add = [&total](int64_t bytes) { total += (std::max<int64_t>(bytes, 1) + 255) & ~(int64_t)255; };
    const int64_t nb = ceil_div(n, kBlock);
    for (int b = 0; b < 2; ++b) {  // master state, double-buffered
        for (int k = 0; k < 3; ++k) add(8 * D * n);  // pos vel vh
        for (int k = 0; k < 3; ++k) add(8 * n);      // dens mass h
        add(n);                                      // type
        add(4 * n);                                  // id
    }
    add(8 * D * n);                          // acc
    for (int k = 0; k < 3; ++k) add(8 * n);  // rho_tmp divv dti
    for (int k = 0; k < 3; ++k) add(16 * n);  // FP32 records px pv pq
    add(4 * n);                               // pcell
    if (precision == WCSPH_PREC_FP64) {
        add(8 * n);  // pr64
        add(8 * n);  // c64
    }
    if (precision == WCSPH_PREC_MIXED) {
        const int64_t r = mixed_records(n);
        add(r * (int64_t)sizeof(PosRec16));
        add(r * (int64_t)sizeof(VelRec16));
        add(r * (int64_t)sizeof(StateRec16));
        add(r * (int64_t)sizeof(h2));
    }
    for (int k = 0; k < 3; ++k) add(4 * n);  // key slot perm
    add(sizeof(Scalars));
    add(8 * nb);  // pairs
    add(4 * nb);  // maxcnt
    add(4 * nb);  // over
    add(8 * nb);  // dtmin
    add(4);       // pending (the force pass's counts not yet folded)
    add(8 * nb * kDiag);
    add(8 * kDiag);
    // cell grid: counts and starts with the dead bucket, scan totals
    const int64_t npart = ceil_div(ncell + 1, kScanTile);
    add(4 * (ncell + 1));
    add(4 * (ncell + 2));
    add(4 * npart);
    add(4 * ceil_div(npart, scan_group_size(npart)));
    if (precision == WCSPH_PREC_FP64) {  // neighbour lists of the FP64 pass
        add(4 * (int64_t)K * n);
        add(4 * n);
    }
    return total;
}

int64_t sim_grid_cells(const wcsph_params* p, const wcsph_sim_config* cfg, int64_t n, const double* pos,
                       const double* h, int32_t* nc) {
    wcsph_sim_config c{};
    if (cfg) c = *cfg;
    const int R = c.cells_per_support == 0 ? 2 : c.cells_per_support;
    const GridPlan G = plan_grid(p->dim, R, c, n, pos, max_h_fp32(n, h));
    if (nc)
        for (int d = 0; d < 3; ++d) nc[d] = G.nc[d];
    return G.ncell;
}

}  // namespace wcsph

// Fold a pending force pass's per-block pair counts into the step scalars,
// behind the queued work (before the host reads them).
static void fold_counts_now(wcsph_sim* s) {
    hipLaunchKernelGGL(k_fold_counts, dim3(1), dim3(kBlock), 0, s->st, s->P, s->sc);
}

// Copy the device failure flags to the pinned mirror behind the queued work.
static void queue_flags(wcsph_sim* s) {
    WCSPH_HIP(hipMemcpyAsync((void*)s->hflags, &s->sc->nonfinite, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             s->st));
}

// Report a diverged state: from the mirror as of the last completed copy
// (callers that need it current synchronise first).
static void raise_if_failed(const wcsph_sim* s) {
    const uint32_t nonfinite = s->hflags[0], overfull = s->hflags[1];
    if (nonfinite == 0 && overfull == 0) return;
    std::string msg = "the simulation has diverged:";
    if (nonfinite) msg += " " + std::to_string(nonfinite) + " particles with a non-finite position at the last binning;";
    if (overfull) msg += " a cell holds more than " + std::to_string(kRankCap) + " particles;";
    msg += " steps since then have stopped the clock and done no force work";
    throw Error{WCSPH_EDIVERGED, msg};
}

// v[0] when all n values are equal and positive, else 0.
static double uniform_value(int64_t n, const double* v) {
    for (int64_t i = 1; i < n; ++i)
        if (v[i] != v[0]) return 0.0;
    return v[0] > 0.0 ? v[0] : 0.0;
}

// Largest smoothing length (as FP32, the precision of the predicate) and
// whether all are equal.
static void scan_h(wcsph_sim* s, const double* h) {
    const int64_t n = s->n;
    const double hmax = max_h_fp32(n, h);
    s->uniform_h = true;
    bool same = true;
    for (int64_t i = 0; i < n; ++i) {
        s->uniform_h = s->uniform_h && (float)h[i] == (float)h[0];
        same = same && h[i] == h[0];
    }
    s->h_const = same ? h[0] : 0.0;
    WCSPH_REQUIRE(hmax > 0.0, WCSPH_EINVAL, "smoothing lengths must be positive");
    // with one smoothing length both gradient forms are the same pair body
    WCSPH_REQUIRE(s->p.grad_mode == 0 || same, WCSPH_EINVAL,
                  "grad_mode 1 (the mean of the gradients at h_i and h_j) needs one smoothing length for all "
                  "particles on the device-resident paths (the FDPS drop-in takes it with any h)");
    s->hmax = (float)hmax;
}

static void alloc_grid(wcsph_sim* s, const double* lo, double L, const int32_t* nc) {
    s->grid.invL = 1.0 / L;
    s->grid.L = L;
    s->grid.Lf = (float)L;
    int64_t ncell = 1;
    for (int d = 0; d < 3; ++d) {
        s->grid.lo[d] = d < s->D ? lo[d] : 0.0;
        s->grid.nc[d] = d < s->D ? nc[d] : 1;
        WCSPH_REQUIRE(s->grid.nc[d] >= 1, WCSPH_EINVAL, "cell grid needs at least one cell per axis");
        ncell *= s->grid.nc[d];
    }
    WCSPH_REQUIRE(ncell < (int64_t)1 << 31, WCSPH_EOVERFLOW, "cell grid exceeds 2^31 cells");
    s->ncell = ncell;
    // counts and starts carry one bucket past the last cell: the dead
    // bucket of particles whose position is not finite
    s->npart = ceil_div(ncell + 1, kScanTile);
    s->count = s->alloc<uint32_t>(ncell + 1);
    s->start = s->alloc<uint32_t>(ncell + 2);
    s->ngroups = ceil_div(s->npart, scan_group_size(s->npart));
    s->tile_total = s->alloc<uint32_t>(s->npart);
    s->group_total = s->alloc<uint32_t>(s->ngroups);
    // the scatter clears the counts the scan consumed, ready for the next pass
    WCSPH_HIP(hipMemsetAsync(s->count, 0, sizeof(uint32_t) * (ncell + 1), s->st));
}

static void build_grid(wcsph_sim* s, const double* pos, const double* h) {
    scan_h(s, h);
    const GridPlan G = plan_grid(s->D, s->R, s->cfg, s->n, pos, s->hmax);
    alloc_grid(s, G.lo, G.L, G.nc);
}

// Checks and constants of the precision path that depend on the grid and h.
static void setup_precision(wcsph_sim* s) {
    if (s->cfg.precision != WCSPH_PREC_MIXED) return;
    const int D = s->D;
    WCSPH_REQUIRE(s->uniform_h, WCSPH_EINVAL, "the mixed-precision path needs one smoothing length for all particles");
    WCSPH_REQUIRE(s->grid.nc[0] <= 2048, WCSPH_EINVAL,
                  "the mixed-precision path needs at most 2048 cells along x (FP16 cell index)");
    WCSPH_REQUIRE(mixed_records(s->n) <= ((int64_t)1 << kMixRecBits), WCSPH_EINVAL,
                  "the mixed-precision path holds at most 2^26 particles");
    const double hh = (double)s->hmax;
    s->mc.inv_h = (float)(1.0 / hh);
    s->mc.ch_scale = (float)(-s->p.alpha / s->p.c0);
    s->mc.inv_rho0 = (float)(1.0 / s->p.rho0);
    s->mc.pr_scale = (float)(s->p.rho0 / (s->p.c0 * s->p.c0));
    const double cnorm = (D == 2 ? 80.0 / (7.0 * M_PI) : 16.0 / M_PI) / (D == 2 ? 8.0 : 16.0);
    s->mc.m_scale = (float)(3.0 * cnorm / (s->p.rho0 * std::pow(hh, D)));  // with the spline derivative's 3
    s->mc.half_Ls = (float)(0.5 * s->grid.L / hh);
}

// Argument records of the step kernels for the current state.
static ScalarsArgs scalars_args(const wcsph_sim* s, int mode) { return ScalarsArgs{s->sc, s->P, mode}; }

template <int D>
static BinArgs<D> bin_args(const wcsph_sim* s, int kick) {
    return BinArgs<D>{s->m[s->cur], s->acc, s->divv, kick, s->n, s->grid, s->key, s->slot, s->count, s->sc};
}

static ScanArgs scan_args(const wcsph_sim* s) {
    return ScanArgs{s->count, s->ncell + 1, s->tile_total, s->group_total, scan_group_size(s->npart), s->start};
}

static ScatterArgs scatter_args(const wcsph_sim* s) {
    return ScatterArgs{s->key, s->slot, s->start, s->perm, s->count, s->n};
}

template <int D, class Extra>
static ReorderArgs<D, Extra> reorder_args(const wcsph_sim* s, int kick, Extra extra) {
    const bool mixed = s->cfg.precision == WCSPH_PREC_MIXED;
    ReorderArgs<D, Extra> a{};
    a.perm = s->perm;
    a.key = s->key;
    a.start = s->start;
    a.src = s->m[s->cur];
    a.dst = s->m[1 - s->cur];
    a.acc_old = s->acc;
    a.s = s->sc;
    a.step = kick;
    a.keep_vel = s->cfg.precision == WCSPH_PREC_FP64 ? 1 : 0;
    a.pk = s->pk;
    a.g = s->grid;
    a.pr64 = s->pr64;
    a.c64 = s->c64;
    a.e = eos_consts(s->p.rho0, s->p.c0, s->p.gamma);
    a.v_scale = mixed ? (float)(1.0 / s->p.c0) : 1.0f;
    a.h_const = s->h_const;
    a.m_const = s->m_const;
    a.extra = extra;
    a.n = s->n;
    return a;
}

template <int D, class Extra>
static void launch_reorder(const wcsph_sim* s, const ReorderArgs<D, Extra>& a) {
    hipLaunchKernelGGL((k_reorder<D, Extra>), dim3(s->nb(a.n)), dim3(kBlock), 0, s->st, a.perm, a.key, a.start,
                       a.src, a.dst, a.acc_old, a.s, a.step, a.keep_vel, a.pk, a.g, a.pr64, a.c64, a.e, a.v_scale,
                       a.h_const, a.m_const, a.extra, a.n);
}

template <int D>
static void bin_and_reorder_d(wcsph_sim* s, int kick) {
    const BinArgs<D> b = bin_args<D>(s, kick);
    hipLaunchKernelGGL(k_bin<D>, dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, b.m, b.acc, b.divv, b.kick, b.n, b.g,
                       b.key, b.slot, b.count, b.s);
    const ScanArgs c = scan_args(s);
    hipLaunchKernelGGL(k_scan_reduce, dim3(s->ngroups), dim3(kBlock), 0, s->st, c.count, c.m, c.tile_total,
                       c.group_total, c.gsize);
    hipLaunchKernelGGL(k_scan_apply, dim3(s->npart), dim3(kBlock), 0, s->st, c.count, c.m, c.tile_total,
                       c.group_total, c.gsize, c.start);
    const ScatterArgs t = scatter_args(s);
    hipLaunchKernelGGL(k_scatter, dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, t.key, t.slot, t.start, t.perm, t.count,
                       t.n);
    if (s->cfg.precision == WCSPH_PREC_MIXED)
        launch_reorder<D>(s, reorder_args<D>(s, kick, MixedExtra<D>{s->mx, s->mc}));
    else
        launch_reorder<D>(s, reorder_args<D>(s, kick, NoExtra{}));
}

// Cell list by counting sort and the gather into cell order. step: the
// first half of a step (kick and drift) runs inside the binning pass, and
// the reorder forms the predicted velocities.
static void bin_and_reorder(wcsph_sim* s, bool step) {
    const int kick = step ? 1 : 0;
    if (s->D == 3) bin_and_reorder_d<3>(s, kick);
    else bin_and_reorder_d<2>(s, kick);
    s->cur = 1 - s->cur;
}

// Final-stage arguments for the current state.
static FinalArgs final_args(const wcsph_sim* s, int kick) {
    FinalArgs f;
    f.m = s->m[s->cur];
    f.g = s->p.g;
    f.nu = s->p.nu;
    f.rho_floor = s->p.rho_floor;
    f.kick = kick;
    f.wall_mode = s->p.wall_mode;
    return f;
}

template <class T>
static PairConsts<T> pair_consts(const wcsph_sim* s) {
    PairConsts<T> k;
    k.alpha = (T)s->p.alpha;
    k.beta = (T)s->p.beta;
    k.eps = (T)s->p.eps;
    k.cnorm = (T)(s->D == 2 ? 80.0 / (7.0 * M_PI) : 16.0 / M_PI);
    k.vsig_mode = s->p.vsig_mode;
    k.grad_mode = 0;  // grad_mode 1 is taken only with one smoothing length, where both forms agree
    return k;
}

// The FP32 force records of the current state (cell order) from the master
// state, as k_reorder writes them for the FP32 and FP64 paths: for the
// neighbour lists of the mixed path, which does not keep them.
template <int D>
__global__ void k_fp32_records(Master m, Packed pk, Grid g, EosConsts e, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = m.pos[d * n + p];
    int c[3] = {0, 0, 0};
    bool esc = false;
    cell_of<D>(g, x, c, esc);
    float xl[3] = {0.f, 0.f, 0.f}, vf[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xl[d] = cell_local(x[d], g.lo[d], c[d], g.L);
        vf[d] = (float)m.vel[d * n + p];
    }
    double pr, cs;
    tait(e, m.dens[p], pr, cs);
    pk.px[p] = make_float4(xl[0], xl[1], xl[2], (float)c[0]);
    pk.pv[p] = make_float4(vf[0], vf[1], vf[2], (float)m.mass[p]);
    pk.pq[p] = make_float4((float)m.dens[p], (float)pr, (float)cs, (float)m.h[p]);
    pk.pcell[p] = cell_key<D>(g, c);
}

static void fp32_records(wcsph_sim* s) {
    const EosConsts e = eos_consts(s->p.rho0, s->p.c0, s->p.gamma);
    if (s->D == 3)
        hipLaunchKernelGGL(k_fp32_records<3>, dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, s->m[s->cur], s->pk, s->grid,
                           e, s->n);
    else
        hipLaunchKernelGGL(k_fp32_records<2>, dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, s->m[s->cur], s->pk, s->grid,
                           e, s->n);
}

static void ensure_lists(wcsph_sim* s) {
    if (!s->nbr) {
        s->nbr = s->alloc<uint32_t>((int64_t)s->K * s->n);
        s->cnt = s->alloc<uint32_t>(s->n);
    }
}

template <int D, int R>
static void launch_neigh(wcsph_sim* s, int count_total) {
    ensure_lists(s);
    hipLaunchKernelGGL((k_neigh<D, R>), dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, s->pk, s->start, s->grid, s->n,
                       s->K, s->hmax, s->nbr, s->cnt, s->P, s->sc);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s->st, s->P, s->sc, count_total ? 5 : 1);
}

// count_total = 0 for list rebuilds that must not add to the pair totals
static void neigh_lists(wcsph_sim* s, int count_total = 1) {
    if (s->D == 3) s->R == 1 ? launch_neigh<3, 1>(s, count_total) : launch_neigh<3, 2>(s, count_total);
    else s->R == 1 ? launch_neigh<2, 1>(s, count_total) : launch_neigh<2, 2>(s, count_total);
}

template <int D>
static void launch_force64(wcsph_sim* s) {
    const Master& m = s->m[s->cur];
    ForceArgs64 a;
    a.pos = m.pos;
    a.vel = m.vel;
    a.h = m.h;
    a.rho = m.dens;
    a.m = m.mass;
    a.pr = s->pr64;
    a.c = s->c64;
    a.nbr = s->nbr;
    a.cnt = s->cnt;
    a.acc = s->acc;
    a.divv = s->divv;
    a.dti = s->dti;
    a.n = s->n;
    a.K = s->K;
    a.nu = s->p.nu;
    a.c_cfl = s->p.c_cfl;
    a.k = pair_consts<double>(s);
    a.wtype = s->p.wall_mode ? m.type : nullptr;
    hipLaunchKernelGGL(k_force64<D>, dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, a);
}

template <int D, int R, int Q, bool UH, int VS>
static void launch_nforce(wcsph_sim* s) {
    NFArgs a;
    a.pk = s->pk;
    a.start = s->start;
    a.g = s->grid;
    a.n = s->n;
    a.hmax = s->hmax;
    a.nu = (float)s->p.nu;
    a.c_cfl = s->p.c_cfl;
    a.k = pair_consts<float>(s);
    a.acc = s->acc;
    a.divv = s->divv;
    a.dti = s->dti;
    a.P = s->P;
    a.s = s->sc;
    a.fused_final = s->fuse_final ? 1 : 0;
    a.fin = final_args(s, 1);
    a.out_index = nullptr;
    a.wtype = s->p.wall_mode ? s->m[s->cur].type : nullptr;
    hipLaunchKernelGGL((k_nforce<D, R, Q, UH, VS>), dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, a);
}

template <int D, bool UH, int VS>
static void nforce_dispatch_hv(wcsph_sim* s) {
    if (s->R == 1) s->Q == 32 ? launch_nforce<D, 1, 32, UH, VS>(s) : launch_nforce<D, 1, 24, UH, VS>(s);
    else s->Q == 32 ? launch_nforce<D, 2, 32, UH, VS>(s) : launch_nforce<D, 2, 24, UH, VS>(s);
}

template <int D, bool UH>
static void nforce_dispatch_h(wcsph_sim* s) {
    if (s->p.vsig_mode) nforce_dispatch_hv<D, UH, 1>(s);
    else nforce_dispatch_hv<D, UH, 0>(s);
}

// Arguments of the mixed force pass over the current state.
static NFMixedArgs nforce_mixed_args(const wcsph_sim* s) {
    NFMixedArgs a;
    a.mx = s->mx;
    a.pos = s->m[s->cur].pos;
    a.dens = s->m[s->cur].dens;
    a.gm1 = (float)(s->p.gamma - 1.0);
    a.start = s->start;
    a.g = s->grid;
    a.n = s->n;
    a.nrec = mixed_records(s->n);
    a.h = s->hmax;
    auto splat = [](double v) { return h2{(_Float16)(float)v, (_Float16)(float)v}; };
    a.k_eps = splat(s->p.eps);
    a.beta = (float)(2.0 * s->p.beta);
    a.k_Ls = splat(s->grid.L / (double)s->hmax);
    a.inv_rho0 = 1.0 / s->p.rho0;
    a.acc = s->acc;
    a.divv = s->divv;
    a.dti = s->dti;
    a.P = s->P;
    a.s = s->sc;
    a.fused_final = s->fuse_final ? 1 : 0;
    a.fin = final_args(s, 1);
    // the epilogue's constants, formed once here rather than per particle
    const double c0 = s->p.c0, h = s->hmax;
    a.ks = c0 * c0 / h;
    a.kv = 4.0 * s->p.nu * c0 / (h * h);
    const double hd = s->h_const > 0.0 ? s->h_const : h;  // the particles' FP64 smoothing length
    a.t2 = 0.125 * hd * hd / s->p.nu;
    a.kdivv = (float)(-(c0 / h));
    a.c0f = (float)c0;
    a.cfl_h = (float)(s->p.c_cfl * h);
    a.hf = (float)hd;
    a.g7 = s->p.gamma == 7.0 ? 1 : 0;
    a.wall_mode = s->p.wall_mode;
    return a;
}

template <int D, int R, int VS, int BT>
static void launch_nforce_mixed(wcsph_sim* s) {
    hipLaunchKernelGGL((k_nforce_mixed<D, R, VS, BT>), dim3(s->nb(s->n)), dim3(kBlock), 0, s->st,
                       nforce_mixed_args(s));
}

template <int D, int BT>
static void nforce_mixed_dispatch_b(wcsph_sim* s) {
    if (s->p.vsig_mode) s->R == 1 ? launch_nforce_mixed<D, 1, 1, BT>(s) : launch_nforce_mixed<D, 2, 1, BT>(s);
    else s->R == 1 ? launch_nforce_mixed<D, 1, 0, BT>(s) : launch_nforce_mixed<D, 2, 0, BT>(s);
}

template <int D>
static void nforce_mixed_dispatch(wcsph_sim* s) {
    if (s->p.beta != 0.0) nforce_mixed_dispatch_b<D, 1>(s);
    else nforce_mixed_dispatch_b<D, 0>(s);
}

template <int D>
static void nforce_dispatch(wcsph_sim* s) {
    if (s->cfg.precision == WCSPH_PREC_MIXED) {
        nforce_mixed_dispatch<D>(s);
        return;
    }
    if (s->uniform_h) nforce_dispatch_h<D, true>(s);
    else nforce_dispatch_h<D, false>(s);
}

// e_before / e_mid / e_after (any may be null) bracket the neighbour and
// force kernels (e_mid directly follows e_before on the fused path).
static void neigh_force(wcsph_sim* s, hipEvent_t e_before, hipEvent_t e_mid, hipEvent_t e_after) {
    if (e_before) WCSPH_HIP(hipEventRecord(e_before, s->st));
    if (s->cfg.precision == WCSPH_PREC_FP64) {
        neigh_lists(s);
        if (e_mid) WCSPH_HIP(hipEventRecord(e_mid, s->st));
        if (s->D == 3) launch_force64<3>(s);
        else launch_force64<2>(s);
    } else {
        if (e_mid) WCSPH_HIP(hipEventRecord(e_mid, s->st));
        if (s->D == 3) nforce_dispatch<3>(s);
        else nforce_dispatch<2>(s);
    }
    if (e_after) WCSPH_HIP(hipEventRecord(e_after, s->st));
    s->forced = true;
}

static void final_stage(wcsph_sim* s, int kick) {
    const FinalArgs f = final_args(s, kick);
    if (s->D == 3)
        hipLaunchKernelGGL(k_final<3>, dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, f, s->acc, s->dti, s->sc, s->P,
                           s->n);
    else
        hipLaunchKernelGGL(k_final<2>, dim3(s->nb(s->n)), dim3(kBlock), 0, s->st, f, s->acc, s->dti, s->sc, s->P,
                           s->n);
}

static hipEvent_t* step_events(wcsph_sim* s) {
    const size_t need = (size_t)(s->tsteps + 1) * 5;
    while (s->tev.size() < need) {
        hipEvent_t e;
        WCSPH_HIP(hipEventCreate(&e));
        s->tev.push_back(e);
    }
    return &s->tev[(size_t)s->tsteps * 5];
}

// Shepard filter over the current cell order (records must be up to date).
// The Shepard filter: FP64 from the master state on the FP64 path, FP32
// from the FP32 force records on the FP32 and mixed paths (the mixed path
// writes them for the pass), then the commit of the filtered densities.
template <int D>
static void shepard_pass_d(wcsph_sim* s) {
    const Master& m = s->m[s->cur];
    const int64_t n = s->n;
    const dim3 grid(s->nb(n)), block(kBlock);
    if (s->cfg.precision == WCSPH_PREC_FP64) {
        const double cnorm = D == 2 ? 80.0 / (7.0 * M_PI) : 16.0 / M_PI;  // src/kernel.h:35-41
        if (s->R == 1)
            hipLaunchKernelGGL((k_shepard<D, 1>), grid, block, 0, s->st, m, s->pk, s->start, s->grid, n, cnorm,
                               s->rho_tmp, s->sc);
        else
            hipLaunchKernelGGL((k_shepard<D, 2>), grid, block, 0, s->st, m, s->pk, s->start, s->grid, n, cnorm,
                               s->rho_tmp, s->sc);
    } else {
        if (s->cfg.precision == WCSPH_PREC_MIXED) fp32_records(s);
        if (s->R == 1)
            hipLaunchKernelGGL((k_shepard32<D, 1>), grid, block, 0, s->st, s->pk, s->start, s->grid, n, s->p.rho0,
                               m.dens, s->rho_tmp, s->sc);
        else
            hipLaunchKernelGGL((k_shepard32<D, 2>), grid, block, 0, s->st, s->pk, s->start, s->grid, n, s->p.rho0,
                               m.dens, s->rho_tmp, s->sc);
    }
    const EosConsts e = eos_consts(s->p.rho0, s->p.c0, s->p.gamma);
    if (s->cfg.precision == WCSPH_PREC_MIXED)
        hipLaunchKernelGGL((k_apply_density<D, MixedExtra<D>>), grid, block, 0, s->st, m, s->pk, s->rho_tmp,
                           s->pr64, s->c64, e, MixedExtra<D>{s->mx, s->mc}, n);
    else
        hipLaunchKernelGGL((k_apply_density<D, NoExtra>), grid, block, 0, s->st, m, s->pk, s->rho_tmp, s->pr64,
                           s->c64, e, NoExtra{}, n);
}

static void shepard_pass(wcsph_sim* s) {
    if (s->D == 3) shepard_pass_d<3>(s);
    else shepard_pass_d<2>(s);
}

static bool shepard_due(const wcsph_sim* s, int64_t step) {
    return s->p.shepard_interval > 0 && step % s->p.shepard_interval == 0;
}

static void one_step(wcsph_sim* s, bool filter) {
    hipEvent_t* ev = s->timing ? step_events(s) : nullptr;
    if (ev) WCSPH_HIP(hipEventRecord(ev[0], s->st));
    hipLaunchKernelGGL(k_scalars, dim3(1), dim3(kScalarsThreads), 0, s->st, s->sc, s->P, 1);
    bin_and_reorder(s, true);
    if (filter) shepard_pass(s);  // after the drift, before the pressure (src/main.cpp:105-106)
    // the fused force passes end with the final stage; the FP64 path runs it
    // as a pass of its own
    const bool fused = s->cfg.precision != WCSPH_PREC_FP64;
    s->fuse_final = fused;
    if (ev) neigh_force(s, ev[1], ev[2], ev[3]);
    else neigh_force(s, nullptr, nullptr, nullptr);
    s->fuse_final = false;
    if (!fused) final_stage(s, 1);
    if (ev) {
        WCSPH_HIP(hipEventRecord(ev[4], s->st));
        ++s->tsteps;
    }
}

// FP64 diagnostics: per-block fixed-order tree sums, then one ordered pass.

template <int D>
__global__ __launch_bounds__(kBlock) void k_diag(Master m, int64_t n, double g, double rho0, double B,
                                                 double gamma, double zfront, double xcol, double xwall,
                                                 double* part) {
    __shared__ double sh[kDiag][kBlock];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[kDiag];
    for (int k = 0; k < kDiag; ++k) v[k] = 0.0;
    v[7] = -1.0e300;
    v[8] = -1.0e300;
    v[9] = -1.0e300;
    if (i < n) {
        const double mass = m.mass[i];
        double v2 = 0.0;
        v[0] = mass;
        for (int d = 0; d < D; ++d) {
            const double u = m.vel[d * n + i];
            v[1 + d] = mass * u;
            v2 += u * u;
        }
        v[4] = 0.5 * mass * v2;
        const double z = m.pos[(D - 1) * n + i];
        v[5] = mass * g * z;
        // internal energy of the clamped Tait law (de/drho = P/rho^2, zero below rho0)
        const double r = m.dens[i] / rho0;
        if (r > 1.0) v[6] = mass * B / rho0 * (pow(r, gamma - 1.0) / (gamma - 1.0) + 1.0 / r - gamma / (gamma - 1.0));
        if (m.type[i] == 0) {
            const double x = m.pos[i];
            if (z < zfront) v[7] = x;
            if (x < xcol) v[8] = z;
            if (x > xwall) v[9] = z;
        }
    }
    for (int k = 0; k < kDiag; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int k = 0; k < 7; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
            for (int k = 7; k < kDiag; ++k) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x < kDiag) part[(int64_t)blockIdx.x * kDiag + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_diag_final(const double* part, int64_t nblk, double* out) {
    if (threadIdx.x >= kDiag) return;
    const int k = threadIdx.x;
    double acc = k < 7 ? 0.0 : -1.0e300;
    for (int64_t b = 0; b < nblk; ++b) {
        const double v = part[b * kDiag + k];
        acc = k < 7 ? acc + v : fmax(acc, v);
    }
    out[k] = acc;
}

extern "C" {

int64_t wcsph_sim_bytes(int32_t dim, int64_t n, const wcsph_sim_config* cfg) {
    const int K = (cfg && cfg->max_neighbours > 0) ? cfg->max_neighbours : 128;
    const int prec = cfg ? cfg->precision : WCSPH_PREC_FP32;
    return sim_footprint(dim, n, prec, K, n);  // a grid of about one cell per particle
}

int64_t wcsph_sim_bytes_state(const wcsph_params* p, const wcsph_sim_config* cfg, int64_t n, const double* pos,
                              const double* h) {
    if (!p || !pos || !h || n <= 0 || (p->dim != 2 && p->dim != 3)) return -1;
    const int K = (cfg && cfg->max_neighbours > 0) ? cfg->max_neighbours : 128;
    const int prec = cfg ? cfg->precision : WCSPH_PREC_FP32;
    return sim_footprint(p->dim, n, prec, K, sim_grid_cells(p, cfg, n, pos, h));
}

int wcsph_sim_create(const wcsph_params* p, const wcsph_sim_config* cfg, int64_t n, wcsph_sim** out) {
    return guarded([&] {
        WCSPH_REQUIRE(out != nullptr, WCSPH_EINVAL, "null output handle");
        *out = nullptr;
        *out = sim_create_in(p, cfg, n, nullptr, nullptr, nullptr);
    });
}

}  // extern "C"

wcsph_sim* wcsph::sim_create_in(const wcsph_params* p, const wcsph_sim_config* cfg, int64_t n, wcsph_arena* arena,
                                hipStream_t stream, volatile uint32_t* flags) {
    {
        check_params(p);
        WCSPH_REQUIRE(n > 0 && n < ((int64_t)1 << 31), WCSPH_EINVAL, "n must be in [1, 2^31)");
        auto s = std::make_unique<wcsph_sim>();
        s->arena = arena;
        s->p = *p;
        if (cfg) s->cfg = *cfg;
        WCSPH_REQUIRE(s->cfg.precision == WCSPH_PREC_FP32 || s->cfg.precision == WCSPH_PREC_FP64 ||
                          s->cfg.precision == WCSPH_PREC_MIXED,
                      WCSPH_EINVAL, "unknown precision");
        s->D = p->dim;
        s->n = n;
        s->K = s->cfg.max_neighbours > 0 ? s->cfg.max_neighbours : 128;
        WCSPH_REQUIRE(s->K <= 4096, WCSPH_EINVAL, "max_neighbours above 4096");
        s->R = s->cfg.cells_per_support == 0 ? 2 : s->cfg.cells_per_support;
        WCSPH_REQUIRE(s->R == 1 || s->R == 2, WCSPH_EINVAL, "cells_per_support must be 1 or 2");
        s->Q = s->cfg.queue_depth == 0 ? 24 : s->cfg.queue_depth;
        WCSPH_REQUIRE(s->Q == 24 || s->Q == 32, WCSPH_EINVAL, "queue_depth must be 24 or 32");
        // the fused pass packs a particle index and a row number in 32 bits
        WCSPH_REQUIRE(n < ((int64_t)1 << 27), WCSPH_EINVAL, "too many particles for one simulation (limit 2^27)");
        s->device = p->device;
        WCSPH_HIP(hipSetDevice(s->device));
        if (stream) {
            s->st = stream;
            s->own_stream = false;
        } else {
            WCSPH_HIP(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        }
        for (auto& e : s->ev) WCSPH_HIP(hipEventCreate(&e));
        const int D = s->D;
        for (int b = 0; b < 2; ++b) {
            Master& m = s->m[b];
            m.pos = s->alloc<double>(D * n);
            m.vel = s->alloc<double>(D * n);
            m.vh = s->alloc<double>(D * n);
            m.dens = s->alloc<double>(n);
            m.mass = s->alloc<double>(n);
            m.h = s->alloc<double>(n);
            m.type = s->alloc<uint8_t>(n);
            m.id = s->alloc<uint32_t>(n);
        }
        s->acc = s->alloc<double>(D * n);
        s->rho_tmp = s->alloc<double>(n);
        s->divv = s->alloc<double>(n);
        s->dti = s->alloc<double>(n);
        s->pk.px = s->alloc<float4>(n);
        s->pk.pv = s->alloc<float4>(n);
        s->pk.pq = s->alloc<float4>(n);
        s->pk.pcell = s->alloc<uint32_t>(n);
        if (s->cfg.precision == WCSPH_PREC_FP64) {
            s->pr64 = s->alloc<double>(n);
            s->c64 = s->alloc<double>(n);
        }
        if (s->cfg.precision == WCSPH_PREC_MIXED) {
            s->mx.pos = s->alloc<PosRec16>(mixed_records(n));
            s->mx.vel = s->alloc<VelRec16>(mixed_records(n));
            s->mx.st = s->alloc<StateRec16>(mixed_records(n));
            s->mx.ch = s->alloc<h2>(mixed_records(n));
            // the high half of the last record stays zero when n is odd
            WCSPH_HIP(hipMemsetAsync(s->mx.pos, 0, sizeof(PosRec16) * mixed_records(n), s->st));
            WCSPH_HIP(hipMemsetAsync(s->mx.vel, 0, sizeof(VelRec16) * mixed_records(n), s->st));
            WCSPH_HIP(hipMemsetAsync(s->mx.st, 0, sizeof(StateRec16) * mixed_records(n), s->st));
            WCSPH_HIP(hipMemsetAsync(s->mx.ch, 0, sizeof(h2) * mixed_records(n), s->st));
        }
        s->key = s->alloc<uint32_t>(n);
        s->slot = s->alloc<uint32_t>(n);
        s->perm = s->alloc<uint32_t>(n);
        s->sc = s->alloc<Scalars>(1);
        s->P.nblk = s->nb(n);
        s->P.pairs = s->alloc<unsigned long long>(s->P.nblk);
        s->P.maxcnt = s->alloc<uint32_t>(s->P.nblk);
        s->P.over = s->alloc<uint32_t>(s->P.nblk);
        s->P.dtmin = s->alloc<double>(s->P.nblk);
        s->P.pending = s->alloc<uint32_t>(1);
        WCSPH_HIP(hipMemsetAsync(s->P.pending, 0, sizeof(uint32_t), s->st));
        s->diag_part = s->alloc<double>(s->nb(n) * kDiag);
        s->diag_out = s->alloc<double>(kDiag);
        WCSPH_HIP(hipMemsetAsync(s->acc, 0, sizeof(double) * D * n, s->st));
        WCSPH_HIP(hipMemsetAsync(s->divv, 0, sizeof(double) * n, s->st));
        WCSPH_HIP(hipMemsetAsync(s->sc, 0, sizeof(Scalars), s->st));
        if (flags) {
            s->hflags = flags;
            s->own_flags = false;
        } else {
            WCSPH_HIP(hipHostMalloc((void**)&s->hflags, 2 * sizeof(uint32_t), hipHostMallocDefault));
        }
        s->hflags[0] = s->hflags[1] = 0u;
        WCSPH_HIP(hipStreamSynchronize(s->st));
        return s.release();
    }
}

extern "C" {

int wcsph_sim_destroy(wcsph_sim* s) {
    return guarded([&] { delete s; });
}

int wcsph_sim_upload(wcsph_sim* s, const double* pos, const double* vel, const double* dens, const double* mass,
                     const double* h, const int32_t* type) {
    return guarded([&] {
        WCSPH_REQUIRE(s && pos && vel && dens && mass && h && type, WCSPH_EINVAL, "null argument");
        WCSPH_REQUIRE(!s->uploaded, WCSPH_ESTATE, "state already uploaded");
        WCSPH_HIP(hipSetDevice(s->device));
        const int64_t n = s->n;
        const int D = s->D;
        for (int64_t i = 0; i < n; ++i) {
            WCSPH_REQUIRE(dens[i] > 0 && mass[i] > 0 && h[i] > 0, WCSPH_EINVAL,
                          "density, mass and smoothing length must be positive");
            WCSPH_REQUIRE(std::isfinite(dens[i]) && std::isfinite(mass[i]) && std::isfinite(h[i]), WCSPH_EINVAL,
                          "density, mass and smoothing length must be finite");
            for (int d = 0; d < D; ++d)
                WCSPH_REQUIRE(std::isfinite(pos[d * n + i]) && std::isfinite(vel[d * n + i]), WCSPH_EINVAL,
                              "positions and velocities must be finite");
            WCSPH_REQUIRE(type[i] == 0 || type[i] == 1, WCSPH_EINVAL, "type must be 0 (fluid) or 1 (wall)");
            if (s->p.wall_mode && type[i] == 1)
                for (int d = 0; d < D; ++d)
                    WCSPH_REQUIRE(vel[d * n + i] == 0.0, WCSPH_EINVAL, "wall_mode 1 needs every wall particle at rest");
        }
        build_grid(s, pos, h);
        setup_precision(s);
        std::vector<uint8_t> t8(n);
        std::vector<uint32_t> id(n);
        for (int64_t i = 0; i < n; ++i) {
            t8[i] = (uint8_t)type[i];
            id[i] = (uint32_t)i;
        }
        const Master& m = s->m[s->cur];
        WCSPH_HIP(hipMemcpy(m.pos, pos, sizeof(double) * D * n, hipMemcpyHostToDevice));
        WCSPH_HIP(hipMemcpy(m.vel, vel, sizeof(double) * D * n, hipMemcpyHostToDevice));
        WCSPH_HIP(hipMemcpy(m.vh, vel, sizeof(double) * D * n, hipMemcpyHostToDevice));
        WCSPH_HIP(hipMemcpy(m.dens, dens, sizeof(double) * n, hipMemcpyHostToDevice));
        WCSPH_HIP(hipMemcpy(m.mass, mass, sizeof(double) * n, hipMemcpyHostToDevice));
        WCSPH_HIP(hipMemcpy(m.h, h, sizeof(double) * n, hipMemcpyHostToDevice));
        if (s->h_const > 0.0)  // the reorder does not permute a constant h: both buffers hold it
            WCSPH_HIP(hipMemcpy(s->m[1 - s->cur].h, h, sizeof(double) * n, hipMemcpyHostToDevice));
        s->m_const = uniform_value(n, mass);  // nor a constant mass
        if (s->m_const > 0.0)
            WCSPH_HIP(hipMemcpy(s->m[1 - s->cur].mass, mass, sizeof(double) * n, hipMemcpyHostToDevice));
        WCSPH_HIP(hipMemcpy(m.type, t8.data(), n, hipMemcpyHostToDevice));
        WCSPH_HIP(hipMemcpy(m.id, id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
        s->uploaded = true;
    });
}

int wcsph_sim_init(wcsph_sim* s) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->uploaded, WCSPH_ESTATE, "upload a state first");
        WCSPH_HIP(hipSetDevice(s->device));
        hipLaunchKernelGGL(k_scalars, dim3(1), dim3(kScalarsThreads), 0, s->st, s->sc, s->P, 0);
        bin_and_reorder(s, false);
        neigh_force(s, nullptr, nullptr, nullptr);
        if (s->p.shepard_interval > 0) shepard_pass(s);  // after the first force (src/main.cpp:75)
        final_stage(s, 0);
        queue_flags(s);
        s->hstep = 0;
        s->steppable = true;
        WCSPH_HIP(hipGetLastError());
    });
}

// Two steps captured once per parity of the double-buffered state, then
// replayed with one submission (the kernel arguments are fixed: the two
// steps return the buffers to the parity they started from).
static hipGraphExec_t two_step_graph(wcsph_sim* s) {
    hipGraphExec_t& ge = s->graph2[s->cur];
    if (ge) return ge;
    const int cur0 = s->cur;
    // buffers allocated on first use must exist before capture (a restored
    // FP64 simulation has not yet run the pass that allocates its lists)
    if (s->cfg.precision == WCSPH_PREC_FP64) ensure_lists(s);
    hipGraph_t g = nullptr;
    WCSPH_HIP(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    one_step(s, false);
    one_step(s, false);
    WCSPH_HIP(hipStreamEndCapture(s->st, &g));
    WCSPH_REQUIRE(s->cur == cur0, WCSPH_ESTATE, "two-step capture did not return to its parity");
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    WCSPH_HIP(e);
    return ge;
}

int wcsph_sim_step(wcsph_sim* s, int32_t nsteps) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->uploaded && s->steppable, WCSPH_ESTATE,
                      "call wcsph_sim_init first (and again after wcsph_sim_force)");
        WCSPH_REQUIRE(nsteps >= 0, WCSPH_EINVAL, "nsteps < 0");
        raise_if_failed(s);
        WCSPH_HIP(hipSetDevice(s->device));
        // pairs of steps without a filter replay the captured graph
        for (int k = 0; k < nsteps;) {
            if (s->use_graphs && !s->timing && k + 2 <= nsteps && !shepard_due(s, s->hstep) &&
                !shepard_due(s, s->hstep + 1)) {
                WCSPH_HIP(hipGraphLaunch(two_step_graph(s), s->st));
                s->hstep += 2;
                k += 2;
            } else {
                one_step(s, shepard_due(s, s->hstep));
                ++s->hstep;
                ++k;
            }
        }
        queue_flags(s);
        WCSPH_HIP(hipGetLastError());
    });
}

int wcsph_sim_set_graphs(wcsph_sim* s, int32_t enable) {
    return guarded([&] {
        WCSPH_REQUIRE(s, WCSPH_EINVAL, "null simulation");
        s->use_graphs = enable != 0;
    });
}

int wcsph_sim_query(wcsph_sim* s, int32_t* idle) {
    return guarded([&] {
        WCSPH_REQUIRE(s && idle, WCSPH_EINVAL, "null argument");
        const hipError_t e = hipStreamQuery(s->st);
        if (e == hipErrorNotReady) {
            *idle = 0;
            return;
        }
        WCSPH_HIP(e);
        *idle = 1;
        raise_if_failed(s);
    });
}

int wcsph_sim_force(wcsph_sim* s) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->uploaded, WCSPH_ESTATE, "upload a state first");
        WCSPH_HIP(hipSetDevice(s->device));
        hipLaunchKernelGGL(k_scalars, dim3(1), dim3(kScalarsThreads), 0, s->st, s->sc, s->P, 2);
        s->steppable = false;  // acc is hydro only and the dt bounds are stale
        WCSPH_HIP(hipEventRecord(s->ev[0], s->st));
        bin_and_reorder(s, false);
        neigh_force(s, s->ev[1], s->ev[2], s->ev[3]);
        queue_flags(s);
        WCSPH_HIP(hipGetLastError());
    });
}

int wcsph_sim_last_force_ms(wcsph_sim* s, float* out) {
    return guarded([&] {
        WCSPH_REQUIRE(s && out, WCSPH_EINVAL, "null argument");
        WCSPH_HIP(hipEventSynchronize(s->ev[3]));
        WCSPH_HIP(hipEventElapsedTime(&out[0], s->ev[0], s->ev[1]));
        WCSPH_HIP(hipEventElapsedTime(&out[1], s->ev[1], s->ev[2]));
        WCSPH_HIP(hipEventElapsedTime(&out[2], s->ev[2], s->ev[3]));
    });
}

int wcsph_sim_grid(wcsph_sim* s, double* lo, double* L, int32_t* nc) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->uploaded && lo && L && nc, WCSPH_EINVAL, "null argument or nothing uploaded");
        for (int d = 0; d < 3; ++d) {
            lo[d] = s->grid.lo[d];
            nc[d] = s->grid.nc[d];
        }
        *L = s->grid.L;
    });
}

int wcsph_sim_set_timing(wcsph_sim* s, int32_t enable) {
    return guarded([&] {
        WCSPH_REQUIRE(s, WCSPH_EINVAL, "null simulation");
        s->timing = enable != 0;
        s->tsteps = 0;
    });
}

int wcsph_sim_timing(wcsph_sim* s, double* out, int64_t* nsteps) {
    return guarded([&] {
        WCSPH_REQUIRE(s && out, WCSPH_EINVAL, "null argument");
        WCSPH_HIP(hipStreamSynchronize(s->st));
        double acc[5] = {0, 0, 0, 0, 0};
        for (int64_t k = 0; k < s->tsteps; ++k) {
            hipEvent_t* e = &s->tev[(size_t)k * 5];
            float ms;
            for (int q = 0; q < 4; ++q) {
                WCSPH_HIP(hipEventElapsedTime(&ms, e[q], e[q + 1]));
                acc[q] += ms;
            }
            WCSPH_HIP(hipEventElapsedTime(&ms, e[0], e[4]));
            acc[4] += ms;
        }
        for (int q = 0; q < 5; ++q) out[q] = acc[q];
        if (nsteps) *nsteps = s->tsteps;
        s->tsteps = 0;
    });
}

int wcsph_sim_sync(wcsph_sim* s) {
    return guarded([&] {
        WCSPH_REQUIRE(s, WCSPH_EINVAL, "null simulation");
        WCSPH_HIP(hipStreamSynchronize(s->st));
        raise_if_failed(s);
    });
}

int wcsph_sim_get_stats(wcsph_sim* s, wcsph_sim_stats* st) {
    return guarded([&] {
        WCSPH_REQUIRE(s && st, WCSPH_EINVAL, "null argument");
        WCSPH_HIP(hipSetDevice(s->device));
        fold_counts_now(s);
        Scalars h{};
        WCSPH_HIP(hipMemcpyAsync(&h, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost, s->st));
        WCSPH_HIP(hipStreamSynchronize(s->st));
        st->time = h.time;
        st->dt = h.dt;
        st->step = h.step;
        st->n = s->n;
        st->pairs_last = (int64_t)h.pairs_last;
        st->pairs_total = (int64_t)h.pairs_total;
        st->escaped = h.escaped;
        st->max_count = (int32_t)h.maxcnt;
        st->overflow = (int32_t)h.overflow;
        st->nonfinite = (int32_t)h.nonfinite;
        st->overfull = (int32_t)h.overfull;
    });
}

int wcsph_sim_download(wcsph_sim* s, double* pos, double* vel, double* dens, double* acc, double* div_v,
                       double* dt_i) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->uploaded, WCSPH_ESTATE, "nothing uploaded");
        WCSPH_HIP(hipStreamSynchronize(s->st));
        const int64_t n = s->n;
        const int D = s->D;
        const Master& m = s->m[s->cur];
        std::vector<uint32_t> id(n);
        WCSPH_HIP(hipMemcpy(id.data(), m.id, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        std::vector<double> tmp((size_t)D * n);
        auto fetch = [&](const double* src, double* dst, int comps) {
            if (!dst) return;
            WCSPH_HIP(hipMemcpy(tmp.data(), src, sizeof(double) * comps * n, hipMemcpyDeviceToHost));
            for (int c = 0; c < comps; ++c)
                for (int64_t p = 0; p < n; ++p) dst[c * n + id[p]] = tmp[c * n + p];
        };
        fetch(m.pos, pos, D);
        fetch(m.vel, vel, D);
        fetch(m.dens, dens, 1);
        fetch(s->acc, acc, D);
        fetch(s->divv, div_v, 1);
        fetch(s->dti, dt_i, 1);
    });
}

int wcsph_sim_neighbours(wcsph_sim* s, int32_t* counts, int32_t* lists, int32_t cap) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->forced && counts, WCSPH_ESTATE, "no force pass yet or null counts");
        WCSPH_REQUIRE(!lists || cap > 0, WCSPH_EINVAL, "cap must be positive with lists");
        if (s->cfg.precision != WCSPH_PREC_FP64) {
            // the fused pass keeps no lists: rebuild them for the current cell
            // order (the mixed path keeps no FP32 records either: write them)
            if (s->cfg.precision == WCSPH_PREC_MIXED) fp32_records(s);
            hipLaunchKernelGGL(k_scalars, dim3(1), dim3(kScalarsThreads), 0, s->st, s->sc, s->P, 2);
            neigh_lists(s, 0);
            WCSPH_HIP(hipGetLastError());
        }
        WCSPH_HIP(hipStreamSynchronize(s->st));
        const int64_t n = s->n;
        const Master& m = s->m[s->cur];
        std::vector<uint32_t> id(n), c(n);
        WCSPH_HIP(hipMemcpy(id.data(), m.id, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        WCSPH_HIP(hipMemcpy(c.data(), s->cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < n; ++p) counts[id[p]] = (int32_t)c[p];
        if (!lists) return;
        std::vector<uint32_t> nb((size_t)s->K * n);
        WCSPH_HIP(hipMemcpy(nb.data(), s->nbr, sizeof(uint32_t) * s->K * n, hipMemcpyDeviceToHost));
        std::vector<int32_t> row;
        for (int64_t p = 0; p < n; ++p) {
            const int k = (int)std::min<uint32_t>(c[p], (uint32_t)s->K);
            row.resize(k);
            for (int t = 0; t < k; ++t) row[t] = (int32_t)id[nb[(size_t)t * n + p]];
            std::sort(row.begin(), row.end());
            int32_t* out = lists + (int64_t)id[p] * cap;
            for (int t = 0; t < cap; ++t) out[t] = t < k ? row[t] : -1;
        }
    });
}

int wcsph_sim_diagnostics(wcsph_sim* s, double zfront, double xcol, double xwall, double* out) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->uploaded && out, WCSPH_EINVAL, "null argument or nothing uploaded");
        WCSPH_HIP(hipSetDevice(s->device));
        const double B = s->p.c0 * s->p.c0 * s->p.rho0 / s->p.gamma;
        const int64_t nblk = s->nb(s->n);
        const Master& m = s->m[s->cur];
        if (s->D == 3)
            hipLaunchKernelGGL(k_diag<3>, dim3(nblk), dim3(kBlock), 0, s->st, m, s->n, s->p.g, s->p.rho0, B,
                               s->p.gamma, zfront, xcol, xwall, s->diag_part);
        else
            hipLaunchKernelGGL(k_diag<2>, dim3(nblk), dim3(kBlock), 0, s->st, m, s->n, s->p.g, s->p.rho0, B,
                               s->p.gamma, zfront, xcol, xwall, s->diag_part);
        hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s->st, s->diag_part, nblk, s->diag_out);
        WCSPH_HIP(hipGetLastError());
        WCSPH_HIP(hipMemcpyAsync(out, s->diag_out, sizeof(double) * kDiag, hipMemcpyDeviceToHost, s->st));
        WCSPH_HIP(hipStreamSynchronize(s->st));
    });
}

}  // extern "C"

// ---------------- snapshots and restart ----------------
namespace {

constexpr char kRestartMagic[8] = {'W', 'C', 'S', 'P', 'H', 'R', 'S', 'T'};
constexpr uint32_t kRestartVersion = 2;

// Restart file header; the payload follows it (see restart_arrays).
struct RestartHeader {
    char magic[8];
    uint32_t version, header_bytes;
    int64_t n;
    wcsph_params params;
    wcsph_sim_config cfg;
    double grid_lo[3], grid_L;
    int32_t grid_nc[3], dim;
    double time, dt, dt_next;
    int64_t step, hstep;
    uint64_t pairs_last, pairs_total;
    uint32_t escaped, maxcnt, overflow, pad;
    uint64_t payload_bytes, checksum;
};

// The host-side copy of a simulation's state, in device (cell) order.
struct HostState {
    std::vector<double> pos, vel, dens, mass, h, acc, divv, dti;
    std::vector<uint32_t> id;
    std::vector<uint8_t> type;
    void resize(int D, int64_t n) {
        pos.resize((size_t)D * n);
        vel.resize((size_t)D * n);
        acc.resize((size_t)D * n);
        for (auto* v : {&dens, &mass, &h, &divv, &dti}) v->resize((size_t)n);
        id.resize((size_t)n);
        type.resize((size_t)n);
    }
};

// Payload order: pos, vel, dens, mass, h, acc, div_v, dt_i (FP64), id (u32), type (u8).
template <class F>
void restart_arrays(HostState& h, F&& f) {
    for (auto* v : {&h.pos, &h.vel, &h.dens, &h.mass, &h.h, &h.acc, &h.divv, &h.dti})
        f((void*)v->data(), v->size() * sizeof(double));
    f((void*)h.id.data(), h.id.size() * sizeof(uint32_t));
    f((void*)h.type.data(), h.type.size());
}

int64_t restart_payload_bytes(int D, int64_t n) { return (int64_t)(3 * D + 5) * 8 * n + 5 * n; }

// FNV-1a over 64-bit little-endian words (tail bytes zero-padded).
uint64_t fnv1a_words(uint64_t hs, const void* data, size_t bytes) {
    const uint8_t* p = (const uint8_t*)data;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        hs = (hs ^ w) * 0x100000001b3ull;
    }
    if (i < bytes) {
        uint64_t w = 0;
        std::memcpy(&w, p + i, bytes - i);
        hs = (hs ^ w) * 0x100000001b3ull;
    }
    return hs;
}

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};

void fetch_state(wcsph_sim* s, HostState& hs) {
    const int D = s->D;
    const int64_t n = s->n;
    const Master& m = s->m[s->cur];
    hs.resize(D, n);
    auto get = [&](void* dst, const void* src, size_t bytes) {
        WCSPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    };
    get(hs.pos.data(), m.pos, sizeof(double) * D * n);
    get(hs.vel.data(), m.vel, sizeof(double) * D * n);
    get(hs.dens.data(), m.dens, sizeof(double) * n);
    get(hs.mass.data(), m.mass, sizeof(double) * n);
    get(hs.h.data(), m.h, sizeof(double) * n);
    get(hs.acc.data(), s->acc, sizeof(double) * D * n);
    get(hs.divv.data(), s->divv, sizeof(double) * n);
    get(hs.dti.data(), s->dti, sizeof(double) * n);
    get(hs.id.data(), m.id, sizeof(uint32_t) * n);
    get(hs.type.data(), m.type, n);
}

}  // namespace

extern "C" {

int wcsph_sim_set_clock(wcsph_sim* s, double time, int64_t step) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->uploaded && s->steppable, WCSPH_ESTATE, "call wcsph_sim_init first");
        WCSPH_REQUIRE(step >= 0 && std::isfinite(time), WCSPH_EINVAL, "step must be >= 0 and time finite");
        WCSPH_HIP(hipSetDevice(s->device));
        Scalars h{};
        WCSPH_HIP(hipMemcpyAsync(&h, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost, s->st));
        WCSPH_HIP(hipStreamSynchronize(s->st));
        h.time = time;
        h.step = step;
        WCSPH_HIP(hipMemcpyAsync(s->sc, &h, sizeof(Scalars), hipMemcpyHostToDevice, s->st));
        WCSPH_HIP(hipStreamSynchronize(s->st));
        s->hstep = step;
    });
}

int wcsph_sim_params(wcsph_sim* s, wcsph_params* p, wcsph_sim_config* cfg, int64_t* n) {
    return guarded([&] {
        WCSPH_REQUIRE(s, WCSPH_EINVAL, "null simulation");
        if (p) *p = s->p;
        if (cfg) *cfg = s->cfg;
        if (n) *n = s->n;
    });
}

int wcsph_sim_write_ascii(wcsph_sim* s, const char* path) {
    return guarded([&] {
        WCSPH_REQUIRE(s && s->uploaded && path, WCSPH_EINVAL, "null argument or nothing uploaded");
        WCSPH_HIP(hipSetDevice(s->device));
        WCSPH_HIP(hipStreamSynchronize(s->st));
        HostState hs;
        fetch_state(s, hs);
        Scalars sc{};
        WCSPH_HIP(hipMemcpy(&sc, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
        const int D = s->D;
        const int64_t n = s->n;
        std::vector<double> pos((size_t)D * n), vel((size_t)D * n), dens(n), mass(n), snds(n), pres(n);
        std::vector<int64_t> id(n), tag(n);
        const double rho0 = s->p.rho0, gamma = s->p.gamma;
        const double B = s->p.c0 * s->p.c0 * rho0 / gamma;  // src/EoS.h:78
        parallel_for(n, [&](int64_t q) {
            const int64_t i = hs.id[q];
            for (int d = 0; d < D; ++d) {
                pos[d * n + i] = hs.pos[d * n + q];
                vel[d * n + i] = hs.vel[d * n + q];
            }
            const double rho = hs.dens[q];
            dens[i] = rho;
            mass[i] = hs.mass[q];
            // src/EoS.h:80-86
            pres[i] = std::max(B * (std::pow(rho / rho0, gamma) - 1.0), 0.0);
            snds[i] = std::sqrt(std::max(B * gamma / rho0 * std::pow(rho / rho0, gamma - 1.0), 1.0e-16));
            id[i] = i;
            tag[i] = hs.type[q];
        });
        const int rc = wcsph_snapshot_write_ascii(path, D, n, sc.time, id.data(), tag.data(), mass.data(),
                                                  pos.data(), vel.data(), dens.data(), snds.data(), pres.data());
        if (rc != WCSPH_OK) throw Error{rc, wcsph_last_error()};
    });
}

int wcsph_sim_save(wcsph_sim* s, const char* path) {
    return guarded([&] {
        WCSPH_REQUIRE(s && path, WCSPH_EINVAL, "null argument");
        WCSPH_REQUIRE(s->uploaded && s->steppable, WCSPH_ESTATE,
                      "a restart file holds the state after wcsph_sim_init or a step");
        WCSPH_HIP(hipSetDevice(s->device));
        WCSPH_HIP(hipStreamSynchronize(s->st));
        raise_if_failed(s);  // a diverged state is not worth resuming
        HostState hs;
        fetch_state(s, hs);
        fold_counts_now(s);  // the header carries the pair counts
        WCSPH_HIP(hipStreamSynchronize(s->st));
        Scalars sc{};
        WCSPH_HIP(hipMemcpy(&sc, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
        std::vector<double> dtm((size_t)s->P.nblk);
        WCSPH_HIP(hipMemcpy(dtm.data(), s->P.dtmin, sizeof(double) * s->P.nblk, hipMemcpyDeviceToHost));
        RestartHeader hd;
        std::memset(&hd, 0, sizeof hd);
        std::memcpy(hd.magic, kRestartMagic, 8);
        hd.version = kRestartVersion;
        hd.header_bytes = sizeof(RestartHeader);
        hd.n = s->n;
        hd.params = s->p;
        hd.cfg = s->cfg;
        for (int d = 0; d < 3; ++d) {
            hd.grid_lo[d] = s->grid.lo[d];
            hd.grid_nc[d] = s->grid.nc[d];
        }
        hd.grid_L = s->grid.L;
        hd.dim = s->D;
        hd.time = sc.time;
        hd.dt = sc.dt;
        hd.dt_next = *std::min_element(dtm.begin(), dtm.end());
        hd.step = sc.step;
        hd.hstep = s->hstep;
        hd.pairs_last = sc.pairs_last;
        hd.pairs_total = sc.pairs_total;
        hd.escaped = sc.escaped;
        hd.maxcnt = sc.maxcnt;
        hd.overflow = sc.overflow;
        hd.payload_bytes = (uint64_t)restart_payload_bytes(s->D, s->n);
        uint64_t ck = 0xcbf29ce484222325ull;
        restart_arrays(hs, [&](void* p, size_t b) { ck = fnv1a_words(ck, p, b); });
        hd.checksum = ck;
        // written beside the target and renamed, so a failed save leaves any
        // earlier file at path intact
        const std::string tmp = std::string(path) + ".part";
        {
            std::unique_ptr<FILE, FileCloser> f(std::fopen(tmp.c_str(), "wb"));
            WCSPH_REQUIRE(f != nullptr, WCSPH_EIO, "cannot open " + tmp + ": " + std::strerror(errno));
            bool ok = std::fwrite(&hd, sizeof hd, 1, f.get()) == 1;
            restart_arrays(hs, [&](void* p, size_t b) { ok = ok && std::fwrite(p, 1, b, f.get()) == b; });
            ok = ok && std::fflush(f.get()) == 0;
            if (!ok) {
                f.reset();
                std::remove(tmp.c_str());
                throw Error{WCSPH_EIO, "write failed: " + tmp};
            }
        }
        WCSPH_REQUIRE(std::rename(tmp.c_str(), path) == 0, WCSPH_EIO,
                      std::string("cannot rename ") + tmp + " to " + path + ": " + std::strerror(errno));
    });
}

int wcsph_sim_restore(const char* path, int32_t device, wcsph_sim** out) {
    return guarded([&] {
        WCSPH_REQUIRE(path && out, WCSPH_EINVAL, "null argument");
        *out = nullptr;
        std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "rb"));
        WCSPH_REQUIRE(f != nullptr, WCSPH_EIO, std::string("cannot open ") + path + ": " + std::strerror(errno));
        RestartHeader hd;
        WCSPH_REQUIRE(std::fread(&hd, sizeof hd, 1, f.get()) == 1, WCSPH_EIO, "restart file shorter than its header");
        WCSPH_REQUIRE(std::memcmp(hd.magic, kRestartMagic, 8) == 0, WCSPH_EIO, "not a restart file (bad magic)");
        WCSPH_REQUIRE(hd.version == kRestartVersion && hd.header_bytes == sizeof(RestartHeader), WCSPH_EIO,
                      "restart file of another version or layout");
        WCSPH_REQUIRE((hd.dim == 2 || hd.dim == 3) && hd.params.dim == hd.dim && hd.n > 0 &&
                          hd.n < ((int64_t)1 << 31),
                      WCSPH_EIO, "restart header: bad dim or particle count");
        WCSPH_REQUIRE(hd.payload_bytes == (uint64_t)restart_payload_bytes(hd.dim, hd.n), WCSPH_EIO,
                      "restart header: payload size does not match the particle count");
        WCSPH_REQUIRE(std::fseek(f.get(), 0, SEEK_END) == 0, WCSPH_EIO, "cannot seek");
        const long size = std::ftell(f.get());
        WCSPH_REQUIRE(size == (long)(sizeof hd + hd.payload_bytes), WCSPH_EIO,
                      "restart file truncated or padded: " + std::to_string(size) + " bytes, expected " +
                          std::to_string(sizeof hd + hd.payload_bytes));
        WCSPH_REQUIRE(std::fseek(f.get(), (long)sizeof hd, SEEK_SET) == 0, WCSPH_EIO, "cannot seek");
        const int D = hd.dim;
        const int64_t n = hd.n;
        HostState hs;
        hs.resize(D, n);
        bool ok = true;
        restart_arrays(hs, [&](void* p, size_t b) { ok = ok && std::fread(p, 1, b, f.get()) == b; });
        WCSPH_REQUIRE(ok, WCSPH_EIO, "restart payload read failed");
        uint64_t ck = 0xcbf29ce484222325ull;
        restart_arrays(hs, [&](void* p, size_t b) { ck = fnv1a_words(ck, p, b); });
        WCSPH_REQUIRE(ck == hd.checksum, WCSPH_EIO, "restart file checksum mismatch");
        std::vector<uint8_t> seen((size_t)n, 0);
        for (int64_t q = 0; q < n; ++q) {
            WCSPH_REQUIRE(hs.id[q] < (uint64_t)n && !seen[hs.id[q]], WCSPH_EIO, "restart file: ids are not a permutation");
            seen[hs.id[q]] = 1;
            WCSPH_REQUIRE(hs.type[q] <= 1, WCSPH_EIO, "restart file: bad particle type");
            if (hd.params.wall_mode && hs.type[q] == 1)
                for (int d = 0; d < hd.params.dim; ++d)
                    WCSPH_REQUIRE(hs.vel[d * n + q] == 0.0, WCSPH_EIO, "restart file: wall_mode 1 with a moving wall");
        }
        wcsph_params p = hd.params;
        p.device = device;
        wcsph_sim* raw = nullptr;
        const int rc = wcsph_sim_create(&p, &hd.cfg, n, &raw);
        if (rc != WCSPH_OK) throw Error{rc, wcsph_last_error()};
        std::unique_ptr<wcsph_sim> s(raw);
        WCSPH_HIP(hipSetDevice(s->device));
        scan_h(s.get(), hs.h.data());
        alloc_grid(s.get(), hd.grid_lo, hd.grid_L, hd.grid_nc);
        setup_precision(s.get());
        const Master& m = s->m[s->cur];
        auto put = [&](void* dst, const void* src, size_t bytes) {
            WCSPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        };
        put(m.pos, hs.pos.data(), sizeof(double) * D * n);
        put(m.vel, hs.vel.data(), sizeof(double) * D * n);
        put(m.vh, hs.vel.data(), sizeof(double) * D * n);
        put(m.dens, hs.dens.data(), sizeof(double) * n);
        put(m.mass, hs.mass.data(), sizeof(double) * n);
        put(m.h, hs.h.data(), sizeof(double) * n);
        if (s->h_const > 0.0) put(s->m[1 - s->cur].h, hs.h.data(), sizeof(double) * n);
        s->m_const = uniform_value(n, hs.mass.data());
        if (s->m_const > 0.0) put(s->m[1 - s->cur].mass, hs.mass.data(), sizeof(double) * n);
        put(s->acc, hs.acc.data(), sizeof(double) * D * n);
        put(s->divv, hs.divv.data(), sizeof(double) * n);
        put(s->dti, hs.dti.data(), sizeof(double) * n);
        put(m.id, hs.id.data(), sizeof(uint32_t) * n);
        put(m.type, hs.type.data(), n);
        Scalars sc{};
        sc.dt = hd.dt;
        sc.time = hd.time;
        std::memcpy(&sc.dtmin_bits, &hd.dt, 8);
        sc.step = hd.step;
        sc.pairs_last = hd.pairs_last;
        sc.pairs_total = hd.pairs_total;
        sc.escaped = hd.escaped;
        sc.maxcnt = hd.maxcnt;
        sc.overflow = hd.overflow;
        put(s->sc, &sc, sizeof sc);
        // the next step's k_scalars takes the minimum of the per-block bounds
        std::vector<double> dtm((size_t)s->P.nblk, 1.0e30);
        dtm[0] = hd.dt_next;
        put(s->P.dtmin, dtm.data(), sizeof(double) * s->P.nblk);
        WCSPH_HIP(hipStreamSynchronize(s->st));
        s->hstep = hd.hstep;
        s->uploaded = s->steppable = true;  // neighbour lists need a force pass first
        *out = s.release();
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Support of the grouped steps (group.hip)

template <int D>
MixedStep<D> wcsph::mixed_step_args(wcsph_sim* s) {
    MixedStep<D> a;
    const int cur0 = s->cur;
    a.sc = scalars_args(s, 1);
    a.bin = bin_args<D>(s, 1);
    a.scan = scan_args(s);
    a.scat = scatter_args(s);
    a.ro = reorder_args<D>(s, 1, MixedExtra<D>{s->mx, s->mc});
    // the force pass reads the reordered state and ends with the final stage
    s->cur = 1 - cur0;
    s->fuse_final = true;
    a.nf = nforce_mixed_args(s);
    s->fuse_final = false;
    s->cur = cur0;
    a.nb = (int32_t)s->nb(s->n);
    a.ngroups = (int32_t)s->ngroups;
    a.npart = (int32_t)s->npart;
    return a;
}

template MixedStep<2> wcsph::mixed_step_args<2>(wcsph_sim*);
template MixedStep<3> wcsph::mixed_step_args<3>(wcsph_sim*);

void wcsph::sim_one_step(wcsph_sim* s) {
    WCSPH_HIP(hipSetDevice(s->device));
    one_step(s, shepard_due(s, s->hstep));
    ++s->hstep;
    queue_flags(s);
}

void wcsph::sim_after_group_steps(wcsph_sim* s, int64_t steps) {
    s->hstep += steps;
    queue_flags(s);
}

void wcsph::sim_check_failed(const wcsph_sim* s) { raise_if_failed(s); }
