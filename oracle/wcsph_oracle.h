/*
 * wcsph_oracle.h -- FP64 CPU restatement of the reference hydro force path.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the reference results
 * are compared with (or as the timed CPU baseline). The product library never links or calls it.
 *
 * Each function cites the reference file:line it restates. The reference's
 * own build (FDPS-based) is not buildable in this image, so every function
 * here is pinned by independent checks in tests/test_oracle.py: closed forms
 * (kernel normalisation by quadrature, gradW against a finite difference of
 * W, EoS special values), invariants (momentum antisymmetry, Galilean and
 * rotation invariance, j-list order/superset invariance, dissipation sign),
 * brute force (neighbour sets), and hand-derived two-particle values.
 *
 * Record layouts follow the reference's FDPS boundary types
 * (src/class.h:63-67 RESULT::Hydro, :167-176 EPI::Hydro, :212-223 EPJ::Hydro)
 * in their 64-bit build: every field is 8 bytes; F64vec has D components.
 * Offsets are expressed in doubles:
 *   EPI : pos[D] vel[D] smth dens pres snds visc tag            stride 2D+6
 *   EPJ : pos[D] vel[D] dens mass smth pres snds visc tag type  stride 2D+8
 *   RES : acc[D] dt div_v                                       stride D+2
 */
#ifndef WCSPH_ORACLE_H
#define WCSPH_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dim;          /* 2 or 3 */
    int vsig_mode;    /* 0: v_sig = max(v_sig, mu) as src/force.h:47; 1: max(v_sig, -mu) */
    double alpha;     /* linear AV coefficient (src/force.h:32 fixes it at 0.01) */
    double beta;      /* quadratic AV coefficient (0 in src/force.h:46) */
    double eps;       /* regulariser 0.01 in src/force.h:45,53 */
    double c_cfl;     /* src/param.h:5 */
    int grad_mode;    /* 0: gradW at the mean smoothing length (src/force.h:49);
                         1: the mean of the gradients at h_i and h_j (src/kernel.cu:265) */
} orc_force_params;

/* src/kernel.h:27-55 (CubicSpline<D>) */
double orc_kernel_W(int dim, double r, double h);
double orc_kernel_gradW_factor(int dim, double r, double h); /* gradW = factor * dr */

/* src/EoS.h:75-90 (Murnaghan / Tait) */
double orc_eos_pressure(double dens, double rho0, double c0, double gamma);
double orc_eos_soundspeed(double dens, double rho0, double c0, double gamma);

/* src/force.h:31-60 (CalcHydroForce::operator()), one FDPS i-group.
 * res is accumulated into as FDPS does after RESULT::Hydro::clear(); call
 * orc_result_clear first for a fresh pass. */
void orc_result_clear(int dim, double* res, int ni);
void orc_hydro_force(const orc_force_params* p, const double* epi, int ni,
                     const double* epj, int nj, double* res);

/* Same functor over many FDPS walks (OpenMP over walks when nthreads > 1).
 * epi_off/epj_off: prefix offsets (n_walk+1 entries) in records. */
void orc_hydro_force_walks(const orc_force_params* p, int n_walk,
                           const double* epi, const int64_t* epi_off,
                           const double* epj, const int64_t* epj_off,
                           double* res, int nthreads);

/* Neighbour predicate (FP32, no contraction) and brute-force neighbour sets.
 * Pair (i,j), i != j, is a neighbour iff r2 > 0 && r2 < (hi+hj)^2 where
 * d = (float)(x_i - x_j) per component and r2 = dx*dx + dy*dy (+ dz*dz).
 * This is the support of the mean-h kernel used at src/force.h:49. */
int orc_nbr_pred(int dim, const double* xi, const double* xj, float hi, float hj);
/* pos: SoA [dim][n]; counts[n]; if lists != NULL, lists[i*cap + k] (sorted
 * ascending by j), returns max count. */
int orc_neighbours_brute(int dim, int64_t n, const double* pos, const float* h,
                         int* counts, int* lists, int cap);

/* The device-resident path's neighbour predicate, defined on a cell grid
 * (origin lo, cell edge L, nc cells per axis; cells of size L >= support):
 * c = clamp(floor((x - lo) / L)), origin = lo + c L (FP64, no contraction),
 * xl = (float)(x - origin); for a pair, per axis
 * d = fmaf((float)(c_i - c_j), (float)L, xl_i) - xl_j, then
 * r2 = fmaf(d_z, d_z, fmaf(d_y, d_y, d_x * d_x)) (all FP32, fused where
 * written, nothing else contracted); neighbour iff r2 > 0 && r2 < H * H with
 * H = h_i + h_j (FP32). Same outputs as orc_neighbours_brute. */
int orc_neighbours_brute_cells(int dim, int64_t n, const double* pos, const float* h,
                               const double* lo, double L, const int* nc,
                               int* counts, int* lists, int cap);

/* The same predicate and counts as orc_neighbours_brute_cells, through a
 * cell list over the grid (for large states); every particle must lie
 * inside the grid: returns -1 if one does not, else the summed counts. */
int64_t orc_neighbour_counts_cells(int dim, int64_t n, const double* pos, const float* h,
                                   const double* lo, double L, const int* nc, int* counts, int nthreads);

/* Fake FDPS multiwalk grouping: Morton-ordered i-groups of at most
 * group_max particles, j list = every particle within r_search of the
 * group's bounding box (a superset of each i's support, as FDPS Symmetry
 * lists are). Two-pass: call with out arrays NULL to size. */
int64_t orc_make_walks(int dim, int64_t n, const double* pos, double r_search,
                       int group_max, int64_t* n_walk_out,
                       int64_t* grp_off, int64_t* grp_idx,
                       int64_t* nbr_off, int64_t* nbr_idx);

/* ---------------- whole-step oracle (src/main.cpp:89-133) ---------------- */
typedef struct {
    int dim;
    int vsig_mode;
    double alpha, beta, eps, c_cfl;
    double rho0, c0, gamma;
    double g;          /* gravity along the last axis, src/init/Dam2.h:104-108 */
    double nu;         /* kinematic viscosity, src/EoS.h:38 */
    double rho_floor;  /* src/init/Dam2.h:111 (0 disables) */
    int shepard_interval; /* steps between Shepard filters, src/param.h:12 (0 disables) */
} orc_sim_params;

typedef struct {
    int64_t n;
    double *x, *v, *vh, *a;  /* SoA [dim][n] */
    double *dens, *pres, *snds, *mass, *h, *divv, *dt_i;
    int *type;               /* 0 fluid (HYDRO), 1 wall (FREEZE) */
    double time, dt;
    int64_t step;
} orc_state;

/* Dam-break initial condition on a lattice (independent of the product's). */
int64_t orc_ic_dam_break(int dim, double dx, double col_w, double col_h,
                         double depth, double tank_l, double tank_h, int nwall,
                         double rho0, double c0, double gamma, double g,
                         double hfac, orc_state* s /* NULL to count */);

/* CalcPressure (src/force.h:3-9) over the state */
void orc_sim_eos(const orc_sim_params* p, orc_state* s);
/* ApplyFilter (src/force.h:10-30): Shepard density re-initialisation,
 * out_i = sum_j m_j W(r_ij, h_i) / sum_j (m_j / rho_j) W(r_ij, h_i), both sums
 * over every j (i itself included) within i's support 2 h_i. x is SoA [dim][n]. */
void orc_shepard(int dim, int64_t n, const double* x, const double* h, const double* mass,
                 const double* dens, double* out, int nthreads);
/* CalcPressure, first force, the filter (if enabled) + gravity (src/main.cpp:69-79) */
void orc_sim_init(const orc_sim_params* p, orc_state* s, int nthreads);
/* one step of src/main.cpp:89-133; the filter on steps with
 * step % shepard_interval == 0, after the drift (src/main.cpp:105);
 * returns the true pair count of the step's force pass */
int64_t orc_sim_step(const orc_sim_params* p, orc_state* s, int nthreads);
/* getTimeStepGlobal (src/integral.h:2-12) */
double orc_global_dt(const orc_sim_params* p, const orc_state* s);
/* force pass over all particles with a cell list; returns true pair count */
int64_t orc_sim_force(const orc_sim_params* p, orc_state* s, int nthreads);
/* The scale of each particle's sums in the same pass (no reference
 * counterpart): acc_scale[i] = sum_j |a_ij|_2 and divv_scale[i] =
 * sum_j |div v_ij| / rho_i over the pair terms of src/force.h:51-54, each
 * term taken alone. Per-particle errors of a force pass are stated against
 * these, because the sums cancel (at depth |a_i| is far below them). The
 * state is not changed. */
void orc_sim_force_scales(const orc_sim_params* p, const orc_state* s, int nthreads, double* acc_scale,
                          double* divv_scale);

/* diagnostics (no reference counterpart): out[0]=mass, out[1..3]=momentum,
 * out[4]=kinetic, out[5]=potential, out[6]=internal (Tait) energy */
void orc_conservation(const orc_sim_params* p, const orc_state* s, double* out);
/* max x of fluid with vertical coordinate < zmax */
double orc_dam_front(const orc_state* s, int dim, double zmax);

#ifdef __cplusplus
}
#endif
#endif
