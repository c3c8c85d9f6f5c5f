/*
 * wcsph_oracle.c -- FP64 CPU restatement of the reference hydro force path.
 * TEST INFRASTRUCTURE ONLY (see wcsph_oracle.h). Built with
 * -ffp-contract=off so the FP32 neighbour predicate is evaluated exactly as
 * the product evaluates it.
 */
#include "wcsph_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* src/kernel.h:27-55 CubicSpline<D>                                   */
/* ------------------------------------------------------------------ */
static double cs_norm(int dim) {
    /* normalisedConstant, src/kernel.h:35-41 */
    return dim == 2 ? 80.0 / (7.0 * ORC_PI) : 16.0 / ORC_PI;
}
static double plus0(double a) { return a > 0 ? a : 0; } /* src/kernel.h:12-14 */

double orc_kernel_W(int dim, double r, double h) {
    /* src/kernel.h:42-48: H = 2h, s = r/H, ((1-s)+^3 - 4(1/2-s)+^3) C/H^D */
    const double H = 2.0 * h;
    const double s = r / H;
    const double a = plus0(1.0 - s), b = plus0(0.5 - s);
    return (a * a * a - 4.0 * b * b * b) * cs_norm(dim) / pow(H, dim);
}

double orc_kernel_gradW_factor(int dim, double r, double h) {
    /* src/kernel.h:49-55: (-3(1-s)+^2 + 12(1/2-s)+^2) C/H^D / (r H), times dr */
    const double H = 2.0 * h;
    const double s = r / H;
    const double a = plus0(1.0 - s), b = plus0(0.5 - s);
    return (-3.0 * a * a + 12.0 * b * b) * cs_norm(dim) / pow(H, dim) / (r * H);
}

/* ------------------------------------------------------------------ */
/* src/EoS.h:75-90 Murnaghan (Tait), B = c0^2 rho0 / gamma              */
/* ------------------------------------------------------------------ */
double orc_eos_pressure(double dens, double rho0, double c0, double gamma) {
    const double B = c0 * c0 * rho0 / gamma;
    const double p = B * (pow(dens / rho0, gamma) - 1.0);
    return p > 0.0 ? p : 0.0;
}
double orc_eos_soundspeed(double dens, double rho0, double c0, double gamma) {
    const double B = c0 * c0 * rho0 / gamma;
    const double c2 = B * gamma / rho0 * pow(dens / rho0, gamma - 1.0);
    return sqrt(c2 > 1.0e-16 ? c2 : 1.0e-16);
}

/* ------------------------------------------------------------------ */
/* src/force.h:31-60 CalcHydroForce<kernel_t>::operator()               */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *pos, *vel;
    double smth, dens, pres, snds, visc, mass;
} orc_pv;

/* one (i,j) term of src/force.h:40-54; returns 1 if the pair was not
 * skipped by the dr*dr <= 0 test (src/force.h:42). */
static int pair_term(const orc_force_params* p, int D, const orc_pv* I, const orc_pv* J,
                     double* acc, double* div_v, double* vsig) {
    double dr[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
    double r2 = 0, drdv = 0;
    for (int k = 0; k < D; ++k) {
        dr[k] = I->pos[k] - J->pos[k];
        dv[k] = I->vel[k] - J->vel[k];
        r2 += dr[k] * dr[k];
        drdv += dr[k] * dv[k];
    }
    if (r2 <= 0.0) return 0;
    const double c_ij = 0.5 * (I->snds + J->snds);                       /* :43 */
    const double dens_ij = 0.5 * (I->dens + J->dens);                    /* :44 */
    const double den = r2 + p->eps * I->smth * J->smth;
    const double mu = (drdv < 0.0) ? 0.5 * (I->smth + J->smth) * drdv / den : 0.0; /* :45 */
    const double AV = (-p->alpha * c_ij * mu + p->beta * mu * mu) / dens_ij;      /* :46 */
    const double vs = p->vsig_mode ? -mu : mu;                                      /* :47 */
    if (vs > *vsig) *vsig = vs;
    /* :49 gradW at the mean smoothing length; grad_mode 1: the mean of the
     * gradients at h_i and h_j, as the reference GPU kernel takes it
     * (src/kernel.cu:265; src/force.h:48 has it commented out) */
    const double r = sqrt(r2);
    const double F = p->grad_mode ? 0.5 * (orc_kernel_gradW_factor(D, r, I->smth) + orc_kernel_gradW_factor(D, r, J->smth))
                                  : orc_kernel_gradW_factor(D, r, 0.5 * (I->smth + J->smth));
    const double pt = I->pres / (I->dens * I->dens) + J->pres / (J->dens * J->dens) + AV;
    double drgw = 0, dvgw = 0;
    for (int k = 0; k < D; ++k) {
        const double gw = F * dr[k];
        drgw += dr[k] * gw;
        dvgw += dv[k] * gw;
        acc[k] -= J->mass * pt * gw;                                        /* :51 */
    }
    const double vc = J->mass * 4.0 * 0.5 * (I->visc + J->visc) * drgw / (I->dens + J->dens) / den;
    for (int k = 0; k < D; ++k) acc[k] += vc * dv[k];                       /* :53 */
    *div_v += -J->mass * dvgw;                                              /* :54 */
    return 1;
}

void orc_result_clear(int dim, double* res, int ni) {
    /* RESULT::Hydro::clear, src/class.h:68-72 */
    const int rs = dim + 2;
    for (int i = 0; i < ni; ++i) {
        for (int k = 0; k < dim; ++k) res[i * rs + k] = 0.0;
        res[i * rs + dim] = 1.0e30;
        res[i * rs + dim + 1] = 0.0;
    }
}

static void load_epi(int D, const double* e, orc_pv* o) {
    o->pos = e; o->vel = e + D;
    o->smth = e[2 * D]; o->dens = e[2 * D + 1]; o->pres = e[2 * D + 2];
    o->snds = e[2 * D + 3]; o->visc = e[2 * D + 4]; o->mass = 0.0;
}
static void load_epj(int D, const double* e, orc_pv* o) {
    o->pos = e; o->vel = e + D;
    o->dens = e[2 * D]; o->mass = e[2 * D + 1]; o->smth = e[2 * D + 2];
    o->pres = e[2 * D + 3]; o->snds = e[2 * D + 4]; o->visc = e[2 * D + 5];
}

void orc_hydro_force(const orc_force_params* p, const double* epi, int ni,
                     const double* epj, int nj, double* res) {
    const int D = p->dim;
    const int si = 2 * D + 6, sj = 2 * D + 8, sr = D + 2;
    for (int i = 0; i < ni; ++i) {
        orc_pv I;
        load_epi(D, epi + (int64_t)i * si, &I);
        double* r = res + (int64_t)i * sr;
        double vsig = 0.0;                                                  /* :36 */
        for (int j = 0; j < nj; ++j) {
            orc_pv J;
            load_epj(D, epj + (int64_t)j * sj, &J);
            pair_term(p, D, &I, &J, r, r + D + 1, &vsig);
        }
        r[D] = p->c_cfl * I.smth / (I.snds + vsig);                         /* :56 */
        r[D + 1] /= I.dens;                                                 /* :57 */
    }
}

void orc_hydro_force_walks(const orc_force_params* p, int n_walk,
                           const double* epi, const int64_t* epi_off,
                           const double* epj, const int64_t* epj_off,
                           double* res, int nthreads) {
    const int D = p->dim;
    const int si = 2 * D + 6, sj = 2 * D + 8, sr = D + 2;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int w = 0; w < n_walk; ++w) {
        const int ni = (int)(epi_off[w + 1] - epi_off[w]);
        const int nj = (int)(epj_off[w + 1] - epj_off[w]);
        orc_result_clear(D, res + epi_off[w] * sr, ni);
        orc_hydro_force(p, epi + epi_off[w] * si, ni, epj + epj_off[w] * sj, nj,
                        res + epi_off[w] * sr);
    }
    (void)nthreads;
}

/* ------------------------------------------------------------------ */
/* Neighbour predicate: support of the mean-h kernel (src/force.h:49)  */
/* ------------------------------------------------------------------ */
int orc_nbr_pred(int dim, const double* xi, const double* xj, float hi, float hj) {
    float r2 = 0.0f;
    for (int k = 0; k < dim; ++k) {
        const float d = (float)(xi[k] - xj[k]);
        const float d2 = d * d;
        r2 = r2 + d2;
    }
    const float H = hi + hj;
    const float H2 = H * H;
    return r2 > 0.0f && r2 < H2;
}

static void gather_pos(int dim, int64_t n, const double* pos, int64_t i, double* out) {
    for (int k = 0; k < dim; ++k) out[k] = pos[k * n + i];
}

int orc_neighbours_brute(int dim, int64_t n, const double* pos, const float* h,
                         int* counts, int* lists, int cap) {
    int mx = 0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int64_t i = 0; i < n; ++i) {
        double xi[3], xj[3];
        gather_pos(dim, n, pos, i, xi);
        int c = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            gather_pos(dim, n, pos, j, xj);
            if (orc_nbr_pred(dim, xi, xj, h[i], h[j])) {
                if (lists && c < cap) lists[i * cap + c] = (int)j;
                ++c;
            }
        }
        counts[i] = c;
        if (c > mx) mx = c;
    }
    return mx;
}

static void cell_local(int dim, const double* x, const double* lo, double L, const int* nc,
                       int* c, float* xl) {
    for (int k = 0; k < dim; ++k) {
        const double u = (x[k] - lo[k]) * (1.0 / L);
        int ck = (int)floor(u);
        if (!(u >= 0.0)) ck = 0;
        if (ck > nc[k] - 1) ck = nc[k] - 1;
        const double cl = (double)ck * L;
        const double origin = lo[k] + cl;
        c[k] = ck;
        xl[k] = (float)(x[k] - origin);
    }
}

int orc_neighbours_brute_cells(int dim, int64_t n, const double* pos, const float* h,
                               const double* lo, double L, const int* nc,
                               int* counts, int* lists, int cap) {
    int* ci = (int*)malloc(sizeof(int) * 3 * (size_t)n);
    float* xl = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        double x[3];
        gather_pos(dim, n, pos, i, x);
        cell_local(dim, x, lo, L, nc, ci + 3 * i, xl + 3 * i);
    }
    const float Lf = (float)L;
    int mx = 0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int64_t i = 0; i < n; ++i) {
        int c = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            float d[3] = {0.0f, 0.0f, 0.0f};
            for (int k = 0; k < dim; ++k)
                d[k] = fmaf((float)(ci[3 * i + k] - ci[3 * j + k]), Lf, xl[3 * i + k]) - xl[3 * j + k];
            float r2 = d[0] * d[0];
            r2 = fmaf(d[1], d[1], r2);
            if (dim == 3) r2 = fmaf(d[2], d[2], r2);
            const float H = h[i] + h[j];
            const float H2 = H * H;
            if (r2 > 0.0f && r2 < H2) {
                if (lists && c < cap) lists[i * cap + c] = (int)j;
                ++c;
            }
        }
        counts[i] = c;
        if (c > mx) mx = c;
    }
    free(ci);
    free(xl);
    return mx;
}

/* The cell-local predicate of orc_neighbours_brute_cells, counted through a
 * cell list over the same grid, for states too large for the O(N^2) pass:
 * a pair passing the predicate has r < H <= 2 h_max, so it lies within
 * ceil(2 h_max / L) cells along each axis. Every particle must lie inside
 * the grid (a clamped one keeps its true offset from the edge cell, which
 * this window does not follow): returns -1 if one does not, else the sum
 * of the counts. */
int64_t orc_neighbour_counts_cells(int dim, int64_t n, const double* pos, const float* h,
                                   const double* lo, double L, const int* nc, int* counts, int nthreads) {
    int* ci = (int*)malloc(sizeof(int) * 3 * (size_t)n);
    float* xl = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    float hmax = 0.0f;
    int outside = 0;
    for (int64_t i = 0; i < n; ++i) {
        double x[3];
        gather_pos(dim, n, pos, i, x);
        cell_local(dim, x, lo, L, nc, ci + 3 * i, xl + 3 * i);
        for (int k = 0; k < dim; ++k)
            if (!(x[k] >= lo[k] && x[k] < lo[k] + nc[k] * L)) outside = 1;
        if (h[i] > hmax) hmax = h[i];
    }
    if (outside) {
        free(ci); free(xl);
        return -1;
    }
    int64_t ncell = 1;
    for (int k = 0; k < dim; ++k) ncell *= nc[k];
    int64_t* cstart = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t* cidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    for (int64_t i = 0; i < n; ++i) {
        const int* c = ci + 3 * i;
        const int64_t cc = dim == 3 ? ((int64_t)c[2] * nc[1] + c[1]) * nc[0] + c[0] : (int64_t)c[1] * nc[0] + c[0];
        cstart[cc + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    memcpy(fill, cstart, sizeof(int64_t) * (size_t)ncell);
    for (int64_t i = 0; i < n; ++i) {
        const int* c = ci + 3 * i;
        const int64_t cc = dim == 3 ? ((int64_t)c[2] * nc[1] + c[1]) * nc[0] + c[0] : (int64_t)c[1] * nc[0] + c[0];
        cidx[fill[cc]++] = i;
    }
    const int R = (int)ceil(2.0 * (double)hmax / L);
    const float Lf = (float)L;
    int64_t total = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : total) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < n; ++i) {
        const int* c = ci + 3 * i;
        int cnt = 0;
        const int zr = dim == 3 ? R : 0;
        for (int dz = -zr; dz <= zr; ++dz)
            for (int dy = -R; dy <= R; ++dy)
                for (int dx = -R; dx <= R; ++dx) {
                    const int cx = c[0] + dx, cy = c[1] + dy, cz = (dim == 3 ? c[2] : 0) + dz;
                    if (cx < 0 || cy < 0 || cx >= nc[0] || cy >= nc[1]) continue;
                    if (dim == 3 && (cz < 0 || cz >= nc[2])) continue;
                    const int64_t cc = dim == 3 ? ((int64_t)cz * nc[1] + cy) * nc[0] + cx : (int64_t)cy * nc[0] + cx;
                    for (int64_t t = cstart[cc]; t < cstart[cc + 1]; ++t) {
                        const int64_t j = cidx[t];
                        if (j == i) continue;
                        float d[3] = {0.0f, 0.0f, 0.0f};
                        for (int k = 0; k < dim; ++k)
                            d[k] = fmaf((float)(ci[3 * i + k] - ci[3 * j + k]), Lf, xl[3 * i + k]) - xl[3 * j + k];
                        float r2 = d[0] * d[0];
                        r2 = fmaf(d[1], d[1], r2);
                        if (dim == 3) r2 = fmaf(d[2], d[2], r2);
                        const float H = h[i] + h[j];
                        if (r2 > 0.0f && r2 < H * H) ++cnt;
                    }
                }
        counts[i] = cnt;
        total += cnt;
    }
    free(ci); free(xl); free(cstart); free(cidx); free(fill);
    (void)nthreads;
    return total;
}

/* ------------------------------------------------------------------ */
/* Fake FDPS multiwalk grouping                                        */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t key; int64_t idx; } keyidx;
static int cmp_keyidx(const void* a, const void* b) {
    const keyidx *x = (const keyidx*)a, *y = (const keyidx*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}
static uint64_t spread3(uint64_t v) { /* 21 bits -> every third bit */
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}
static uint64_t spread2(uint64_t v) { /* 32 bits -> every second bit */
    v &= 0xffffffffULL;
    v = (v | v << 16) & 0x0000ffff0000ffffULL;
    v = (v | v << 8) & 0x00ff00ff00ff00ffULL;
    v = (v | v << 4) & 0x0f0f0f0f0f0f0f0fULL;
    v = (v | v << 2) & 0x3333333333333333ULL;
    v = (v | v << 1) & 0x5555555555555555ULL;
    return v;
}

int64_t orc_make_walks(int dim, int64_t n, const double* pos, double r_search,
                       int group_max, int64_t* n_walk_out,
                       int64_t* grp_off, int64_t* grp_idx,
                       int64_t* nbr_off, int64_t* nbr_idx) {
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) {
        lo[k] = hi[k] = pos[k * n];
        for (int64_t i = 1; i < n; ++i) {
            const double v = pos[k * n + i];
            if (v < lo[k]) lo[k] = v;
            if (v > hi[k]) hi[k] = v;
        }
    }
    const double q = r_search * 0.25;
    keyidx* ki = (keyidx*)malloc(sizeof(keyidx) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t c[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) c[k] = (uint64_t)((pos[k * n + i] - lo[k]) / q);
        ki[i].key = dim == 3 ? (spread3(c[0]) | spread3(c[1]) << 1 | spread3(c[2]) << 2)
                             : (spread2(c[0]) | spread2(c[1]) << 1);
        ki[i].idx = i;
    }
    qsort(ki, (size_t)n, sizeof(keyidx), cmp_keyidx);
    const int64_t n_walk = (n + group_max - 1) / group_max;
    *n_walk_out = n_walk;

    /* coarse grid of cell size r_search for the j search */
    int64_t nc[3] = {1, 1, 1};
    for (int k = 0; k < dim; ++k) nc[k] = (int64_t)((hi[k] - lo[k]) / r_search) + 1;
    const int64_t ncell = nc[0] * nc[1] * nc[2];
    int64_t* cstart = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t* cidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* ccell = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) c[k] = (int64_t)((pos[k * n + i] - lo[k]) / r_search);
        ccell[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
        cstart[ccell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    {
        int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
        memcpy(fill, cstart, sizeof(int64_t) * (size_t)ncell);
        for (int64_t i = 0; i < n; ++i) cidx[fill[ccell[i]]++] = i;
        free(fill);
    }

    int64_t total_j = 0;
    if (grp_off) grp_off[0] = 0;
    if (nbr_off) nbr_off[0] = 0;
    for (int64_t w = 0; w < n_walk; ++w) {
        const int64_t b = w * group_max;
        const int64_t e = (b + group_max < n) ? b + group_max : n;
        double blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) { blo[k] = 1e300; bhi[k] = -1e300; }
        for (int64_t t = b; t < e; ++t) {
            const int64_t i = ki[t].idx;
            if (grp_idx) grp_idx[t] = i;
            for (int k = 0; k < dim; ++k) {
                const double v = pos[k * n + i];
                if (v < blo[k]) blo[k] = v;
                if (v > bhi[k]) bhi[k] = v;
            }
        }
        if (grp_off) grp_off[w + 1] = e;
        int64_t c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) {
            c0[k] = (int64_t)((blo[k] - r_search - lo[k]) / r_search) - 1;
            c1[k] = (int64_t)((bhi[k] + r_search - lo[k]) / r_search) + 1;
            if (c0[k] < 0) c0[k] = 0;
            if (c1[k] > nc[k] - 1) c1[k] = nc[k] - 1;
        }
        for (int64_t cz = c0[2]; cz <= c1[2]; ++cz)
            for (int64_t cy = c0[1]; cy <= c1[1]; ++cy)
                for (int64_t cx = c0[0]; cx <= c1[0]; ++cx) {
                    const int64_t c = (cz * nc[1] + cy) * nc[0] + cx;
                    for (int64_t t = cstart[c]; t < cstart[c + 1]; ++t) {
                        const int64_t j = cidx[t];
                        double d2 = 0;
                        for (int k = 0; k < dim; ++k) {
                            const double v = pos[k * n + j];
                            double d = 0;
                            if (v < blo[k]) d = blo[k] - v;
                            else if (v > bhi[k]) d = v - bhi[k];
                            d2 += d * d;
                        }
                        if (d2 < r_search * r_search) {
                            if (nbr_idx) nbr_idx[total_j] = j;
                            ++total_j;
                        }
                    }
                }
        if (nbr_off) nbr_off[w + 1] = total_j;
    }
    free(ki); free(cstart); free(cidx); free(ccell);
    return total_j;
}

/* ------------------------------------------------------------------ */
/* Initial condition: dam break on a lattice                            */
/* ------------------------------------------------------------------ */
static int64_t lround_pos(double v) { return (int64_t)floor(v + 0.5); }

int64_t orc_ic_dam_break(int dim, double dx, double col_w, double col_h,
                         double depth, double tank_l, double tank_h, int nwall,
                         double rho0, double c0, double gamma, double g,
                         double hfac, orc_state* s) {
    const int64_t nw = lround_pos(col_w / dx), nh = lround_pos(col_h / dx);
    const int64_t nl = lround_pos(tank_l / dx), nt = lround_pos(tank_h / dx);
    const int64_t nd = dim == 3 ? lround_pos(depth / dx) : 1;
    const double B = c0 * c0 * rho0 / gamma;
    const double vol = pow(dx, dim);
    int64_t cnt = 0;
    /* pass: fluid then walls; vertical axis is the last one */
    for (int pass = 0; pass < 2; ++pass) {
        const int64_t ylo = dim == 3 ? (pass ? -nwall : 0) : 0;
        const int64_t yhi = dim == 3 ? (pass ? nd + nwall : nd) : 1;
        const int64_t xlo = pass ? -nwall : 0, xhi = pass ? nl + nwall : nw;
        const int64_t zlo = pass ? -nwall : 0, zhi = pass ? nt : nh;
        for (int64_t kz = zlo; kz < zhi; ++kz)
            for (int64_t ky = ylo; ky < yhi; ++ky)
                for (int64_t kx = xlo; kx < xhi; ++kx) {
                    int wall = kx < 0 || kx >= nl || kz < 0 || (dim == 3 && (ky < 0 || ky >= nd));
                    if (pass == 1 && !wall) continue;
                    if (s) {
                        const double px = (kx + 0.5) * dx, py = (ky + 0.5) * dx, pz = (kz + 0.5) * dx;
                        const double vert = pz; /* kz is the vertical index in 2D and 3D */
                        double rho = rho0;
                        if (vert < col_h && px < col_w)
                            rho = rho0 * pow(1.0 + rho0 * g * (col_h - vert) / B, 1.0 / gamma);
                        const int64_t n = s->n;
                        if (dim == 3) {
                            s->x[cnt] = px; s->x[n + cnt] = py; s->x[2 * n + cnt] = pz;
                        } else {
                            s->x[cnt] = px; s->x[n + cnt] = pz;
                        }
                        for (int k = 0; k < dim; ++k) {
                            s->v[k * n + cnt] = 0; s->vh[k * n + cnt] = 0; s->a[k * n + cnt] = 0;
                        }
                        s->dens[cnt] = rho;
                        s->mass[cnt] = rho * vol;
                        s->h[cnt] = hfac * dx;
                        s->type[cnt] = pass;
                        s->divv[cnt] = 0; s->dt_i[cnt] = 1e30;
                        s->pres[cnt] = 0; s->snds[cnt] = 0;
                    }
                    ++cnt;
                }
    }
    return cnt;
}

/* ------------------------------------------------------------------ */
/* Whole-step oracle                                                   */
/* ------------------------------------------------------------------ */
static void state_pv(const orc_sim_params* p, const orc_state* s, int64_t i,
                     double* pos, double* vel, orc_pv* o) {
    const int D = p->dim;
    for (int k = 0; k < D; ++k) { pos[k] = s->x[k * s->n + i]; vel[k] = s->v[k * s->n + i]; }
    o->pos = pos; o->vel = vel;
    o->smth = s->h[i]; o->dens = s->dens[i]; o->pres = s->pres[i];
    o->snds = s->snds[i]; o->visc = p->nu; o->mass = s->mass[i];
}

/* The force pass over all particles with a cell list. With acc_scale (and
 * divv_scale) non-null it leaves the state alone and writes instead, per
 * particle, the sum over j of the magnitudes of the pair terms
 * (sum_j |a_ij|_2 and sum_j |div v_ij|, each term from pair_term alone):
 * the scale against which the cancellation in the sums is judged. */
static int64_t force_pass(const orc_sim_params* p, orc_state* s, int nthreads, double* acc_scale,
                          double* divv_scale) {
    const int D = p->dim;
    const int64_t n = s->n;
    orc_force_params fp = {D, p->vsig_mode, p->alpha, p->beta, p->eps, p->c_cfl, 0};
    double hmax = 0, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i) if (s->h[i] > hmax) hmax = s->h[i];
    for (int k = 0; k < D; ++k) {
        lo[k] = hi[k] = s->x[k * n];
        for (int64_t i = 1; i < n; ++i) {
            const double v = s->x[k * n + i];
            if (v < lo[k]) lo[k] = v;
            if (v > hi[k]) hi[k] = v;
        }
    }
    const double L = 2.0 * hmax * 1.0001;
    int64_t nc[3] = {1, 1, 1};
    for (int k = 0; k < D; ++k) nc[k] = (int64_t)((hi[k] - lo[k]) / L) + 1;
    const int64_t ncell = nc[0] * nc[1] * nc[2];
    int64_t* cstart = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t* cidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* ccell = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) c[k] = (int64_t)((s->x[k * n + i] - lo[k]) / L);
        ccell[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
        cstart[ccell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    memcpy(fill, cstart, sizeof(int64_t) * (size_t)ncell);
    for (int64_t i = 0; i < n; ++i) cidx[fill[ccell[i]]++] = i;
    free(fill);
    int64_t pairs = 0;
    float* hf = (float*)malloc(sizeof(float) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) hf[i] = (float)s->h[i];
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : pairs) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < n; ++i) {
        double pi[3], vi[3], pj[3], vj[3];
        orc_pv I, J;
        state_pv(p, s, i, pi, vi, &I);
        double acc[3] = {0, 0, 0}, divv = 0, vsig = 0;
        int64_t c[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) c[k] = (int64_t)((pi[k] - lo[k]) / L);
        const int64_t zr = D == 3 ? 1 : 0;
        for (int64_t dz = -zr; dz <= zr; ++dz)
            for (int64_t dy = -1; dy <= 1; ++dy)
                for (int64_t dx = -1; dx <= 1; ++dx) {
                    const int64_t cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= nc[0] || cy >= nc[1] || cz >= nc[2]) continue;
                    const int64_t cc = (cz * nc[1] + cy) * nc[0] + cx;
                    for (int64_t t = cstart[cc]; t < cstart[cc + 1]; ++t) {
                        const int64_t j = cidx[t];
                        state_pv(p, s, j, pj, vj, &J);
                        if (acc_scale) {
                            double ta[3] = {0, 0, 0}, td = 0, tv = 0;
                            if (pair_term(&fp, D, &I, &J, ta, &td, &tv)) {
                                double t2 = 0;
                                for (int k = 0; k < D; ++k) t2 += ta[k] * ta[k];
                                acc[0] += sqrt(t2);
                                divv += fabs(td);
                            }
                        } else {
                            pair_term(&fp, D, &I, &J, acc, &divv, &vsig);
                        }
                        if (j != i && orc_nbr_pred(D, pi, pj, hf[i], hf[j])) ++pairs;
                    }
                }
        if (acc_scale) {
            acc_scale[i] = acc[0];
            divv_scale[i] = divv / I.dens;
            continue;
        }
        for (int k = 0; k < D; ++k) s->a[k * n + i] = acc[k];
        s->dt_i[i] = p->c_cfl * I.smth / (I.snds + vsig);
        s->divv[i] = divv / I.dens;
    }
    free(hf); free(cstart); free(cidx); free(ccell);
    return pairs;
}

int64_t orc_sim_force(const orc_sim_params* p, orc_state* s, int nthreads) {
    return force_pass(p, s, nthreads, NULL, NULL);
}

void orc_sim_force_scales(const orc_sim_params* p, const orc_state* s, int nthreads, double* acc_scale,
                          double* divv_scale) {
    force_pass(p, (orc_state*)s, nthreads, acc_scale, divv_scale);
}

void orc_shepard(int dim, int64_t n, const double* x, const double* h, const double* mass,
                 const double* dens, double* out, int nthreads) {
    /* src/force.h:10-30 with the Gather tree of src/main.cpp:65 (j within 2 h_i) */
    double hmax = 0, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i) if (h[i] > hmax) hmax = h[i];
    for (int k = 0; k < dim; ++k) {
        lo[k] = hi[k] = x[k * n];
        for (int64_t i = 1; i < n; ++i) {
            if (x[k * n + i] < lo[k]) lo[k] = x[k * n + i];
            if (x[k * n + i] > hi[k]) hi[k] = x[k * n + i];
        }
    }
    const double L = 2.0 * hmax * 1.0001;
    int64_t nc[3] = {1, 1, 1};
    for (int k = 0; k < dim; ++k) nc[k] = (int64_t)((hi[k] - lo[k]) / L) + 1;
    const int64_t ncell = nc[0] * nc[1] * nc[2];
    int64_t* cstart = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t* cidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t* ccell = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) c[k] = (int64_t)((x[k * n + i] - lo[k]) / L);
        ccell[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
        cstart[ccell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    memcpy(fill, cstart, sizeof(int64_t) * (size_t)ncell);
    for (int64_t i = 0; i < n; ++i) cidx[fill[ccell[i]]++] = i;
    free(fill);
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) c[k] = (int64_t)((x[k * n + i] - lo[k]) / L);
        double one = 0.0, sum = 0.0;
        const int64_t zr = dim == 3 ? 1 : 0;
        for (int64_t dz = -zr; dz <= zr; ++dz)
            for (int64_t dy = -1; dy <= 1; ++dy)
                for (int64_t dx = -1; dx <= 1; ++dx) {
                    const int64_t cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= nc[0] || cy >= nc[1] || cz >= nc[2]) continue;
                    const int64_t cc = (cz * nc[1] + cy) * nc[0] + cx;
                    for (int64_t t = cstart[cc]; t < cstart[cc + 1]; ++t) {
                        const int64_t j = cidx[t];
                        double r2 = 0.0;
                        for (int k = 0; k < dim; ++k) {
                            const double d = x[k * n + i] - x[k * n + j];
                            r2 += d * d;
                        }
                        const double W = orc_kernel_W(dim, sqrt(r2), h[i]);  /* src/force.h:17,23 */
                        one += W * mass[j] / dens[j];                         /* src/force.h:18 */
                        sum += mass[j] * W;
                    }
                }
        out[i] = sum / one;                                                  /* src/force.h:24 */
    }
    free(cstart); free(cidx); free(ccell);
}

static void sim_shepard(const orc_sim_params* p, orc_state* s, int nthreads) {
    double* out = (double*)malloc(sizeof(double) * (size_t)(s->n > 0 ? s->n : 1));
    orc_shepard(p->dim, s->n, s->x, s->h, s->mass, s->dens, out, nthreads);
    memcpy(s->dens, out, sizeof(double) * (size_t)s->n);   /* copyFromForce, src/class.h:116-118 */
    free(out);
}

void orc_sim_eos(const orc_sim_params* p, orc_state* s) {
    /* CalcPressure, src/force.h:3-9 */
    for (int64_t i = 0; i < s->n; ++i) {
        s->pres[i] = orc_eos_pressure(s->dens[i], p->rho0, p->c0, p->gamma);
        s->snds[i] = orc_eos_soundspeed(s->dens[i], p->rho0, p->c0, p->gamma);
    }
}

static void sim_gravity_post(const orc_sim_params* p, orc_state* s, int final_kick) {
    const int D = p->dim;
    const int64_t n = s->n;
    for (int64_t i = 0; i < n; ++i) {
        s->a[(D - 1) * n + i] -= p->g;                 /* src/init/Dam2.h:104-108 */
        if (final_kick && s->type[i] == 0)              /* src/class.h:160-163 */
            for (int k = 0; k < D; ++k)
                s->v[k * n + i] = s->vh[k * n + i] + 0.5 * s->dt * s->a[k * n + i];
        if (s->dens[i] < p->rho_floor) s->dens[i] = p->rho_floor; /* Dam2.h:111 */
    }
}

void orc_sim_init(const orc_sim_params* p, orc_state* s, int nthreads) {
    /* src/main.cpp:69-79: pressure, force, filter, gravity */
    orc_sim_eos(p, s);
    orc_sim_force(p, s, nthreads);
    if (p->shepard_interval > 0) sim_shepard(p, s, nthreads);
    sim_gravity_post(p, s, 0);
}

double orc_global_dt(const orc_sim_params* p, const orc_state* s) {
    /* src/integral.h:2-12 */
    const int D = p->dim;
    double dt = 1.0e30;
    for (int64_t i = 0; i < s->n; ++i) {
        if (s->type[i] != 0) continue;
        double a2 = 0;
        for (int k = 0; k < D; ++k) a2 += s->a[k * s->n + i] * s->a[k * s->n + i];
        double t = 1.0e30;
        const double t1 = 0.25 * sqrt(s->h[i] / sqrt(a2));
        if (t1 < t) t = t1;
        const double t2 = 0.125 * s->h[i] * s->h[i] / p->nu;
        if (t2 < t) t = t2;
        const double ti = s->dt_i[i] < t ? s->dt_i[i] : t;
        if (ti < dt) dt = ti;
    }
    return dt;
}

int64_t orc_sim_step(const orc_sim_params* p, orc_state* s, int nthreads) {
    /* src/main.cpp:89-133 */
    const int D = p->dim;
    const int64_t n = s->n;
    s->dt = orc_global_dt(p, s);
    const double dt = s->dt;
    for (int64_t i = 0; i < n; ++i) {
        const int fluid = s->type[i] == 0;
        if (fluid)                                                   /* initialKick */
            for (int k = 0; k < D; ++k) s->vh[k * n + i] = s->v[k * n + i] + 0.5 * dt * s->a[k * n + i];
        s->dens[i] += -s->dens[i] * dt * s->divv[i];                /* fullDrift */
        if (fluid) {
            for (int k = 0; k < D; ++k) s->x[k * n + i] += dt * s->vh[k * n + i];
            for (int k = 0; k < D; ++k) s->v[k * n + i] += dt * s->a[k * n + i]; /* predict */
        }
    }
    s->time += dt;
    if (p->shepard_interval > 0 && s->step % p->shepard_interval == 0) sim_shepard(p, s, nthreads); /* :105 */
    orc_sim_eos(p, s);
    const int64_t pairs = orc_sim_force(p, s, nthreads);
    sim_gravity_post(p, s, 1);
    s->step += 1;
    return pairs;
}

void orc_conservation(const orc_sim_params* p, const orc_state* s, double* out) {
    const int D = p->dim;
    const int64_t n = s->n;
    for (int k = 0; k < 7; ++k) out[k] = 0;
    const double B = p->c0 * p->c0 * p->rho0 / p->gamma;
    for (int64_t i = 0; i < n; ++i) {
        const double m = s->mass[i];
        out[0] += m;
        double v2 = 0;
        for (int k = 0; k < D; ++k) {
            out[1 + k] += m * s->v[k * n + i];
            v2 += s->v[k * n + i] * s->v[k * n + i];
        }
        out[4] += 0.5 * m * v2;
        out[5] += m * p->g * s->x[(D - 1) * n + i];
        /* internal energy of the clamped Tait law: de/drho = P/rho^2, e(rho0) = 0,
         * e = B/rho0 [ r^(g-1)/(g-1) + 1/r - g/(g-1) ] for r = rho/rho0 >= 1, 0 below */
        const double r = s->dens[i] / p->rho0;
        if (r > 1.0)
            out[6] += m * B / p->rho0 * (pow(r, p->gamma - 1.0) / (p->gamma - 1.0) + 1.0 / r - p->gamma / (p->gamma - 1.0));
    }
}

double orc_dam_front(const orc_state* s, int dim, double zmax) {
    double f = -1e300;
    for (int64_t i = 0; i < s->n; ++i)
        if (s->type[i] == 0 && s->x[(dim - 1) * s->n + i] < zmax && s->x[i] > f) f = s->x[i];
    return f;
}
