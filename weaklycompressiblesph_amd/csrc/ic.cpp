// Dam-break initial condition on a square (2D) or cubic (3D) lattice.
//
// Geometry of the dam-break configurations: a water column col_w x depth x
// col_h standing at x in [0, col_w] against the left wall of a tank
// tank_l x depth x tank_h, with nwall layers of fixed wall particles on the
// side walls and the floor (no lid). Lattice sites sit at (k + 1/2) dx.
// Fluid and the wall under / beside the column start from hydrostatic
// density rho0 (1 + rho0 g (col_h - z) / B)^(1/gamma), B = c0^2 rho0 / gamma;
// mass m = rho dx^D; smoothing length h = hfac dx; velocity zero.
// The reference's own problem (src/init/Dam2.h:9-103) is a different tank
// with an obstacle and uniform density; this generator follows the
// dam-break configurations this library is measured on.
#include <cmath>

#include "common.hpp"

namespace wcsph {
namespace {

struct Lattice {
    int dim, nwall;
    int64_t nw, nh, nd, nl, nt;
};

Lattice make_lattice(const wcsph_dam_break_desc& d) {
    auto cells = [&](double len) { return (int64_t)std::floor(len / d.dx + 0.5); };
    Lattice L;
    L.dim = d.dim;
    L.nwall = d.nwall;
    L.nw = cells(d.col_w);
    L.nh = cells(d.col_h);
    L.nd = d.dim == 3 ? cells(d.depth) : 1;
    L.nl = cells(d.tank_l);
    L.nt = cells(d.tank_h);
    return L;
}

// Visit fluid sites (z-major, then y, then x), then wall sites.
template <class F>
void visit(const Lattice& L, F&& f) {
    const int64_t w = L.nwall;
    const int64_t ylo_f = 0, yhi_f = L.nd;
    for (int64_t z = 0; z < L.nh; ++z)
        for (int64_t y = ylo_f; y < yhi_f; ++y)
            for (int64_t x = 0; x < L.nw; ++x) f(x, y, z, 0);
    const int64_t ylo = L.dim == 3 ? -w : 0, yhi = L.dim == 3 ? L.nd + w : 1;
    for (int64_t z = -w; z < L.nt; ++z)
        for (int64_t y = ylo; y < yhi; ++y)
            for (int64_t x = -w; x < L.nl + w; ++x) {
                const bool inside = x >= 0 && x < L.nl && z >= 0 && (L.dim == 2 || (y >= 0 && y < L.nd));
                if (!inside) f(x, y, z, 1);
            }
}

void check_desc(const wcsph_dam_break_desc* d) {
    WCSPH_REQUIRE(d != nullptr, WCSPH_EINVAL, "null dam-break description");
    WCSPH_REQUIRE(d->dim == 2 || d->dim == 3, WCSPH_EINVAL, "dim must be 2 or 3");
    WCSPH_REQUIRE(d->dx > 0 && d->col_w > 0 && d->col_h > 0 && d->tank_l >= d->col_w &&
                      d->tank_h > 0 && d->nwall >= 0 && d->hfac > 0 && (d->dim == 2 || d->depth > 0),
                  WCSPH_EINVAL, "invalid dam-break geometry");
}

}  // namespace
}  // namespace wcsph

using namespace wcsph;

extern "C" {

int wcsph_dam_break_count(const wcsph_dam_break_desc* d, int64_t* n_fluid, int64_t* n_wall) {
    return guarded([&] {
        check_desc(d);
        const Lattice L = make_lattice(*d);
        int64_t nf = 0, nw = 0;
        visit(L, [&](int64_t, int64_t, int64_t, int wall) { (wall ? nw : nf)++; });
        if (n_fluid) *n_fluid = nf;
        if (n_wall) *n_wall = nw;
    });
}

int wcsph_dam_break_fill(const wcsph_dam_break_desc* d, const wcsph_params* p, double* pos, double* vel,
                         double* dens, double* mass, double* h, int32_t* type) {
    return guarded([&] {
        check_desc(d);
        WCSPH_REQUIRE(p && pos && vel && dens && mass && h && type, WCSPH_EINVAL, "null argument");
        WCSPH_REQUIRE(p->dim == d->dim, WCSPH_EINVAL, "params and description differ in dim");
        const Lattice L = make_lattice(*d);
        int64_t n = 0;
        visit(L, [&](int64_t, int64_t, int64_t, int) { ++n; });
        const int D = d->dim;
        const double B = p->c0 * p->c0 * p->rho0 / p->gamma;
        const double vol = std::pow(d->dx, D);
        int64_t i = 0;
        visit(L, [&](int64_t x, int64_t y, int64_t z, int wall) {
            const double px = (x + 0.5) * d->dx, py = (y + 0.5) * d->dx, pz = (z + 0.5) * d->dx;
            double rho = p->rho0;
            if (pz < d->col_h && px < d->col_w)
                rho = p->rho0 * std::pow(1.0 + p->rho0 * p->g * (d->col_h - pz) / B, 1.0 / p->gamma);
            pos[i] = px;
            if (D == 3) {
                pos[n + i] = py;
                pos[2 * n + i] = pz;
            } else {
                pos[n + i] = pz;
            }
            for (int k = 0; k < D; ++k) vel[k * n + i] = 0.0;
            dens[i] = rho;
            mass[i] = rho * vol;
            h[i] = d->hfac * d->dx;
            type[i] = wall;
            ++i;
        });
    });
}

}  // extern "C"
