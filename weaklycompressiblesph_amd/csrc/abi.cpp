// Library-wide C-ABI helpers: error strings, version, reference parameters.
#include <cmath>
#include <string>

#include "common.hpp"

namespace wcsph {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace wcsph

extern "C" {

const char* wcsph_last_error(void) { return wcsph::g_last_error.c_str(); }

const char* wcsph_version(void) { return "wcsph-mi355x 0.1 (gfx950)"; }

int wcsph_params_reference(int32_t dim, wcsph_params* p) {
    if (!p || (dim != 2 && dim != 3)) {
        wcsph::set_error("wcsph_params_reference: dim must be 2 or 3 and p non-null");
        return WCSPH_EINVAL;
    }
    *p = wcsph_params{};
    p->dim = dim;
    p->vsig_mode = 0;   // src/force.h:47
    p->alpha = 0.01;    // src/force.h:32
    p->beta = 0.0;      // src/force.h:46
    p->eps = 0.01;      // src/force.h:45,53
    p->c_cfl = 0.3;     // src/param.h:5
    p->rho0 = 1000.0;   // src/init/Dam2.h:39
    p->gamma = 7.0;     // src/init/Dam2.h:38
    p->c0 = 20.0 * std::sqrt(2.0 * 9.8 * 0.55);  // src/init/Dam2.h:42
    p->g = 9.8;         // src/init/Dam2.h:106
    p->nu = 1.0e-6;     // src/EoS.h:38
    p->rho_floor = 100.0;  // src/init/Dam2.h:111
    p->device = 0;
    p->shepard_interval = 30;  // src/param.h:12
    return WCSPH_OK;
}

}  // extern "C"
