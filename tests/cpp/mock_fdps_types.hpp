// Minimal stand-ins for the FDPS scalar/vector types and the reference's
// particle records, used only to test include/wcsph_fdps.hpp. The data
// members and their order follow src/class.h:63-67 (RESULT::Hydro),
// :167-176 (EPI::Hydro) and :212-223 (EPJ::Hydro); FDPS's F64vec is a plain
// aggregate of 2 or 3 doubles in its 64-bit build.
#pragma once
#include <cstdint>

namespace PS {
typedef int32_t S32;
typedef int64_t S64;
typedef double F64;
#ifdef PARTICLE_SIMULATOR_TWO_DIMENSION
struct F64vec { F64 x, y; };
#else
struct F64vec { F64 x, y, z; };
#endif
}  // namespace PS

enum TYPE { HYDRO, FREEZE };

namespace STD {
namespace RESULT {
struct Hydro { PS::F64vec acc; PS::F64 dt; PS::F64 div_v; };
}
namespace EPI {
struct Hydro { PS::F64vec pos, vel; PS::F64 smth, dens, pres, snds, visc; PS::S64 tag; };
}
namespace EPJ {
struct Hydro { PS::F64vec pos, vel; PS::F64 dens, mass, smth, pres, snds, visc; PS::S64 tag; TYPE type; };
}
}  // namespace STD
