// wcsph_fdps.hpp -- FDPS multiwalk callbacks with the reference's exact
// signatures, forwarding to the C-ABI of wcsph_abi.h.
//
// Include this header after the reference's particle classes (src/class.h,
// which defines STD::EPI::Hydro, STD::EPJ::Hydro and STD::RESULT::Hydro)
// in place of src/kernel.cu. Nothing else in the caller changes:
//
//   hydr_tree.calcForceAllAndWriteBackMultiWalk(GPU_WC::HydrDispatchKernel,
//       GPU_WC::HydrRetrieveKernel, 1, sph_system, dinfo, N_WALK_LIMIT, true,
//       PS::MAKE_LIST_FOR_REUSE);                         // src/main.cpp:112
//
// The records are passed through unchanged: static_asserts below pin the
// reference's layouts (src/class.h:63-67, :167-176, :212-223) to the C
// records, so FDPS's buffers are read in place with no repacking on the
// caller's side.
//
// Parameters: call GPU_WC::Init(params) once before the first force pass.
// Without it, the first dispatch initialises device 0 with the reference CPU
// functor's constants (wcsph_params_reference), the way the reference sets
// itself up lazily on its first call (src/kernel.cu:290-305) -- but without
// reading a device id from stdin.
#pragma once
#include <cstddef>

#include "wcsph_abi.h"

namespace GPU_WC {

#ifdef PARTICLE_SIMULATOR_TWO_DIMENSION
typedef wcsph_epi_hydro_2d epi_rec;
typedef wcsph_epj_hydro_2d epj_rec;
typedef wcsph_result_hydro_2d res_rec;
static const int kDim = 2;
#else
typedef wcsph_epi_hydro_3d epi_rec;
typedef wcsph_epj_hydro_3d epj_rec;
typedef wcsph_result_hydro_3d res_rec;
static const int kDim = 3;
#endif

#define WCSPH_SAME_FIELD(A, B, fa, fb) \
    static_assert(offsetof(A, fa) == offsetof(B, fb), #A "::" #fa " is not where the C record has it")

static_assert(sizeof(STD::EPI::Hydro) == sizeof(epi_rec), "EPI::Hydro size differs from the C record");
static_assert(sizeof(STD::EPJ::Hydro) == sizeof(epj_rec), "EPJ::Hydro size differs from the C record");
static_assert(sizeof(STD::RESULT::Hydro) == sizeof(res_rec), "RESULT::Hydro size differs from the C record");
WCSPH_SAME_FIELD(STD::EPI::Hydro, epi_rec, pos, pos);
WCSPH_SAME_FIELD(STD::EPI::Hydro, epi_rec, vel, vel);
WCSPH_SAME_FIELD(STD::EPI::Hydro, epi_rec, smth, smth);
WCSPH_SAME_FIELD(STD::EPI::Hydro, epi_rec, dens, dens);
WCSPH_SAME_FIELD(STD::EPI::Hydro, epi_rec, pres, pres);
WCSPH_SAME_FIELD(STD::EPI::Hydro, epi_rec, snds, snds);
WCSPH_SAME_FIELD(STD::EPI::Hydro, epi_rec, visc, visc);
WCSPH_SAME_FIELD(STD::EPJ::Hydro, epj_rec, pos, pos);
WCSPH_SAME_FIELD(STD::EPJ::Hydro, epj_rec, vel, vel);
WCSPH_SAME_FIELD(STD::EPJ::Hydro, epj_rec, dens, dens);
WCSPH_SAME_FIELD(STD::EPJ::Hydro, epj_rec, mass, mass);
WCSPH_SAME_FIELD(STD::EPJ::Hydro, epj_rec, smth, smth);
WCSPH_SAME_FIELD(STD::EPJ::Hydro, epj_rec, pres, pres);
WCSPH_SAME_FIELD(STD::EPJ::Hydro, epj_rec, snds, snds);
WCSPH_SAME_FIELD(STD::EPJ::Hydro, epj_rec, visc, visc);
WCSPH_SAME_FIELD(STD::RESULT::Hydro, res_rec, acc, acc);
WCSPH_SAME_FIELD(STD::RESULT::Hydro, res_rec, dt, dt);
WCSPH_SAME_FIELD(STD::RESULT::Hydro, res_rec, div_v, div_v);
#undef WCSPH_SAME_FIELD

namespace detail {
inline bool& initialised() {
    static bool flag = false;
    return flag;
}
inline int ensure_init() {
    if (initialised()) return WCSPH_OK;
    wcsph_params p;
    int rc = wcsph_params_reference(kDim, &p);
    if (rc == WCSPH_OK) rc = wcsph_init(&p);
    if (rc == WCSPH_OK) initialised() = true;
    return rc;
}
}  // namespace detail

// Set the physics constants and device for the callbacks below.
inline int Init(const wcsph_params& p) {
    const int rc = wcsph_init(&p);
    if (rc == WCSPH_OK) detail::initialised() = true;
    return rc;
}

// src/prototype.h:8 / src/kernel.cu:289
inline PS::S32 HydrDispatchKernel(const PS::S32 tag, const PS::S32 n_walk, const STD::EPI::Hydro** epi,
                                  const PS::S32* n_epi, const STD::EPJ::Hydro** epj, const PS::S32* n_epj) {
    const int rc = detail::ensure_init();
    if (rc != WCSPH_OK) return rc;
#ifdef PARTICLE_SIMULATOR_TWO_DIMENSION
    return wcsph_hydro_dispatch_2d(tag, n_walk, reinterpret_cast<const epi_rec* const*>(epi), n_epi,
                                   reinterpret_cast<const epj_rec* const*>(epj), n_epj);
#else
    return wcsph_hydro_dispatch_3d(tag, n_walk, reinterpret_cast<const epi_rec* const*>(epi), n_epi,
                                   reinterpret_cast<const epj_rec* const*>(epj), n_epj);
#endif
}

// src/prototype.h:9 / src/kernel.cu:383
inline PS::S32 HydrRetrieveKernel(const PS::S32 tag, const PS::S32 n_walk, const PS::S32* ni,
                                  STD::RESULT::Hydro** force) {
#ifdef PARTICLE_SIMULATOR_TWO_DIMENSION
    return wcsph_hydro_retrieve_2d(tag, n_walk, ni, reinterpret_cast<res_rec**>(force));
#else
    return wcsph_hydro_retrieve_3d(tag, n_walk, ni, reinterpret_cast<res_rec**>(force));
#endif
}

}  // namespace GPU_WC
