"""ctypes view of include/wcsph_abi.h (libwcsph.so).

The library is the product: there is no Python or CPU fallback. Loading
fails with a clear error when libwcsph.so has not been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwcsph.so")

WCSPH_OK, WCSPH_EINVAL, WCSPH_EHIP, WCSPH_ESTATE, WCSPH_ENOMEM, WCSPH_EOVERFLOW, WCSPH_EIO, WCSPH_EDIVERGED = \
    0, 1, 2, 3, 4, 5, 6, 7
PREC_FP32, PREC_MIXED, PREC_FP64 = 0, 1, 2
PRECISIONS = {"fp32": PREC_FP32, "mixed": PREC_MIXED, "fp64": PREC_FP64}


class WcsphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"wcsph error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """wcsph_params."""
    _fields_ = [("dim", C.c_int32), ("vsig_mode", C.c_int32), ("alpha", C.c_double), ("beta", C.c_double),
                ("eps", C.c_double), ("c_cfl", C.c_double), ("rho0", C.c_double), ("c0", C.c_double),
                ("gamma", C.c_double), ("g", C.c_double), ("nu", C.c_double), ("rho_floor", C.c_double),
                ("device", C.c_int32), ("shepard_interval", C.c_int32), ("grad_mode", C.c_int32),
                ("wall_mode", C.c_int32)]


class SimConfig(C.Structure):
    """wcsph_sim_config."""
    _fields_ = [("precision", C.c_int32), ("max_neighbours", C.c_int32), ("box_lo", C.c_double * 3),
                ("box_hi", C.c_double * 3), ("cells_per_support", C.c_int32), ("queue_depth", C.c_int32)]


class SimStats(C.Structure):
    """wcsph_sim_stats."""
    _fields_ = [("time", C.c_double), ("dt", C.c_double), ("step", C.c_int64), ("n", C.c_int64),
                ("pairs_last", C.c_int64), ("pairs_total", C.c_int64), ("escaped", C.c_int64),
                ("max_count", C.c_int32), ("overflow", C.c_int32), ("nonfinite", C.c_int32),
                ("overfull", C.c_int32)]


class DamBreakDesc(C.Structure):
    """wcsph_dam_break_desc."""
    _fields_ = [("dim", C.c_int32), ("nwall", C.c_int32), ("dx", C.c_double), ("col_w", C.c_double),
                ("col_h", C.c_double), ("depth", C.c_double), ("tank_l", C.c_double), ("tank_h", C.c_double),
                ("hfac", C.c_double)]


class Request(C.Structure):
    """wcsph_request."""
    _fields_ = [("params", Params), ("geom", DamBreakDesc), ("precision", C.c_int32), ("steps", C.c_int32),
                ("arrival_s", C.c_double)]


class RequestResult(C.Structure):
    """wcsph_request_result."""
    _fields_ = [("arrival_s", C.c_double), ("start_s", C.c_double), ("end_s", C.c_double), ("n", C.c_int64),
                ("steps", C.c_int64), ("pairs", C.c_int64), ("sim_time", C.c_double), ("front", C.c_double),
                ("column", C.c_double), ("kinetic", C.c_double), ("status", C.c_int32),
                ("precision_used", C.c_int32)]


class PassDesc(C.Structure):
    """wcsph_pass_desc."""
    _fields_ = [("n", C.c_int64), ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3), ("h_max", C.c_double)]


class ServeConfig(C.Structure):
    """wcsph_serve_config."""
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_slots", C.c_int32), ("chunk_steps", C.c_int32)]


def epi_dtype(dim):
    """EPI::Hydro (src/class.h:167-176)."""
    return np.dtype([("pos", "<f8", (dim,)), ("vel", "<f8", (dim,)), ("smth", "<f8"), ("dens", "<f8"),
                     ("pres", "<f8"), ("snds", "<f8"), ("visc", "<f8"), ("tag", "<i8")])


def epj_dtype(dim):
    """EPJ::Hydro (src/class.h:212-223), TYPE enum as int32 plus padding."""
    return np.dtype([("pos", "<f8", (dim,)), ("vel", "<f8", (dim,)), ("dens", "<f8"), ("mass", "<f8"),
                     ("smth", "<f8"), ("pres", "<f8"), ("snds", "<f8"), ("visc", "<f8"), ("tag", "<i8"),
                     ("type", "<i4"), ("_pad", "<i4")])


def res_dtype(dim):
    """RESULT::Hydro (src/class.h:63-67)."""
    return np.dtype([("acc", "<f8", (dim,)), ("dt", "<f8"), ("div_v", "<f8")])


EXPORTS = [
    "wcsph_params_reference", "wcsph_last_error", "wcsph_version", "wcsph_copy_probe", "wcsph_valu_probe",
    "wcsph_init", "wcsph_finalize",
    "wcsph_hydro_dispatch_3d", "wcsph_hydro_dispatch_2d", "wcsph_hydro_retrieve_3d", "wcsph_hydro_retrieve_2d",
    "wcsph_hydro_last_kernel_ms", "wcsph_hydro_last_timing", "wcsph_dam_break_count", "wcsph_dam_break_fill", "wcsph_sim_bytes",
    "wcsph_sim_bytes_state",
    "wcsph_sim_create", "wcsph_sim_destroy", "wcsph_sim_upload", "wcsph_sim_init", "wcsph_sim_step",
    "wcsph_sim_force", "wcsph_sim_sync", "wcsph_sim_get_stats", "wcsph_sim_download", "wcsph_sim_neighbours",
    "wcsph_sim_diagnostics", "wcsph_sim_last_force_ms", "wcsph_sim_set_timing", "wcsph_sim_timing",
    "wcsph_sim_grid", "wcsph_sim_set_graphs", "wcsph_sim_query", "wcsph_serve",
    "wcsph_snapshot_write_ascii", "wcsph_snapshot_read_ascii", "wcsph_sim_write_ascii", "wcsph_sim_save",
    "wcsph_sim_restore", "wcsph_sim_set_clock", "wcsph_sim_params", "wcsph_force_pass",
    "wcsph_force_pass_noalloc",
]

_lib = None

VP = C.c_void_p
I32, I64, F64 = C.c_int32, C.c_int64, C.c_double


def lib():
    """Load libwcsph.so (once) and declare the prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950) first")
    L = C.CDLL(LIB_PATH)
    sig = {
        "wcsph_params_reference": (C.c_int, [I32, C.POINTER(Params)]),
        "wcsph_last_error": (C.c_char_p, []),
        "wcsph_version": (C.c_char_p, []),
        "wcsph_copy_probe": (C.c_int, [I32, I64, I32, C.POINTER(F64)]),
        "wcsph_valu_probe": (C.c_int, [I32, I32, I32, C.POINTER(F64)]),
        "wcsph_init": (C.c_int, [C.POINTER(Params)]),
        "wcsph_finalize": (C.c_int, []),
        "wcsph_hydro_dispatch_3d": (C.c_int, [I32, I32, VP, VP, VP, VP]),
        "wcsph_hydro_dispatch_2d": (C.c_int, [I32, I32, VP, VP, VP, VP]),
        "wcsph_hydro_retrieve_3d": (C.c_int, [I32, I32, VP, VP]),
        "wcsph_hydro_retrieve_2d": (C.c_int, [I32, I32, VP, VP]),
        "wcsph_hydro_last_kernel_ms": (C.c_int, [I32, C.POINTER(C.c_float)]),
        "wcsph_hydro_last_timing": (C.c_int, [I32, VP]),
        "wcsph_dam_break_count": (C.c_int, [C.POINTER(DamBreakDesc), C.POINTER(I64), C.POINTER(I64)]),
        "wcsph_dam_break_fill": (C.c_int, [C.POINTER(DamBreakDesc), C.POINTER(Params), VP, VP, VP, VP, VP, VP]),
        "wcsph_sim_bytes": (I64, [I32, I64, C.POINTER(SimConfig)]),
        "wcsph_sim_bytes_state": (I64, [C.POINTER(Params), C.POINTER(SimConfig), I64, C.c_void_p, C.c_void_p]),
        "wcsph_sim_create": (C.c_int, [C.POINTER(Params), C.POINTER(SimConfig), I64, C.POINTER(VP)]),
        "wcsph_sim_destroy": (C.c_int, [VP]),
        "wcsph_sim_upload": (C.c_int, [VP, VP, VP, VP, VP, VP, VP]),
        "wcsph_sim_init": (C.c_int, [VP]),
        "wcsph_sim_step": (C.c_int, [VP, I32]),
        "wcsph_sim_force": (C.c_int, [VP]),
        "wcsph_sim_sync": (C.c_int, [VP]),
        "wcsph_sim_get_stats": (C.c_int, [VP, C.POINTER(SimStats)]),
        "wcsph_sim_download": (C.c_int, [VP, VP, VP, VP, VP, VP, VP]),
        "wcsph_sim_neighbours": (C.c_int, [VP, VP, VP, I32]),
        "wcsph_sim_diagnostics": (C.c_int, [VP, F64, F64, F64, VP]),
        "wcsph_sim_last_force_ms": (C.c_int, [VP, VP]),
        "wcsph_sim_set_timing": (C.c_int, [VP, I32]),
        "wcsph_sim_grid": (C.c_int, [VP, VP, VP, VP]),
        "wcsph_sim_set_graphs": (C.c_int, [VP, I32]),
        "wcsph_sim_query": (C.c_int, [VP, C.POINTER(I32)]),
        "wcsph_serve": (C.c_int, [C.POINTER(ServeConfig), VP, I32, VP]),
        "wcsph_sim_timing": (C.c_int, [VP, VP, C.POINTER(I64)]),
        "wcsph_snapshot_write_ascii": (C.c_int, [C.c_char_p, I32, I64, F64, VP, VP, VP, VP, VP, VP, VP, VP]),
        "wcsph_snapshot_read_ascii": (C.c_int, [C.c_char_p, I32, C.POINTER(I64), C.POINTER(F64), VP, VP, VP, VP,
                                                VP, VP, VP, VP]),
        "wcsph_sim_write_ascii": (C.c_int, [VP, C.c_char_p]),
        "wcsph_sim_save": (C.c_int, [VP, C.c_char_p]),
        "wcsph_sim_restore": (C.c_int, [C.c_char_p, I32, C.POINTER(VP)]),
        "wcsph_sim_set_clock": (C.c_int, [VP, F64, I64]),
        "wcsph_sim_params": (C.c_int, [VP, C.POINTER(Params), C.POINTER(SimConfig), C.POINTER(I64)]),
        "wcsph_force_pass": (C.c_int, [C.POINTER(Params), C.POINTER(PassDesc)] + [VP] * 10),
        "wcsph_force_pass_noalloc": (C.c_int, [C.POINTER(Params), C.POINTER(PassDesc)] + [VP] * 10),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(code):
    """Raise WcsphError for a nonzero return code."""
    if code != WCSPH_OK:
        raise WcsphError(code, lib().wcsph_last_error().decode())
    return code


def ptr(a):
    return None if a is None else a.ctypes.data
