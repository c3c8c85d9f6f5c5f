"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the reference results are
compared with or as the timed CPU baseline. The product package never imports it.

The oracle restates the reference's hydro force path:
  src/force.h:31-60  CalcHydroForce  -> hydro_force / hydro_force_walks
  src/kernel.h:27-55 CubicSpline<D>  -> kernel_W / kernel_gradW_factor
  src/EoS.h:75-90    Murnaghan       -> eos_pressure / eos_soundspeed
  src/integral.h:2-12                -> State.global_dt
  src/class.h:146-163, src/main.cpp:89-133 -> State.step
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C")


class ForceParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("vsig_mode", C.c_int), ("alpha", C.c_double),
                ("beta", C.c_double), ("eps", C.c_double), ("c_cfl", C.c_double), ("grad_mode", C.c_int)]


class SimParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("vsig_mode", C.c_int), ("alpha", C.c_double),
                ("beta", C.c_double), ("eps", C.c_double), ("c_cfl", C.c_double),
                ("rho0", C.c_double), ("c0", C.c_double), ("gamma", C.c_double),
                ("g", C.c_double), ("nu", C.c_double), ("rho_floor", C.c_double),
                ("shepard_interval", C.c_int)]


class _CState(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in
                ("x", "v", "vh", "a", "dens", "pres", "snds", "mass", "h", "divv", "dt_i", "type")] + \
               [("time", C.c_double), ("dt", C.c_double), ("step", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_variant(outdir, name, flags):
    """Compile the oracle with other optimisation flags (the CPU baseline's
    -O3 -march=native builds, made on the machine that runs them) into
    outdir/liboracle_<name>.so; returns its path."""
    out = os.path.join(outdir, f"liboracle_{name}.so")
    subprocess.run(["gcc", *flags, "-fPIC", "-fopenmp", "-std=c99", "-shared", "-o", out,
                    os.path.join(_HERE, "wcsph_oracle.c"), "-lm"], check=True)
    return out


def use_library(path=None):
    """Route the oracle's functions to the build at `path` (None: the
    default liboracle.so). Returns the path that was in use before."""
    global _lib, _lib_path
    prev = _lib_path
    _lib = _load(path or _LIB)
    _lib_path = path
    return prev


_lib_path = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = _load(_LIB)
    return _lib


def _load(path):
    L = C.CDLL(path)
    L.orc_kernel_W.restype = C.c_double
    L.orc_kernel_W.argtypes = [C.c_int, C.c_double, C.c_double]
    L.orc_kernel_gradW_factor.restype = C.c_double
    L.orc_kernel_gradW_factor.argtypes = [C.c_int, C.c_double, C.c_double]
    L.orc_eos_pressure.restype = C.c_double
    L.orc_eos_pressure.argtypes = [C.c_double] * 4
    L.orc_eos_soundspeed.restype = C.c_double
    L.orc_eos_soundspeed.argtypes = [C.c_double] * 4
    L.orc_hydro_force.argtypes = [C.POINTER(ForceParams), f64p, C.c_int, f64p, C.c_int, f64p]
    L.orc_result_clear.argtypes = [C.c_int, f64p, C.c_int]
    L.orc_hydro_force_walks.argtypes = [C.POINTER(ForceParams), C.c_int, f64p, i64p, f64p, i64p,
                                        f64p, C.c_int]
    L.orc_nbr_pred.restype = C.c_int
    L.orc_neighbours_brute.restype = C.c_int
    L.orc_neighbours_brute.argtypes = [C.c_int, C.c_int64, f64p, f32p, i32p, C.c_void_p, C.c_int]
    L.orc_neighbours_brute_cells.restype = C.c_int
    L.orc_neighbours_brute_cells.argtypes = [C.c_int, C.c_int64, f64p, f32p, f64p, C.c_double, i32p, i32p,
                                             C.c_void_p, C.c_int]
    L.orc_neighbour_counts_cells.restype = C.c_int64
    L.orc_neighbour_counts_cells.argtypes = [C.c_int, C.c_int64, f64p, f32p, f64p, C.c_double, i32p, i32p,
                                             C.c_int]
    L.orc_make_walks.restype = C.c_int64
    L.orc_make_walks.argtypes = [C.c_int, C.c_int64, f64p, C.c_double, C.c_int,
                                 C.POINTER(C.c_int64), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_shepard.argtypes = [C.c_int, C.c_int64, f64p, f64p, f64p, f64p, f64p, C.c_int]
    L.orc_ic_dam_break.restype = C.c_int64
    L.orc_ic_dam_break.argtypes = [C.c_int] + [C.c_double] * 6 + [C.c_int] + [C.c_double] * 5 + \
                                  [C.POINTER(_CState)]
    L.orc_sim_init.argtypes = [C.POINTER(SimParams), C.POINTER(_CState), C.c_int]
    L.orc_sim_step.restype = C.c_int64
    L.orc_sim_step.argtypes = [C.POINTER(SimParams), C.POINTER(_CState), C.c_int]
    L.orc_sim_eos.argtypes = [C.POINTER(SimParams), C.POINTER(_CState)]
    L.orc_sim_force.restype = C.c_int64
    L.orc_sim_force.argtypes = [C.POINTER(SimParams), C.POINTER(_CState), C.c_int]
    L.orc_sim_force_scales.argtypes = [C.POINTER(SimParams), C.POINTER(_CState), C.c_int, f64p, f64p]
    L.orc_global_dt.restype = C.c_double
    L.orc_global_dt.argtypes = [C.POINTER(SimParams), C.POINTER(_CState)]
    L.orc_conservation.argtypes = [C.POINTER(SimParams), C.POINTER(_CState), f64p]
    L.orc_dam_front.restype = C.c_double
    L.orc_dam_front.argtypes = [C.POINTER(_CState), C.c_int, C.c_double]
    return L


# ----------------------------------------------------------------------------
# record layouts of the FDPS boundary types (src/class.h:63-67,167-176,212-223)
# ----------------------------------------------------------------------------
def epi_dtype(dim):
    return np.dtype([("pos", "<f8", (dim,)), ("vel", "<f8", (dim,)), ("smth", "<f8"), ("dens", "<f8"),
                     ("pres", "<f8"), ("snds", "<f8"), ("visc", "<f8"), ("tag", "<i8")])


def epj_dtype(dim):
    return np.dtype([("pos", "<f8", (dim,)), ("vel", "<f8", (dim,)), ("dens", "<f8"), ("mass", "<f8"),
                     ("smth", "<f8"), ("pres", "<f8"), ("snds", "<f8"), ("visc", "<f8"), ("tag", "<i8"),
                     ("type", "<i4"), ("_pad", "<i4")])


def res_dtype(dim):
    return np.dtype([("acc", "<f8", (dim,)), ("dt", "<f8"), ("div_v", "<f8")])


def force_params(dim, alpha=0.01, beta=0.0, eps=0.01, c_cfl=0.3, vsig_mode=0, grad_mode=0):
    """Defaults are the reference CPU functor's constants (src/force.h:32,45-47, src/param.h:5);
    grad_mode 1 takes the gradient as the reference GPU kernel does (src/kernel.cu:265)."""
    return ForceParams(dim, vsig_mode, alpha, beta, eps, c_cfl, grad_mode)


def kernel_W(dim, r, h):
    return lib().orc_kernel_W(dim, r, h)


def kernel_gradW_factor(dim, r, h):
    return lib().orc_kernel_gradW_factor(dim, r, h)


def eos_pressure(dens, rho0, c0, gamma):
    return lib().orc_eos_pressure(dens, rho0, c0, gamma)


def eos_soundspeed(dens, rho0, c0, gamma):
    return lib().orc_eos_soundspeed(dens, rho0, c0, gamma)


def _as_f64(rec):
    return np.ascontiguousarray(rec).view(np.float64).reshape(-1)


def hydro_force(params, epi, epj):
    """One FDPS i-group through src/force.h:31-60 (after RESULT::Hydro::clear)."""
    dim = params.dim
    res = np.zeros(len(epi), res_dtype(dim))
    r = _as_f64(res).copy()
    lib().orc_result_clear(dim, r, len(epi))
    lib().orc_hydro_force(C.byref(params), _as_f64(epi), len(epi), _as_f64(epj), len(epj), r)
    return r.view(res_dtype(dim))


def hydro_force_walks(params, epi, epi_off, epj, epj_off, nthreads=1):
    dim = params.dim
    r = np.zeros(len(epi) * (dim + 2), np.float64)
    lib().orc_hydro_force_walks(C.byref(params), len(epi_off) - 1, _as_f64(epi),
                                np.ascontiguousarray(epi_off, np.int64), _as_f64(epj),
                                np.ascontiguousarray(epj_off, np.int64), r, nthreads)
    return r.view(res_dtype(dim))


def neighbours_brute(pos, h, cap=0):
    """pos: (dim, n) float64; h: (n,) float32. Returns counts (and lists if cap)."""
    dim, n = pos.shape
    counts = np.zeros(n, np.int32)
    lists = np.full((n, cap), -1, np.int32) if cap else None
    lib().orc_neighbours_brute(dim, n, np.ascontiguousarray(pos, np.float64),
                               np.ascontiguousarray(h, np.float32), counts,
                               lists.ctypes.data if cap else None, cap)
    return (counts, lists) if cap else counts


def neighbours_brute_cells(pos, h, lo, L, nc, cap=0):
    """Brute force with the cell-local FP32 predicate of the device path."""
    dim, n = pos.shape
    counts = np.zeros(n, np.int32)
    lists = np.full((n, cap), -1, np.int32) if cap else None
    lib().orc_neighbours_brute_cells(dim, n, np.ascontiguousarray(pos, np.float64),
                                     np.ascontiguousarray(h, np.float32), np.ascontiguousarray(lo[:dim], np.float64),
                                     float(L), np.ascontiguousarray(nc[:dim], np.int32), counts,
                                     lists.ctypes.data if cap else None, cap)
    return (counts, lists) if cap else counts


def neighbour_counts_cells(pos, h, lo, L, nc, nthreads=1):
    """Counts of the cell-local predicate through a cell list (large states;
    every particle inside the grid, else ValueError)."""
    dim, n = pos.shape
    counts = np.zeros(n, np.int32)
    tot = lib().orc_neighbour_counts_cells(dim, n, np.ascontiguousarray(pos, np.float64),
                                           np.ascontiguousarray(h, np.float32),
                                           np.ascontiguousarray(lo[:dim], np.float64), float(L),
                                           np.ascontiguousarray(nc[:dim], np.int32), counts, nthreads)
    if tot < 0:
        raise ValueError("a particle lies outside the grid")
    return counts


def make_walks(pos, r_search, group_max=128):
    """Fake FDPS multiwalk grouping. Returns (grp_off, grp_idx, nbr_off, nbr_idx)."""
    dim, n = pos.shape
    pos = np.ascontiguousarray(pos, np.float64)
    nw = C.c_int64()
    tot = lib().orc_make_walks(dim, n, pos, r_search, group_max, C.byref(nw), None, None, None, None)
    go = np.zeros(nw.value + 1, np.int64)
    gi = np.zeros(n, np.int64)
    no = np.zeros(nw.value + 1, np.int64)
    ni = np.zeros(tot, np.int64)
    lib().orc_make_walks(dim, n, pos, r_search, group_max, C.byref(nw), go.ctypes.data,
                         gi.ctypes.data, no.ctypes.data, ni.ctypes.data)
    return go, gi, no, ni


# ----------------------------------------------------------------------------
# whole-step oracle
# ----------------------------------------------------------------------------
def sim_params(dim, *, alpha=0.1, beta=0.0, eps=0.01, c_cfl=0.25, vsig_mode=0, rho0=1000.0,
               c0=None, col_h=2.0, gamma=7.0, g=9.81, nu=1.0e-6, rho_floor=0.0, shepard_interval=0):
    if c0 is None:
        c0 = 10.0 * np.sqrt(2.0 * g * col_h)
    return SimParams(dim, vsig_mode, alpha, beta, eps, c_cfl, rho0, c0, gamma, g, nu, rho_floor, shepard_interval)


def shepard(pos, h, mass, dens, nthreads=1):
    """Shepard density filter (src/force.h:10-30); pos (dim, n) FP64."""
    dim, n = pos.shape
    out = np.zeros(n)
    lib().orc_shepard(dim, n, np.ascontiguousarray(pos, np.float64), np.ascontiguousarray(h, np.float64),
                      np.ascontiguousarray(mass, np.float64), np.ascontiguousarray(dens, np.float64), out, nthreads)
    return out


class State:
    """SoA FP64 particle state owned by numpy, viewed by the C oracle."""

    def __init__(self, dim, n):
        self.dim, self.n = dim, n
        self.x = np.zeros((dim, n)); self.v = np.zeros((dim, n)); self.vh = np.zeros((dim, n))
        self.a = np.zeros((dim, n))
        for k in ("dens", "pres", "snds", "mass", "h", "divv", "dt_i"):
            setattr(self, k, np.zeros(n))
        self.type = np.zeros(n, np.int32)
        self.c = _CState(n, *[getattr(self, k).ctypes.data for k in
                              ("x", "v", "vh", "a", "dens", "pres", "snds", "mass", "h", "divv", "dt_i",
                               "type")], 0.0, 0.0, 0)

    @property
    def time(self):
        return self.c.time

    @property
    def step_count(self):
        return self.c.step

    @classmethod
    def dam_break(cls, dim, dx, *, col_w=1.0, col_h=2.0, depth=0.5, tank_l=4.0, tank_h=3.0, nwall=3,
                  rho0=1000.0, c0=None, gamma=7.0, g=9.81, hfac=1.3):
        if c0 is None:
            c0 = 10.0 * np.sqrt(2.0 * g * col_h)
        args = (dim, dx, col_w, col_h, depth, tank_l, tank_h, nwall, rho0, c0, gamma, g, hfac)
        n = lib().orc_ic_dam_break(*args, None)
        s = cls(dim, n)
        lib().orc_ic_dam_break(*args, C.byref(s.c))
        return s

    def init(self, p, nthreads=1):
        lib().orc_sim_init(C.byref(p), C.byref(self.c), nthreads)

    def step(self, p, nthreads=1):
        """One step; returns the pair interactions of its force pass."""
        return lib().orc_sim_step(C.byref(p), C.byref(self.c), nthreads)

    def eos(self, p):
        lib().orc_sim_eos(C.byref(p), C.byref(self.c))

    def force(self, p, nthreads=1):
        return lib().orc_sim_force(C.byref(p), C.byref(self.c), nthreads)

    def force_scales(self, p, nthreads=1):
        """Per-particle scales of the force pass's sums (the state is not
        changed): sum_j |a_ij|_2 and sum_j |div v_ij| / rho_i over the pair
        terms, each term alone. Per-particle errors are stated against them."""
        sa, sd = np.zeros(self.n), np.zeros(self.n)
        lib().orc_sim_force_scales(C.byref(p), C.byref(self.c), nthreads, sa, sd)
        return sa, sd

    def global_dt(self, p):
        return lib().orc_global_dt(C.byref(p), C.byref(self.c))

    def conservation(self, p):
        out = np.zeros(7)
        lib().orc_conservation(C.byref(p), C.byref(self.c), out)
        return out

    def dam_front(self, zmax):
        return lib().orc_dam_front(C.byref(self.c), self.dim, zmax)
