"""MI355X (gfx950) hydro force path for weakly compressible SPH.

Python mirror of the C-ABI in include/wcsph_abi.h (libwcsph.so):

* ``HydrDispatchKernel`` / ``HydrRetrieveKernel`` keep the reference's FDPS
  multiwalk callback names and argument meaning (src/prototype.h:7-10,
  src/kernel.cu:289,383) and, like them, return an int status;
* ``Simulation`` drives the device-resident step (cell list, neighbour and
  force passes, FP64 kick-drift-kick, Shepard filter), replacing the loop of
  src/main.cpp:89-133;
* ``write_snapshot_ascii`` / ``read_snapshot_ascii`` and
  ``Simulation.run_with_output`` write the reference's ASCII snapshots on its
  schedule (src/io.h:2-37); ``Simulation.save`` / ``Simulation.restore`` are
  the restart files.
"""
import os

import ctypes as C

import numpy as np

from ._abi import (PREC_FP32, PREC_FP64, PREC_MIXED, PRECISIONS, DamBreakDesc, Params, SimConfig, SimStats,
                   WcsphError, check, epi_dtype, epj_dtype, lib, ptr, res_dtype)

__all__ = ["Params", "SimConfig", "WcsphError", "reference_params", "dam_break_params", "init", "finalize",
           "HydrDispatchKernel", "HydrRetrieveKernel", "last_kernel_ms", "epi_dtype", "epj_dtype", "res_dtype",
           "dam_break_desc", "dam_break_state", "Simulation", "PREC_FP32", "PREC_MIXED", "PREC_FP64",
           "write_snapshot_ascii", "read_snapshot_ascii", "snapshot_path", "restart_path"]


def reference_params(dim, device=0):
    """Constants of the reference CPU functor and its dam-break problem."""
    p = Params()
    check(lib().wcsph_params_reference(dim, C.byref(p)))
    p.device = device
    return p


def dam_break_params(dim, *, col_h=2.0, alpha=0.1, beta=0.0, c_cfl=0.25, g=9.81, rho0=1000.0, gamma=7.0,
                     nu=1.0e-6, c0=None, rho_floor=0.0, vsig_mode=0, eps=0.01, shepard_interval=0, device=0,
                     grad_mode=0, wall_mode=0):
    """Physics of the dam-break configurations: c0 = 10 sqrt(2 g H), alpha = 0.1, CFL 0.25, g = 9.81;
    the Shepard filter is off unless shepard_interval > 0 (the reference uses 30); wall_mode 1 is
    the frozen-wall step (wcsph_params.wall_mode)."""
    p = reference_params(dim, device)
    p.alpha, p.beta, p.c_cfl, p.g, p.rho0, p.gamma, p.nu = alpha, beta, c_cfl, g, rho0, gamma, nu
    p.c0 = 10.0 * np.sqrt(2.0 * g * col_h) if c0 is None else c0
    p.rho_floor, p.vsig_mode, p.eps = rho_floor, vsig_mode, eps
    p.shepard_interval = shepard_interval
    p.grad_mode = grad_mode
    p.wall_mode = wall_mode
    return p


# ----------------------------------------------------------------------------
# FDPS multiwalk drop-in
# ----------------------------------------------------------------------------
def init(params):
    check(lib().wcsph_init(C.byref(params)))
    _dim_hint[0] = params.dim


def finalize():
    check(lib().wcsph_finalize())


def _ptr_array(arrays):
    return (C.c_void_p * max(len(arrays), 1))(*[a.ctypes.data if len(a) else None for a in arrays])


def HydrDispatchKernel(tag, n_walk, epi, n_epi, epj, n_epj):
    """GPU_WC::HydrDispatchKernel (src/kernel.cu:289). epi/epj: sequences of
    structured arrays in the EPI/EPJ layouts; returns 0 or an error code."""
    dim = epi[0].dtype["pos"].shape[0] if n_walk else lib_dim()
    fn = lib().wcsph_hydro_dispatch_3d if dim == 3 else lib().wcsph_hydro_dispatch_2d
    ni = np.ascontiguousarray(n_epi, np.int32)
    nj = np.ascontiguousarray(n_epj, np.int32)
    _keep = (_ptr_array(epi), _ptr_array(epj))
    return fn(tag, n_walk, C.cast(_keep[0], C.c_void_p), ptr(ni), C.cast(_keep[1], C.c_void_p), ptr(nj))


def HydrRetrieveKernel(tag, n_walk, ni, force):
    """GPU_WC::HydrRetrieveKernel (src/kernel.cu:383). force: sequence of
    writable RESULT arrays; returns 0 or an error code."""
    dim = force[0].dtype["acc"].shape[0] if n_walk else lib_dim()
    fn = lib().wcsph_hydro_retrieve_3d if dim == 3 else lib().wcsph_hydro_retrieve_2d
    n = np.ascontiguousarray(ni, np.int32)
    arr = _ptr_array(force)
    return fn(tag, n_walk, ptr(n), C.cast(arr, C.c_void_p))


_dim_hint = [3]


def lib_dim():
    return _dim_hint[0]


def last_kernel_ms(tag=0):
    ms = C.c_float()
    check(lib().wcsph_hydro_last_kernel_ms(tag, C.byref(ms)))
    return ms.value


def last_timing(tag=0):
    """Timing of the tag's last batch: host packing, device span (uploads,
    kernels, result copy), force kernels and host unpacking, in ms."""
    out = np.zeros(5)
    check(lib().wcsph_hydro_last_timing(tag, ptr(out)))
    return dict(pack_ms=out[0], device_ms=out[1], kernel_ms=out[2], unpack_ms=out[3], chunks=int(out[4]))


def copy_probe_gbps(nbytes=1 << 30, reps=10, device=0):
    """Measured HBM streaming-copy bandwidth (read + write GB/s)."""
    v = C.c_double()
    check(lib().wcsph_copy_probe(device, nbytes, reps, C.byref(v)))
    return v.value


def valu_probe_tflops(kind, reps=5, device=0):
    """Measured VALU FMA rate in TFLOP/s: kind "fp32" (v_fma_f32 with VGPR
    and constant operands), "pk_f16" (v_pk_fma_f16) or "fp32_sgpr"
    (v_fma_f32 with one wave-uniform SGPR operand)."""
    k = {"fp32": 0, "pk_f16": 1, "fp32_sgpr": 2}[kind]
    v = C.c_double()
    check(lib().wcsph_valu_probe(device, k, reps, C.byref(v)))
    return v.value


def last_error():
    return lib().wcsph_last_error().decode()


# ----------------------------------------------------------------------------
# dam-break initial condition (generated by the library, on the host)
# ----------------------------------------------------------------------------
def dam_break_desc(dim, dx, *, col_w=1.0, col_h=2.0, depth=0.5, tank_l=4.0, tank_h=3.0, nwall=3, hfac=1.3):
    return DamBreakDesc(dim, nwall, dx, col_w, col_h, depth, tank_l, tank_h, hfac)


def dam_break_state(desc, params):
    """Host SoA arrays of the dam-break lattice: dict(pos, vel, dens, mass, h, type)."""
    nf, nw = C.c_int64(), C.c_int64()
    check(lib().wcsph_dam_break_count(C.byref(desc), C.byref(nf), C.byref(nw)))
    n = nf.value + nw.value
    D = desc.dim
    st = dict(pos=np.zeros((D, n)), vel=np.zeros((D, n)), dens=np.zeros(n), mass=np.zeros(n), h=np.zeros(n),
              type=np.zeros(n, np.int32))
    check(lib().wcsph_dam_break_fill(C.byref(desc), C.byref(params), *[ptr(st[k]) for k in
                                                                       ("pos", "vel", "dens", "mass", "h", "type")]))
    st["n_fluid"], st["n_wall"] = nf.value, nw.value
    return st


# ----------------------------------------------------------------------------
# device-resident simulation
# ----------------------------------------------------------------------------
class Simulation:
    """One device-resident SPH system (all particle data in HBM)."""

    def __init__(self, params, n, *, precision="fp32", max_neighbours=128, box_lo=None, box_hi=None,
                 cells_per_support=0, queue_depth=0):
        self.params = params
        self.dim = params.dim
        self.n = int(n)
        cfg = SimConfig()
        cfg.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        cfg.max_neighbours = max_neighbours
        cfg.cells_per_support = cells_per_support
        cfg.queue_depth = queue_depth
        if box_lo is not None:
            for k in range(self.dim):
                cfg.box_lo[k], cfg.box_hi[k] = box_lo[k], box_hi[k]
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().wcsph_sim_create(C.byref(params), C.byref(cfg), self.n, C.byref(h)))
        self._h = h

    @classmethod
    def from_state(cls, params, state, **kw):
        sim = cls(params, len(state["dens"]), **kw)
        sim.upload(state)
        return sim

    def close(self):
        if getattr(self, "_h", None):
            check(lib().wcsph_sim_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, st):
        D, n = self.dim, self.n
        arrs = [np.ascontiguousarray(st["pos"], np.float64).reshape(D, n),
                np.ascontiguousarray(st["vel"], np.float64).reshape(D, n),
                np.ascontiguousarray(st["dens"], np.float64), np.ascontiguousarray(st["mass"], np.float64),
                np.ascontiguousarray(st["h"], np.float64), np.ascontiguousarray(st["type"], np.int32)]
        check(lib().wcsph_sim_upload(self._h, *[ptr(a) for a in arrs]))

    def init(self):
        check(lib().wcsph_sim_init(self._h))

    def step(self, nsteps=1):
        check(lib().wcsph_sim_step(self._h, nsteps))

    def force(self):
        check(lib().wcsph_sim_force(self._h))

    def sync(self):
        check(lib().wcsph_sim_sync(self._h))

    def stats(self):
        s = SimStats()
        check(lib().wcsph_sim_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in SimStats._fields_}

    def download(self):
        D, n = self.dim, self.n
        out = dict(pos=np.zeros((D, n)), vel=np.zeros((D, n)), dens=np.zeros(n), acc=np.zeros((D, n)),
                   div_v=np.zeros(n), dt=np.zeros(n))
        check(lib().wcsph_sim_download(self._h, *[ptr(out[k]) for k in ("pos", "vel", "dens", "acc", "div_v", "dt")]))
        return out

    def neighbours(self, cap=0):
        counts = np.zeros(self.n, np.int32)
        lists = np.zeros((self.n, cap), np.int32) if cap else None
        check(lib().wcsph_sim_neighbours(self._h, ptr(counts), ptr(lists), cap))
        return (counts, lists) if cap else counts

    def diagnostics(self, zfront, xcol, xwall=1.0e300):
        """FP64 device reductions: conserved totals, dam front (max x of fluid
        below zfront), column height (max height of fluid at x < xcol) and
        run-up (max height of fluid at x > xwall)."""
        out = np.zeros(10)
        check(lib().wcsph_sim_diagnostics(self._h, zfront, xcol, xwall, ptr(out)))
        keys = ["mass", "mom_x", "mom_y", "mom_z", "kinetic", "potential", "internal", "front", "column", "runup"]
        return dict(zip(keys, out))

    def sample_marks(self, marks, zfront, xcol, xwall=1.0e300, chunk=50, on_sample=None):
        """Step until the clock passes each time in `marks` (ascending),
        taking single steps into each mark so the sample falls within one dt
        past it, and return one diagnostics dict per mark (with its "time"
        and "step"). Between marks steps are queued `chunk` at a time;
        on_sample(stats, diagnostics), if given, sees every chunk."""
        out = []
        todo = list(marks)
        while todo:
            s = self.stats()
            self.step(1 if s["time"] + (chunk + 1) * s["dt"] >= todo[0] else chunk)
            s = self.stats()
            d = self.diagnostics(zfront, xcol, xwall)
            if on_sample is not None:
                on_sample(s, d)
            if s["time"] >= todo[0]:
                out.append(dict(d, time=s["time"], step=s["step"]))
                todo.pop(0)
        return out

    def set_graphs(self, enable=True):
        check(lib().wcsph_sim_set_graphs(self._h, 1 if enable else 0))

    def idle(self):
        v = C.c_int32()
        check(lib().wcsph_sim_query(self._h, C.byref(v)))
        return bool(v.value)

    def grid(self):
        lo = np.zeros(3)
        L = C.c_double()
        nc = np.zeros(3, np.int32)
        check(lib().wcsph_sim_grid(self._h, ptr(lo), C.byref(L), ptr(nc)))
        return lo, L.value, nc

    def set_timing(self, enable=True):
        check(lib().wcsph_sim_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        """Summed device ms since set_timing(True): stages and whole steps."""
        out = np.zeros(5)
        n = C.c_int64()
        check(lib().wcsph_sim_timing(self._h, ptr(out), C.byref(n)))
        keys = ["kick_bin_reorder", "neighbour", "force", "final", "step"]
        d = dict(zip(keys, out.tolist()))
        d["steps"] = n.value
        return d

    def last_force_ms(self):
        out = np.zeros(3, np.float32)
        check(lib().wcsph_sim_last_force_ms(self._h, ptr(out)))
        return {"binning": float(out[0]), "neighbour": float(out[1]), "force": float(out[2])}

    # -- snapshots and restart ------------------------------------------------
    def write_ascii(self, path):
        """The current state as the reference's ASCII snapshot (original
        order; tag = type; pressure and sound speed from the Tait law)."""
        check(lib().wcsph_sim_write_ascii(self._h, os.fsencode(path)))

    def save(self, path):
        """Restart file of the state after init() or step()."""
        check(lib().wcsph_sim_save(self._h, os.fsencode(path)))

    @classmethod
    def restore(cls, path, device=0):
        """A simulation resumed from a restart file; stepping it continues bit
        for bit as the saved one would have."""
        h = C.c_void_p()
        check(lib().wcsph_sim_restore(os.fsencode(path), device, C.byref(h)))
        sim = cls.__new__(cls)
        sim._h = h
        params, cfg, n = Params(), SimConfig(), C.c_int64()
        check(lib().wcsph_sim_params(h, C.byref(params), C.byref(cfg), C.byref(n)))
        sim.params, sim.cfg, sim.dim, sim.n = params, cfg, params.dim, n.value
        return sim

    def set_clock(self, time, step):
        """Clock after init() for a state read back from a snapshot (the
        reference resumes at the file's time and step + 1, src/io.h:60-84)."""
        check(lib().wcsph_sim_set_clock(self._h, float(time), int(step)))

    def run_with_output(self, end_time, n_snapshots, outdir, *, restart=True, nproc=1, rank=0):
        """Step to end_time writing snapshots on the reference's schedule
        (OutputFileWithTimeInterval, src/io.h:2-37, called at
        src/main.cpp:79,130): a snapshot now and after every step that
        reaches the next output time, which then advances by
        end_time / n_snapshots; stepping stops at the first step that reaches
        end_time (src/main.cpp:89). Snapshot k goes to snapshot_path(outdir, k)
        and, with restart=True, a restart file to restart_path(outdir, k),
        rotating over 10 like the reference's binary dumps. Steps are queued
        in chunks of half the estimated steps left to the next output time
        (the clock is read back after each chunk), then one at a time.
        Returns the number of snapshots written."""
        os.makedirs(outdir, exist_ok=True)
        interval = end_time / n_snapshots
        t_next = self.stats()["time"]
        k = 0
        while True:
            st = self.stats()
            if st["time"] >= t_next:
                self.write_ascii(snapshot_path(outdir, k, nproc, rank))
                if restart:
                    self.save(restart_path(outdir, k, nproc, rank))
                t_next += interval
                k += 1
            if st["time"] >= end_time:
                return k
            left = min(t_next, end_time) - st["time"]
            chunk = int(0.5 * left / st["dt"]) if st["dt"] > 0 else 0
            self.step(max(chunk, 1))

# ----------------------------------------------------------------------------
# snapshot files
# ----------------------------------------------------------------------------
def snapshot_path(outdir, index, nproc=1, rank=0):
    """FDPS's per-rank name for snapshot `index` (src/io.h:13-14):
    outdir/<index>_<nproc>_<rank>.dat, five digits each."""
    return os.path.join(outdir, f"{index:05d}_{nproc:05d}_{rank:05d}.dat")


def restart_path(outdir, index, nproc=1, rank=0):
    """Name of the restart file written with snapshot `index`, rotating over
    10 like the reference's binary dumps (src/io.h:19-20)."""
    return os.path.join(outdir, f"{index % 10:05d}_{nproc:05d}_{rank:05d}.bin")


def write_snapshot_ascii(path, time, mass, pos, vel, dens, *, snds=None, pres=None, id=None, tag=None):
    """Write the reference's ASCII snapshot (FileHeader + RealPtcl::writeAscii,
    src/class.h:45-58,129-135). pos/vel are [dim][n]."""
    pos = np.ascontiguousarray(pos, np.float64)
    dim, n = pos.shape
    arrs = dict(mass=np.ascontiguousarray(mass, np.float64), pos=pos,
                vel=np.ascontiguousarray(vel, np.float64).reshape(dim, n),
                dens=np.ascontiguousarray(dens, np.float64))
    for k, v, t in (("snds", snds, np.float64), ("pres", pres, np.float64), ("id", id, np.int64),
                    ("tag", tag, np.int64)):
        arrs[k] = None if v is None else np.ascontiguousarray(v, t)
    check(lib().wcsph_snapshot_write_ascii(os.fsencode(path), dim, n, float(time), ptr(arrs["id"]),
                                           ptr(arrs["tag"]), ptr(arrs["mass"]), ptr(arrs["pos"]), ptr(arrs["vel"]),
                                           ptr(arrs["dens"]), ptr(arrs["snds"]), ptr(arrs["pres"])))


def read_snapshot_ascii(path, dim):
    """Read a snapshot of that format: dict of time, id, tag, mass, pos
    [dim][n], vel [dim][n], dens, snds, pres."""
    n, t = C.c_int64(0), C.c_double(0.0)
    fn = os.fsencode(path)
    check(lib().wcsph_snapshot_read_ascii(fn, dim, C.byref(n), C.byref(t), *([None] * 8)))
    m = n.value
    out = dict(id=np.zeros(m, np.int64), tag=np.zeros(m, np.int64), mass=np.zeros(m), pos=np.zeros((dim, m)),
               vel=np.zeros((dim, m)), dens=np.zeros(m), snds=np.zeros(m), pres=np.zeros(m))
    check(lib().wcsph_snapshot_read_ascii(fn, dim, C.byref(n), C.byref(t),
                                          *[ptr(out[k]) for k in ("id", "tag", "mass", "pos", "vel", "dens", "snds",
                                                                  "pres")]))
    out["time"] = t.value
    return out


# ----------------------------------------------------------------------------
# serving loop
# ----------------------------------------------------------------------------
from ._abi import Request, RequestResult, ServeConfig  # noqa: E402


def request_geometry(dim, n_target, aspect, *, col_w=0.5, depth=0.25):
    """Dam-break geometry of a serving request: column col_w wide and
    aspect*col_w high (depth deep in 3D) in a tank 3 col_w long and
    1.5 aspect*col_w high, with dx chosen so the lattice holds about
    n_target particles (fluid and wall)."""
    col_h = aspect * col_w
    base = dict(col_w=col_w, col_h=col_h, depth=depth, tank_l=3 * col_w, tank_h=1.5 * col_h)
    vol = col_w * col_h * (depth if dim == 3 else 1.0)
    dx = (vol / (0.5 * n_target)) ** (1.0 / dim)
    for _ in range(6):  # secant-style refinement on the total count
        nf, nw = C.c_int64(), C.c_int64()
        check(lib().wcsph_dam_break_count(C.byref(dam_break_desc(dim, dx, **base)), C.byref(nf), C.byref(nw)))
        n = nf.value + nw.value
        if abs(n - n_target) < 0.02 * n_target:
            break
        dx *= (n / n_target) ** (1.0 / dim)
    return dam_break_desc(dim, dx, **base)


def make_request(dim, n_target, aspect, alpha, steps, arrival_s, precision="mixed", device=0):
    geom = request_geometry(dim, n_target, aspect)
    p = dam_break_params(dim, col_h=geom.col_h, alpha=alpha, device=device)
    r = Request()
    r.params, r.geom = p, geom
    r.precision = PRECISIONS[precision]
    r.steps, r.arrival_s = int(steps), float(arrival_s)
    return r


def poisson_workload(seed, count, rate, *, n_range=(20_000, 2_000_000), aspect=(1.0, 4.0), alpha=(0.02, 0.2),
                     steps=(500, 5000), dims=(2, 3), precision="mixed"):
    """Seeded request stream: Poisson arrivals at `rate` per second, dimension
    uniform over dims, particle count log-uniform over n_range, column
    aspect, alpha and step count uniform over their ranges."""
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        dim = int(rng.choice(dims))
        n = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        reqs.append(make_request(dim, n, rng.uniform(*aspect), rng.uniform(*alpha), int(rng.integers(steps[0],
                                 steps[1] + 1)), t, precision))
    return reqs


def request_bytes(r, precision=None):
    """Exact device bytes of request r as the serving loop admits it (its
    initial state's grid computed on the host), at its own precision or the
    one named."""
    st = dam_break_state(r.geom, r.params)
    cfg = SimConfig()
    cfg.precision = r.precision if precision is None else PRECISIONS[precision]
    b = lib().wcsph_sim_bytes_state(C.byref(r.params), C.byref(cfg), len(st["dens"]), ptr(st["pos"]), ptr(st["h"]))
    if b < 0:
        raise WcsphError(1, "wcsph_sim_bytes_state rejected the request")
    return b


def serve(requests, *, hbm_budget_bytes=64 << 30, max_slots=4, chunk_steps=50):
    """Run a request stream through the continuous-batching loop. Returns one
    dict per request (same order): arrival/start/end wall times, latency,
    particles, steps, pair interactions and the final diagnostics."""
    n = len(requests)
    arr = (Request * max(n, 1))(*requests)
    out = (RequestResult * max(n, 1))()
    cfg = ServeConfig(int(hbm_budget_bytes), int(max_slots), int(chunk_steps))
    check(lib().wcsph_serve(C.byref(cfg), C.cast(arr, C.c_void_p), n, C.cast(out, C.c_void_p)))
    res = []
    for k in range(n):
        o = out[k]
        d = {f: getattr(o, f) for f, _ in RequestResult._fields_}
        d["latency_s"] = o.end_s - o.arrival_s
        d["dim"] = requests[k].geom.dim
        res.append(d)
    return res
