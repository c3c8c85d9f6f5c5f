#!/usr/bin/env python3
"""Throughput of the device-resident hydro force path on one MI355X.

Workload (default): the 3D dam break -- water column 1.0 x 0.5 x 2.0 m in a
4.0 x 0.5 x 3.0 m tank with 3 wall layers, dx = 0.0126 m (502,440 fluid +
562,938 wall = 1,065,378 particles), h = 1.3 dx, alpha = 0.1,
c0 = 10 sqrt(2 g H), CFL 0.25 -- stepped on the GPU. A step is one full
time step: kick-drift, cell-list rebuild by counting sort, neighbour pass,
force pass, gravity/final kick and the global dt reduction, all in HBM.

value = neighbour pair interactions evaluated in the timed steps / wall time
of those steps (replayed from the library's two-step hipGraphs, as users
step), for the mixed-precision path (packed FP16 pair geometry, FP32
density, pressure and accumulation, FP64 integration); the FP32 plain path
runs the same steps beside it ("paths", "mixed_over_fp32"). The timed
window starts at step 1,001 whatever --warmup is: the dam break is first
run for 1,000 - warmup steps into its developed flow (--develop), then the
warmup steps, then --steps timed steps (default: steps 1,001-2,000 of the
configuration's 2,000). The roofline entry is for the headline force
kernel, from HIP events recorded on the simulation's stream around each
launch of a following window of up to 100 steps launched one by one; the
measured VALU FMA rates (FP32 and packed FP16) and the VALU issue fraction
from the committed rocprofv3 summary sit beside it. The CPU baseline times
the FP64 oracle's force pass (the reference CPU functor, src/force.h:31-60,
over cell-list candidates, OpenMP) on the same developed state, on the
host's cores, built there with -O3 -march=native and again with the
reference's -ffast-math -funroll-loops. The "fdps" entry times the FDPS
drop-in (dispatch + retrieve, host records in and out) on the 3D state after
200 steps. The serving entry runs a seeded Poisson stream of 16 dam-break
requests at 8/s through the continuous-batching loop (grouped step kernels,
16 slots), batched and one at a time (p50/p99 completion latency, aggregate
pairs/s); "serving_sustained" runs 48 requests at 3/s, near the rate the
mix sustains.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = {3: 73, 2: 63}       # algorithmic flops per pair interaction
PEAK_VALU_TFLOPS = 157.3             # MI355X published FP32 / packed-FP16 vector peak
PEAK_HBM_GBS = 8000.0


def host_cpu():
    """The host's CPU as the baseline record states it: logical CPUs (nproc),
    the CPUs this process may run on, the model name and the physical core
    count (distinct (package, core) pairs of /proc/cpuinfo)."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "model": None,
            "physical_cores": None}
    try:
        cores, pkg, model = set(), None, None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                pkg = v
            elif k == "core id":
                cores.add((pkg, v))
        info["model"] = model
        info["physical_cores"] = len(cores) or None
    except OSError:
        pass
    return info


def oracle_force_rate(O, z, budget_s, threads):
    dim, n = int(z["dim"]), len(z["dens"])
    s = O.State(dim, n)
    s.x[:] = z["pos"]
    s.v[:] = z["vel"]
    s.dens[:] = z["dens"]
    s.mass[:] = z["mass"]
    s.h[:] = z["h"]
    s.type[:] = z["type"]
    op = O.sim_params(dim, **{k: float(z[k]) for k in ("alpha", "beta", "eps", "c_cfl", "rho0", "c0", "gamma", "g",
                                                        "nu", "rho_floor")}, vsig_mode=int(z["vsig_mode"]))
    s.eos(op)
    reps, pairs, t = 0, 0, 0.0
    while t < budget_s or reps == 0:
        t0 = time.perf_counter()
        pairs += s.force(op, threads)
        t += time.perf_counter() - t0
        reps += 1
        if t > 0.5 * budget_s and t / reps * (reps + 1) > 1.5 * budget_s:
            break
    return pairs / t, reps, t


def cpu_baseline_worker(path, budget_s, threads):
    """Child process of cpu_baseline: builds the oracle variants on this host
    and times their force passes on the state saved at `path`; prints JSON."""
    import tempfile
    import numpy as np
    import oracle as O
    z = dict(np.load(path))
    builds = {"O3_native": ["-O3", "-march=native"],
              "O3_native_fast_math": ["-O3", "-march=native", "-ffast-math", "-funroll-loops"]}
    rates = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, flags in builds.items():
            O.use_library(O.build_variant(tmp, name, flags))
            rate, reps, t = oracle_force_rate(O, z, budget_s / len(builds), threads)
            rates[name] = {"value": rate, "passes": reps, "seconds": t, "flags": " ".join(flags + ["-fopenmp"])}
    print(json.dumps(rates), flush=True)


def cpu_baseline(sim, params, budget_s, threads):
    """The oracle's FP64 force pass (src/force.h:31-60 over 27-cell
    candidates, OpenMP) on the developed state, built on this host with
    -O3 -march=native, and again with the reference's -ffast-math
    -funroll-loops (src/Makefile:3); value is the first. It runs in a child
    process with OMP_PROC_BIND=close and OMP_PLACES=cores: OpenMP binds the
    thread that starts it, and threads this process creates afterwards
    would inherit that binding."""
    import subprocess
    import tempfile
    import numpy as np
    out = sim.download()
    st0 = sim.initial_state
    host = host_cpu()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "state.npz")
        np.savez(path, dim=sim.dim, pos=out["pos"], vel=out["vel"], dens=out["dens"], mass=st0["mass"],
                 h=st0["h"], type=st0["type"], alpha=params.alpha, beta=params.beta, eps=params.eps,
                 c_cfl=params.c_cfl, rho0=params.rho0, c0=params.c0, gamma=params.gamma, g=params.g, nu=params.nu,
                 rho_floor=params.rho_floor, vsig_mode=params.vsig_mode)
        env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", OMP_NUM_THREADS=str(threads))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", path,
                            "--cpu-seconds", str(budget_s), "--cpu-threads", str(threads)],
                           env=env, check=True, capture_output=True, text=True)
    rates = json.loads(r.stdout.strip().splitlines()[-1])
    head = rates["O3_native"]
    return {"value": head["value"], "unit": "pair-interactions/s", "cores": threads, "kind": "port",
            "threads_used": threads, "omp_proc_bind": "close", "omp_places": "cores", **host, "builds": rates,
            "sample": f"{head['passes']} FP64 force pass(es) of the oracle (src/force.h:31-60 over 27-cell "
                      f"candidates) on the developed {sim.n}-particle state, {head['seconds']:.1f} s, built "
                      f"-O3 -march=native -fopenmp on this host; the -ffast-math -funroll-loops build "
                      f"(src/Makefile:3) is under builds"}


def cpu_baseline_c0(threads, steps=1000):
    """The oracle's full 2D dam break at dx = 0.02 m (5,000 fluid + 1,518 wall
    particles, alpha 0.1, CFL 0.25, FP64): init, then `steps` steps, each
    with its dt, kick-drift, EoS, force pass and final kick."""
    import oracle as O
    s = O.State.dam_break(2, 0.02)
    op = O.sim_params(2)
    t0 = time.perf_counter()
    s.init(op, threads)
    pairs = sum(s.step(op, threads) for _ in range(steps))
    t = time.perf_counter() - t0
    return {"value": pairs / t, "unit": "pair-interactions/s", "cores": threads, "kind": "port",
            "steps_per_s": steps / t, "particles": s.n,
            "sample": f"full run of the 2D dam break dx=0.02 ({s.n} particles, {steps} FP64 steps of the oracle: "
                      f"src/main.cpp:89-133 step order, src/force.h:31-60 pair functor), {t:.2f} s"}


def accuracy_study(W, dim, dx, marks, chunk, early=100):
    """Mixed and FP32 paths against the FP64 GPU path from the same initial
    condition: dam front (max x of fluid within 2 dx of the floor), residual
    column height (fluid within 2 dx of the left wall), run-up on the far
    wall (fluid within 2 dx of it) at the time marks, the same three over the
    bulk fluid (rho >= 0.9 rho0, without spray) at the first step past each
    mark, energy drift, and particle-wise errors after `early` steps (later,
    particle trajectories of any two precisions decorrelate). Returns one
    JSON-able dict."""
    import numpy as np
    p = W.dam_break_params(dim)
    desc = W.dam_break_desc(dim, dx)
    st = W.dam_break_state(desc, p)
    band, xwall = 2 * dx, desc.tank_l - 2 * dx
    fluid = st["type"] == 0
    runs = {}

    def bulk(d):
        # the same three measures over the bulk fluid (rho >= 0.9 rho0): spray
        # particles, which have expanded to a fraction of rho0, are left out
        x, z = d["pos"][0], d["pos"][dim - 1]
        b = fluid & (d["dens"] >= 0.9 * p.rho0)
        sel = lambda m, v: float(v[m].max()) if m.any() else -1e300
        return [sel(b & (z < band), x), sel(b & (x < band), z), sel(b & (x > xwall), z)]

    for prec in ("fp64", "fp32", "mixed"):
        t_start = time.perf_counter()
        sim = W.Simulation.from_state(p, st, precision=prec)
        sim.init()
        rows, todo, last, bulk_rows = [], list(marks), time.time(), []
        sim.step(early)  # particle-wise errors are compared here, before the flow
        snap = sim.download()  # becomes sensitive to rounding
        while todo:
            s = sim.stats()
            # single steps into each mark, so every path samples it within one dt
            sim.step(1 if s["time"] + (chunk + 1) * s["dt"] >= todo[0] else chunk)
            s = sim.stats()
            d = sim.diagnostics(band, band, xwall)
            rows.append([s["time"], d["front"], d["column"], d["runup"], d["kinetic"] + d["potential"] + d["internal"],
                         d["mass"], s["step"]])
            if s["time"] >= todo[0]:
                bulk_rows.append([s["time"]] + bulk(sim.download()))
                todo.pop(0)
            if time.time() - last > 15:
                print(f"# {prec}: t={s['time']:.3f} s, step {s['step']}", file=sys.stderr, flush=True)
                last = time.time()
        runs[prec] = {"rows": np.array(rows), "snap": snap, "wall_s": time.perf_counter() - t_start,
                      "steps": int(rows[-1][6]), "bulk": np.array(bulk_rows)}
        sim.close()
    ref = runs["fp64"]
    out = {"workload": f"{dim}D dam break dx={dx}, {len(st['dens'])} particles", "marks_s": list(marks), "paths": {}}
    for prec, r in runs.items():
        R = r["rows"]
        e0 = R[0, 4]
        entry = {"steps": r["steps"], "wall_s": r["wall_s"], "ms_per_step": 1e3 * r["wall_s"] / r["steps"],
                 "energy_drift": float((R[-1, 4] - e0) / abs(e0)), "mass_drift": float((R[-1, 5] - R[0, 5]) / R[0, 5])}
        for k, name in ((1, "front"), (2, "column"), (3, "runup")):
            vals = np.interp(marks, R[:, 0], R[:, k])
            entry[name] = [float(v) for v in vals]
            if prec != "fp64":
                rv = np.interp(marks, ref["rows"][:, 0], ref["rows"][:, k])
                entry[name + "_rel_err"] = [float(abs(a - b) / abs(b)) for a, b in zip(vals, rv)]
        Bk = r["bulk"]
        entry["bulk_at_marks"] = {"time": [float(t) for t in Bk[:, 0]],
                                  **{name: [float(v) for v in Bk[:, k]] for k, name in
                                     ((1, "front"), (2, "column"), (3, "runup"))}}
        if prec != "fp64":
            Rb = ref["bulk"]
            for k, name in ((1, "front"), (2, "column"), (3, "runup")):
                entry["bulk_at_marks"][name + "_rel_err"] = [
                    float(abs(a - b) / abs(b)) if b > -1e299 and a > -1e299 else None for a, b in zip(Bk[:, k], Rb[:, k])]
        if prec != "fp64":
            a, b = r["snap"], ref["snap"]
            disp = b["pos"] - st["pos"]
            entry[f"after_{early}_steps"] = {
                "pos_rel_err": float(np.linalg.norm(a["pos"] - b["pos"]) / np.linalg.norm(disp)),
                "vel_rel_err": float(np.linalg.norm(a["vel"] - b["vel"]) / np.linalg.norm(b["vel"])),
                "dens_rel_err": float(np.linalg.norm(a["dens"] - b["dens"]) / np.linalg.norm(b["dens"] - st["dens"])),
                "acc_rel_err": float(np.linalg.norm(a["acc"] - b["acc"]) / np.linalg.norm(b["acc"]))}
        out["paths"][prec] = entry
    return out


def serving_study(W, seed, count, rate, max_slots, chunk, precision, budget_gib, compare=True):
    """Continuous-batching run of a seeded Poisson request stream, and (with
    compare) the same stream served one request at a time (max_slots = 1)."""
    import numpy as np
    reqs = W.poisson_workload(seed, count, rate, precision=precision)
    runs, results = {}, []
    for slots in ((max_slots, 1) if compare else (max_slots,)):
        t0 = time.perf_counter()
        res = W.serve(reqs, hbm_budget_bytes=int(budget_gib * (1 << 30)), max_slots=slots, chunk_steps=chunk)
        wall = time.perf_counter() - t0
        lat = np.array([r["latency_s"] for r in res])
        pairs = sum(r["pairs"] for r in res)
        makespan = max(r["end_s"] for r in res)
        runs[f"slots_{slots}"] = {
            "makespan_s": makespan, "wall_s": wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "queue_wait_p50_s": float(np.median([r["start_s"] - r["arrival_s"] for r in res])),
            "pair_interactions": pairs, "pairs_per_s": pairs / makespan,
            "status": sorted({r["status"] for r in res}),
            "fp32_fallbacks": sum(r["precision_used"] != reqs[k].precision for k, r in enumerate(res)),
        }
        results.append(res)
        print(f"# slots {slots}: makespan {makespan:.2f} s", file=sys.stderr, flush=True)
    keys = ("steps", "pairs", "sim_time", "front", "column", "kinetic")
    identical = all(a[k] == b[k] for a, b in zip(*results) for k in keys) if compare else None
    return {"workload": f"{count} dam-break requests, Poisson arrivals at {rate}/s (seed {seed}), 2D/3D, "
                        f"20k-2M particles, 500-5000 steps, {precision}",
            "requests": [{"dim": r.geom.dim, "dx": r.geom.dx, "steps": r.steps, "arrival_s": r.arrival_s}
                         for r in reqs],
            "max_slots": max_slots, "chunk_steps": chunk, "hbm_budget_gib": budget_gib, "runs": runs,
            "results_identical_batched_vs_one_at_a_time": identical}


def microbench(W, dim, dx, reps=100, warm=10, seed=12345):
    """Force passes on a perturbed lattice: the dam-break lattice with seeded
    jitter U(+-0.05 dx) on fluid positions, velocities N(0, (0.02 c0)^2)
    and densities rho0 (1 + N(0, 1e-3^2)) on fluid particles. Each pass
    re-bins, reorders and evaluates the hydro force (wcsph_sim_force); the
    force kernel is timed with HIP events, warm passes first."""
    import numpy as np
    params = W.dam_break_params(dim)
    st = W.dam_break_state(W.dam_break_desc(dim, dx), params)
    rng = np.random.default_rng(seed)
    fl = st["type"] == 0
    nf = int(fl.sum())
    st["pos"][:, fl] += rng.uniform(-0.05, 0.05, (dim, nf)) * dx
    st["vel"][:, fl] = rng.normal(0.0, 0.02 * params.c0, (dim, nf))
    st["dens"][fl] = params.rho0 * (1.0 + rng.normal(0.0, 1e-3, nf))
    n = len(st["dens"])
    out = {"workload": f"{dim}D dam-break lattice dx={dx} m, {n} particles, jittered (seed {seed}), "
                       f"{reps} force passes after {warm}", "paths": {}}
    for prec in ("mixed", "fp32"):
        with W.Simulation.from_state(params, st, precision=prec) as sim:
            for _ in range(warm):
                sim.force()
            f_ms, b_ms = [], []
            for _ in range(reps):
                sim.force()
                t = sim.last_force_ms()
                f_ms.append(t["force"])
                b_ms.append(t["binning"])
            pairs = sim.stats()["pairs_last"]
        f = float(np.mean(f_ms))
        out["paths"][prec] = {"pairs_per_pass": pairs, "force_ms": f, "force_ms_min": float(np.min(f_ms)),
                              "binning_reorder_ms": float(np.mean(b_ms)),
                              "force_kernel_pairs_per_s": pairs / (f * 1e-3),
                              "achieved_tflops": FLOP_PER_PAIR[dim] * pairs / (f * 1e-3) / 1e12}
    out["mixed_over_fp32"] = out["paths"]["mixed"]["force_kernel_pairs_per_s"] / \
        out["paths"]["fp32"]["force_kernel_pairs_per_s"]
    return out


def fdps_study(W, dim, dx, warmup, group_edge=5.0):
    """The FDPS drop-in (GPU_WC::HydrDispatchKernel / HydrRetrieveKernel)
    end to end on a developed state: host records in, host results out,
    PCIe and host packing included. FDPS itself is not here, so the walks
    are formed FDPS-style: particles grouped by cubes of group_edge * dx
    (about 128 per group in the fluid, FDPS's Ngroup), each group's j list
    all particles within its bounding box grown by 2 h_max (FDPS's r_search
    of src/class.h:123-125)."""
    import numpy as np
    params = W.dam_break_params(dim)
    st = W.dam_break_state(W.dam_break_desc(dim, dx), params)
    with W.Simulation.from_state(params, st, precision="fp32") as sim:
        sim.init()
        sim.step(warmup)
        sim.force()
        out = sim.download()
        pairs = sim.stats()["pairs_last"]
    n = out["dens"].shape[0]
    h = st["h"][0]
    rho = out["dens"]
    B = params.c0 ** 2 * params.rho0 / params.gamma
    r = rho / params.rho0
    pres = np.maximum(B * (r ** params.gamma - 1.0), 0.0)
    snds = np.sqrt(np.maximum(B * params.gamma / params.rho0 * r ** (params.gamma - 1.0), 1e-16))
    epi_all = np.zeros(n, W.epi_dtype(dim))
    epj_all = np.zeros(n, W.epj_dtype(dim))
    for a in (epi_all, epj_all):
        a["pos"], a["vel"], a["smth"], a["dens"], a["pres"], a["snds"], a["visc"] = \
            out["pos"].T, out["vel"].T, h, rho, pres, snds, params.nu
    epj_all["mass"] = st["mass"]
    t0 = time.perf_counter()
    x = out["pos"].T
    G = group_edge * dx
    lo = x.min(axis=0)
    c = np.floor((x - lo) / G).astype(np.int64)
    nc = c.max(axis=0) + 1
    key = c[:, 0]
    for d in range(1, dim):
        key = key * nc[d] + c[:, d]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    bounds = np.flatnonzero(np.diff(ks)) + 1
    groups = np.split(order, bounds)
    cell_of = {int(k): g for k, g in zip(ks[np.r_[0, bounds]], groups)}
    halo = 2.0 * h * (1.0 + 1e-6)
    epi, epj = [], []
    offs = [np.array(o) for o in np.ndindex(*([3] * dim))]
    for g in groups:
        cg = c[g[0]]
        cand = []
        for o in offs:
            cc = cg + o - 1
            if np.any(cc < 0) or np.any(cc >= nc):
                continue
            kk = cc[0]
            for d in range(1, dim):
                kk = kk * nc[d] + cc[d]
            m = cell_of.get(int(kk))
            if m is not None:
                cand.append(m)
        cand = np.concatenate(cand)
        blo, bhi = x[g].min(axis=0) - halo, x[g].max(axis=0) + halo
        inside = np.all((x[cand] >= blo) & (x[cand] <= bhi), axis=1)
        epi.append(epi_all[g])
        epj.append(epj_all[cand[inside]])
    t_build = time.perf_counter() - t0
    nw = len(epi)
    ni = [len(e) for e in epi]
    nj = [len(e) for e in epj]
    res = [np.zeros(k, W.res_dtype(dim)) for k in ni]
    W.init(params)
    limit = 16384  # FDPS's n_walk_limit (src/kernel.cuh:4)
    # the walk pointer arrays are built once, as FDPS hands them over: the
    # timed region is the two C calls per batch
    import ctypes as C
    lib = W._abi.lib()
    disp = lib.wcsph_hydro_dispatch_3d if dim == 3 else lib.wcsph_hydro_dispatch_2d
    retr = lib.wcsph_hydro_retrieve_3d if dim == 3 else lib.wcsph_hydro_retrieve_2d
    batches = []
    for b in range(0, nw, limit):
        e = min(nw, b + limit)
        pi = (C.c_void_p * (e - b))(*[a.ctypes.data for a in epi[b:e]])
        pj = (C.c_void_p * (e - b))(*[a.ctypes.data for a in epj[b:e]])
        pr = (C.c_void_p * (e - b))(*[a.ctypes.data for a in res[b:e]])
        nib = np.ascontiguousarray(ni[b:e], np.int32)
        njb = np.ascontiguousarray(nj[b:e], np.int32)
        batches.append((e - b, pi, pj, pr, nib, njb))
    runs = []
    for rep in range(3):
        parts = {"pack_ms": 0.0, "device_ms": 0.0, "kernel_ms": 0.0, "unpack_ms": 0.0, "chunks": 0}
        wall = 0.0
        for (m, pi, pj, pr, nib, njb) in batches:
            t0 = time.perf_counter()
            W.check(disp(0, m, C.cast(pi, C.c_void_p), nib.ctypes.data, C.cast(pj, C.c_void_p), njb.ctypes.data))
            W.check(retr(0, m, nib.ctypes.data, C.cast(pr, C.c_void_p)))
            wall += time.perf_counter() - t0
            for k, v in W.last_timing(0).items():
                parts[k] += v
        runs.append((wall, parts))
    W.finalize()
    wall, parts = min(runs, key=lambda r: r[0])
    kms = parts["kernel_ms"]
    cand_pairs = sum(a * b for a, b in zip(ni, nj))
    return {"workload": f"{dim}D dam break dx={dx}, {n} particles after {warmup} steps, FDPS-style walks "
                        f"({nw} groups of <= {max(ni)} i, {int(np.mean(nj))} j on average)",
            "pairs": pairs, "candidate_pairs": cand_pairs,
            "pairs_per_s_end_to_end": pairs / wall, "end_to_end_ms": 1e3 * wall,
            "kernel_ms": kms, "pairs_per_s_kernel": pairs / (kms * 1e-3),
            "host_records_MB": (sum(ni) * epi_all.itemsize + sum(nj) * epj_all.itemsize) / 1e6,
            "shipped_MB": (sum(ni) * (8 * dim + 20) + sum(nj) * (8 * dim + 24)) / 1e6,
            "stages_ms": parts,
            "walk_build_s": t_build,
            "note": "end to end = wall time of the dispatch + retrieve C calls, best of 3: host packing "
                    "of the AoS records, H2D, kernels, D2H and unpacking"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--develop", type=int, default=-1,
                    help="steps run before the warmup to bring the dam break into its developed flow "
                         "(default: 1000 - warmup, so the timed window starts at step 1001)")
    ap.add_argument("--precision", default="mixed", choices=["fp32", "mixed", "fp64"],
                    help="headline path (the mixed and FP32 paths are both run unless --only)")
    ap.add_argument("--only", action="store_true", help="run the headline path alone")
    ap.add_argument("--dim", type=int, default=3, choices=[2, 3])
    ap.add_argument("--dx", type=float, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget (0 skips it)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cells-per-support", type=int, default=0, help="1 or 2 (0: library default)")
    ap.add_argument("--shepard", type=int, default=0,
                    help="Shepard density filter every N steps (the reference uses 30; 0: off, the default)")
    ap.add_argument("--queue-depth", type=int, default=0, help="24 or 32 (0: library default)")
    ap.add_argument("--accuracy", action="store_true",
                    help="run the FP32 and mixed paths against the FP64 path to --t-end and print their errors")
    ap.add_argument("--t-end", type=float, default=1.0)
    ap.add_argument("--fdps", action="store_true",
                    help="time the FDPS drop-in end to end (host records, PCIe included) on a developed state")
    ap.add_argument("--microbench", action="store_true",
                    help="force passes on the jittered lattice (seed 12345), mixed and FP32")
    ap.add_argument("--serve", action="store_true",
                    help="run a seeded request stream through the serving loop (p50/p99 latency, makespan)")
    ap.add_argument("--serve-requests", type=int, default=16, help="requests of the serving run (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--sustained-requests", type=int, default=48,
                    help="requests of the default run's second serving stream, at --sustained-rate (0 skips it)")
    ap.add_argument("--sustained-rate", type=float, default=3.0,
                    help="arrival rate near what the mix sustains, requests/s")
    ap.add_argument("--serve-seed", type=int, default=2026)
    ap.add_argument("--cpu-baseline-worker", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_worker:
        cpu_baseline_worker(args.cpu_baseline_worker, args.cpu_seconds, args.cpu_threads or 1)
        return
    if args.develop < 0:
        args.develop = max(0, 1000 - args.warmup)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import weaklycompressiblesph_amd as W

    dim = args.dim
    dx = args.dx or (0.0126 if dim == 3 else 0.02)
    if args.fdps:
        print(json.dumps(fdps_study(W, dim, dx, args.warmup)), flush=True)
        return
    if args.microbench:
        print(json.dumps(microbench(W, dim, dx)), flush=True)
        return
    if args.serve:
        print(json.dumps(serving_study(W, args.serve_seed, args.serve_requests, args.serve_rate, args.serve_slots,
                                       50, args.precision, 64.0)), flush=True)
        return
    if args.accuracy:
        marks = [m for m in (0.25, 0.5, 1.0) if m <= args.t_end] or [args.t_end]
        print(json.dumps(accuracy_study(W, dim, dx, marks, 50)), flush=True)
        return
    params = W.dam_break_params(dim, shepard_interval=args.shepard)
    desc = W.dam_break_desc(dim, dx)
    st = W.dam_break_state(desc, params)
    n = len(st["dens"])
    print(f"# {dim}D dam break dx={dx}: {st['n_fluid']} fluid + {st['n_wall']} wall = {n} particles",
          file=sys.stderr, flush=True)
    # the headline path first, then the others it is compared with
    order = [args.precision] + [p for p in ("mixed", "fp32") if p != args.precision and not args.only]
    runs, keep = {}, None
    for prec in order:
        sim, r = run_path(W, params, st, prec, args)
        runs[prec] = r
        if keep is None:
            keep = sim
        else:
            sim.close()
    head = runs[args.precision]
    K = args.steps
    dtype = {"fp32": "fp32", "mixed": "fp16x2/fp32acc", "fp64": "fp64"}[args.precision]
    force_ms = head["force_ms_per_launch"]
    achieved = FLOP_PER_PAIR[dim] * head["force_pairs_per_launch"] / (force_ms * 1e-3) / 1e12
    kname = {"fp32": "k_nforce", "mixed": "k_nforce_mixed", "fp64": "k_force64"}[args.precision]
    line = {
        "metric": "pair-interactions/s",
        "value": head["value"],
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (dam-break lattice initial condition, developed by the warmup steps)",
        "config": {"workload": f"{dim}D dam break, dx={dx} m, {n} particles ({st['n_fluid']} fluid), "
                               f"{args.precision} force path, full device-resident steps",
                   "particles": n, "dx": dx, "precision": args.precision,
                   "cells_per_support": args.cells_per_support or 2, "shepard_interval": args.shepard,
                   "develop_steps": args.develop,
                   "step_window": f"steps {args.develop + args.warmup + 1}-{args.develop + args.warmup + K}"},
        "paths": runs,
        "roofline": {"bound": "valu", "kernel": kname, "achieved": achieved, "peak": PEAK_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_VALU_TFLOPS, "traffic": None,
                     "flop_per_pair": FLOP_PER_PAIR[dim], "avg_launch_ms": force_ms,
                     "note": "algorithmic flops of the pair body over the force launch time (HIP events on the "
                             "simulation stream); the peak is MI355X's published FP32/packed-FP16 vector rate"},
    }
    if "mixed" in runs and "fp32" in runs:
        line["mixed_over_fp32"] = {
            "step_pairs_per_s": runs["mixed"]["value"] / runs["fp32"]["value"],
            "force_kernel_pairs_per_s": runs["mixed"]["force_kernel_pairs_per_s"] /
                                        runs["fp32"]["force_kernel_pairs_per_s"]}
    if not args.only:
        # the same steps with frozen walls (wall_mode 1): waves of walls alone skip their wall-wall pairs,
        # which add nothing the step uses; every fluid result and dt is the same. Pairs here are the pairs
        # evaluated, so the figure is the step time, not the headline's pair rate.
        fw = {}
        for prec in ("mixed", "fp32"):
            pw = W.dam_break_params(dim, shepard_interval=args.shepard, wall_mode=1)
            sim, r = run_path(W, pw, st, prec, args)
            sim.close()
            fw[prec] = {"ms_per_step": r["ms_per_step"], "steps_per_s": r["steps_per_s"],
                        "force_ms_per_launch": r["force_ms_per_launch"],
                        "pairs_evaluated_per_step": r["pairs_per_step"],
                        "speedup_per_step_vs_wall_mode_0": runs[prec]["ms_per_step"] / r["ms_per_step"]}
        line["frozen_walls"] = fw
    if dim == 3 and abs(dx - 0.0126) < 1e-12:
        prof = profiled_traffic(kname)
        if prof:
            line["roofline"]["traffic"] = prof["bytes_per_launch"]
            line["roofline"]["traffic_source"] = prof["source"]
            if "valu_issue_frac" in prof:
                line["roofline"]["valu_issue_frac"] = prof["valu_issue_frac"]
    line["hbm_copy_GBps"] = W.copy_probe_gbps(1 << 30, 10)
    # measured VALU FMA rates: the attainable peaks of the instruction kinds the force passes use (FP32
    # FMA with VGPR operands, packed FP16 FMA) and FP32 FMA with a wave-uniform SGPR operand, which gfx950
    # issues at half the rate of the VGPR form
    peaks = {"fp32_fma": W.valu_probe_tflops("fp32"), "pk_f16_fma": W.valu_probe_tflops("pk_f16")}
    line["valu_probe_TFLOPs"] = dict(peaks, fp32_fma_sgpr_operand=W.valu_probe_tflops("fp32_sgpr"))
    line["roofline"]["frac_of_measured"] = {k: achieved / v for k, v in peaks.items()}
    if rank == 0 and args.cpu_seconds > 0:
        # the box's CPU share is 16 threads per GPU, whatever nproc shows
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)) or 1)
        print(f"# cpu baseline on {threads} threads", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(keep, params, args.cpu_seconds, threads)
        line["cpu_baseline_c0"] = cpu_baseline_c0(threads)
    keep.close()
    if rank == 0 and dim == 3 and not args.only:
        print("# FDPS drop-in", file=sys.stderr, flush=True)
        fd = fdps_study(W, 3, 0.0126, 200)
        line["fdps"] = {"workload": fd["workload"], "end_to_end_ms": fd["end_to_end_ms"],
                        "kernel_ms": fd["kernel_ms"], "pack_ms": fd["stages_ms"]["pack_ms"],
                        "unpack_ms": fd["stages_ms"]["unpack_ms"],
                        "pairs_per_s_end_to_end": fd["pairs_per_s_end_to_end"],
                        "pairs_per_s_kernel": fd["pairs_per_s_kernel"]}
    if rank == 0 and args.serve_requests > 0 and not args.only:
        print(f"# serving {args.serve_requests} requests", file=sys.stderr, flush=True)
        sv = serving_study(W, args.serve_seed, args.serve_requests, args.serve_rate, args.serve_slots, 50,
                           "mixed", 64.0, compare=True)
        line["serving"] = {"workload": sv["workload"], "max_slots": sv["max_slots"], "runs": sv["runs"],
                           "results_identical_batched_vs_one_at_a_time":
                               sv["results_identical_batched_vs_one_at_a_time"]}
        if args.sustained_requests > 0:
            print(f"# serving {args.sustained_requests} requests at {args.sustained_rate}/s", file=sys.stderr,
                  flush=True)
            su = serving_study(W, args.serve_seed, args.sustained_requests, args.sustained_rate, args.serve_slots,
                               50, "mixed", 64.0, compare=False)
            line["serving_sustained"] = {"workload": su["workload"], "max_slots": su["max_slots"],
                                         **su["runs"][f"slots_{args.serve_slots}"]}
    if rank == 0:
        print(json.dumps(line), flush=True)


def run_path(W, params, st, prec, args):
    """Warm a simulation up, then time args.steps full steps as a user runs
    them (two-step hipGraph replays, wall clock between synchronisations).
    A second window of up to 100 steps, launched one by one with HIP events
    around each stage on the simulation stream, gives the per-stage device
    times and the force kernel's launch duration. Returns the simulation and
    its figures."""
    sim = W.Simulation.from_state(params, st, precision=prec, cells_per_support=args.cells_per_support,
                                  queue_depth=args.queue_depth)
    sim.initial_state = st
    sim.init()
    t0 = time.perf_counter()
    sim.step(args.develop + args.warmup)
    sim.sync()
    print(f"# {prec}: {args.develop} developing + {args.warmup} warmup steps {time.perf_counter() - t0:.2f} s",
          file=sys.stderr, flush=True)
    s0 = sim.stats()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    t1 = time.perf_counter()
    s1 = sim.stats()
    K = args.steps
    wall = t1 - t0
    pairs = s1["pairs_total"] - s0["pairs_total"]
    # stage window (events disable the graph replays)
    Kt = max(2, min(K, 100))
    sim.set_timing(True)
    sim.step(Kt)
    tm = sim.timing()
    s2 = sim.stats()
    sim.set_timing(False)
    pairs_t = s2["pairs_total"] - s1["pairs_total"]
    force_ms = tm["force"] / Kt
    return sim, {
        "value": pairs / wall,
        "ms_per_step": 1e3 * wall / K,
        "steps_per_s": K / wall,
        "pairs_per_step": pairs / K,
        "force_ms_per_launch": force_ms,
        "force_pairs_per_launch": pairs_t / Kt,
        "force_kernel_pairs_per_s": pairs_t / Kt / (force_ms * 1e-3),
        "stage_window_steps": Kt,
        "stage_ms_per_step": {k: tm[k] / Kt for k in ("kick_bin_reorder", "neighbour", "force", "final", "step")},
        "sim_time_s": s1["time"],
        "escaped": s1["escaped"],
        "overflow": s1["overflow"],
    }


def profiled_traffic(kernel):
    """HBM bytes per launch of the force kernel on this workload, from the
    committed rocprofv3 summary: FETCH_SIZE x 2 + WRITE_SIZE (KiB per
    dispatch; the factors are those of a 1 GiB copy kernel read the same way)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_dam_break_3d_1.07M.md")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        f = line.split()
        if len(f) >= 4 and f[0].startswith(kernel + "<"):
            for k in ("FETCH_SIZE", "WRITE_SIZE"):
                if k in f and f.index(k) + 2 < len(f):  # counter, dispatches, mean per dispatch
                    vals[k] = float(f[f.index(k) + 2])
    if len(vals) < 2:
        return None
    b = (2.0 * vals["FETCH_SIZE"] + vals["WRITE_SIZE"]) * 1024.0
    out = {"bytes_per_launch": b, "source": os.path.relpath(path, ROOT)}
    # "... so 84% of the 256 CUs x 1 wave64 VALU instruction per cycle issue rate"
    m = re.search(r"so (\d+)% of the 256 CUs x 1 wave64 VALU instruction per cycle", open(path).read())
    if m:
        out["valu_issue_frac"] = int(m.group(1)) / 100.0
    return out


if __name__ == "__main__":
    main()
