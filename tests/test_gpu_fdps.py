"""GPU parity of the FDPS multiwalk drop-in (wcsph_hydro_dispatch/retrieve)
against the FP64 oracle's restatement of src/force.h:31-60.

Walks come from the oracle's fake FDPS driver (Morton-ordered i-groups of
at most 128, superset j lists, shuffled), batched and called in FDPS's order
(retrieve(k-1) before dispatch(k)). Tolerances: normwise
relative error <= 1e-5 on acc and div_v (drho/dt = -rho div_v); dt to 1e-6.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-5


def developed_state(O, W, dim, dx, steps, **geom):
    p = W.dam_break_params(dim)
    s = O.State.dam_break(dim, dx, c0=p.c0, g=p.g, **geom)
    op = O.sim_params(dim, alpha=p.alpha, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu)
    s.init(op, 8)
    for _ in range(steps):
        s.step(op, 8)
    return s, p


def records(O, s, nu, rng=None):
    dim, n = s.x.shape
    epi = np.zeros(n, O.epi_dtype(dim))
    epj = np.zeros(n, O.epj_dtype(dim))
    for a in (epi, epj):
        a["pos"], a["vel"], a["smth"], a["dens"], a["pres"], a["snds"], a["visc"] = \
            s.x.T, s.v.T, s.h, s.dens, s.pres, s.snds, nu
        a["tag"] = np.arange(n)
    epj["mass"] = s.mass
    epj["type"] = s.type
    if rng is not None:  # per-particle h and m variation exercises the mean-h kernel
        f = 1 + 0.1 * rng.uniform(-1, 1, n)
        epi["smth"] *= f
        epj["smth"] *= f
        epj["mass"] *= 1 + 0.2 * rng.uniform(-1, 1, n)
    return epi, epj


def walks(O, epi_all, epj_all, group_max, rng):
    pos = np.ascontiguousarray(epi_all["pos"].T)
    go, gi, no, ni = O.make_walks(pos, r_search=2 * epi_all["smth"].max(), group_max=group_max)
    epi, epj = [], []
    for w in range(len(go) - 1):
        epi.append(epi_all[gi[go[w]:go[w + 1]]])
        js = ni[no[w]:no[w + 1]].copy()
        rng.shuffle(js)
        epj.append(epj_all[js])
    return epi, epj


def run_fdps_order(W, epi, epj, limit, tag=0):
    res = [np.zeros(len(e), W.res_dtype(e.dtype["pos"].shape[0])) for e in epi]
    nw = len(epi)
    prev = None
    for b in range(0, nw, limit):
        e = slice(b, min(nw, b + limit))
        if prev is not None:
            assert W.HydrRetrieveKernel(tag, prev.stop - prev.start, [len(x) for x in epi[prev]], res[prev]) == 0, \
                W.last_error()
        assert W.HydrDispatchKernel(tag, e.stop - e.start, epi[e], [len(x) for x in epi[e]], epj[e],
                                    [len(x) for x in epj[e]]) == 0, W.last_error()
        prev = e
    assert W.HydrRetrieveKernel(tag, prev.stop - prev.start, [len(x) for x in epi[prev]], res[prev]) == 0
    return res


def oracle_walks(O, fp, epi, epj):
    io = np.concatenate([[0], np.cumsum([len(e) for e in epi])])
    jo = np.concatenate([[0], np.cumsum([len(e) for e in epj])])
    return O.hydro_force_walks(fp, np.concatenate(epi), io, np.concatenate(epj), jo, nthreads=8)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


CASES = [
    # dim, dx, steps, param set
    (2, 0.02, 40, "reference"),
    (2, 0.02, 40, "config"),
    (3, 0.04, 30, "reference"),
    (3, 0.04, 30, "config-beta-vsig"),
    # grad_mode 1: the reference GPU kernel's mean of the gradients at h_i
    # and h_j (src/kernel.cu:265), with per-particle h varying by 10%
    (2, 0.02, 40, "config-grad"),
    (3, 0.04, 30, "config-grad-vsig"),
]


@pytest.mark.parametrize("dim,dx,steps,pset", CASES)
def test_dispatch_retrieve_matches_oracle(O, W, dim, dx, steps, pset):
    rng = np.random.default_rng(7 + dim)
    s, p = developed_state(O, W, dim, dx, steps, **({} if dim == 2 else dict(tank_l=2.0, col_h=1.0, tank_h=1.5)))
    if pset == "reference":
        q = W.reference_params(dim)
    else:
        q = W.dam_break_params(dim, alpha=0.1, beta=0.2 if "beta" in pset else 0.0,
                               vsig_mode=1 if "vsig" in pset else 0, grad_mode=1 if "grad" in pset else 0)
    fp = O.force_params(dim, alpha=q.alpha, beta=q.beta, eps=q.eps, c_cfl=q.c_cfl, vsig_mode=q.vsig_mode,
                        grad_mode=q.grad_mode)
    epi_all, epj_all = records(O, s, p.nu, rng if pset != "reference" else None)
    epi, epj = walks(O, epi_all, epj_all, 128, rng)
    W.init(q)
    got = np.concatenate(run_fdps_order(W, epi, epj, limit=37))
    ref = oracle_walks(O, fp, epi, epj)
    W.finalize()
    assert rel(got["acc"], ref["acc"]) <= TOL
    assert rel(got["div_v"], ref["div_v"]) <= TOL
    assert np.allclose(got["dt"], ref["dt"], rtol=1e-6, atol=0)


def test_dispatch_contract_errors(O, W):
    s, p = developed_state(O, W, 2, 0.05, 2)
    epi_all, epj_all = records(O, s, p.nu)
    epi, epj = walks(O, epi_all, epj_all, 64, np.random.default_rng(1))
    ni = [len(e) for e in epi]
    nj = [len(e) for e in epj]
    W.init(W.reference_params(2))
    res = [np.zeros(len(e), W.res_dtype(2)) for e in epi]
    E = W._abi
    assert W.HydrRetrieveKernel(0, len(epi), ni, res) == E.WCSPH_ESTATE      # nothing dispatched
    assert W.HydrDispatchKernel(0, len(epi), epi, ni, epj, nj) == 0
    assert W.HydrDispatchKernel(0, len(epi), epi, ni, epj, nj) == E.WCSPH_ESTATE  # one batch per tag
    bad = list(ni)
    bad[0] += 1
    assert W.HydrRetrieveKernel(0, len(epi), bad, res) == E.WCSPH_EINVAL      # sizes differ
    assert W.HydrRetrieveKernel(0, len(epi), ni, res) == 0                    # still retrievable
    # two tags in flight at once
    assert W.HydrDispatchKernel(1, len(epi), epi, ni, epj, nj) == 0
    assert W.HydrDispatchKernel(2, len(epi), epi, ni, epj, nj) == 0
    r1 = [np.zeros(len(e), W.res_dtype(2)) for e in epi]
    r2 = [np.zeros(len(e), W.res_dtype(2)) for e in epi]
    assert W.HydrRetrieveKernel(2, len(epi), ni, r2) == 0
    assert W.HydrRetrieveKernel(1, len(epi), ni, r1) == 0
    a, b, c = (np.concatenate(x) for x in (res, r1, r2))
    assert np.array_equal(a, b) and np.array_equal(a, c)                   # deterministic
    assert W.last_kernel_ms(1) > 0
    # empty batch and empty walks
    assert W.HydrDispatchKernel(0, 0, [], [], [], []) == 0
    assert W.HydrRetrieveKernel(0, 0, [], []) == 0
    e0 = [epi[0][:0], epi[1]]
    j0 = [epj[0][:0], epj[1]]
    r = [np.zeros(0, W.res_dtype(2)), np.zeros(len(epi[1]), W.res_dtype(2))]
    assert W.HydrDispatchKernel(0, 2, e0, [0, len(epi[1])], j0, [0, len(epj[1])]) == 0
    assert W.HydrRetrieveKernel(0, 2, [0, len(epi[1])], r) == 0
    assert np.array_equal(r[1], res[1])
    # a walk with i particles and no j candidates: the functor's loop is
    # empty (src/force.h:31-60), so acc = 0, div_v = 0, dt = C_CFL h_i / c_i;
    # retrieve assigns every field (src/kernel.cu:392-404), stale values or not
    lone = epi[1][:5]
    r = [np.full(5, 7.0, W.res_dtype(2))]
    assert W.HydrDispatchKernel(0, 1, [lone], [5], [epj[1][:0]], [0]) == 0
    assert W.HydrRetrieveKernel(0, 1, [5], r) == 0
    fp = W.reference_params(2)
    assert np.all(r[0]["acc"] == 0.0) and np.all(r[0]["div_v"] == 0.0)
    assert np.allclose(r[0]["dt"], fp.c_cfl * lone["smth"] / lone["snds"], rtol=1e-6)
    W.finalize()


@pytest.mark.parametrize("dim", [2, 3])
def test_cpp_bridge_in_fdps_order(O, W, dim, tmp_path):
    # the C++ header with the reference's exact signatures, driven in FDPS order
    s, p = developed_state(O, W, dim, 0.05 if dim == 2 else 0.06, 10,
                           **({} if dim == 2 else dict(tank_l=2.0, col_h=1.0, tank_h=1.5)))
    epi_all, epj_all = records(O, s, p.nu)
    epi, epj = walks(O, epi_all, epj_all, 128, np.random.default_rng(3))
    q = W.dam_break_params(dim)
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        f.write(np.int32(len(epi)).tobytes())
        f.write(np.array([len(e) for e in epi], np.int32).tobytes())
        f.write(np.array([len(e) for e in epj], np.int32).tobytes())
        f.write(np.concatenate(epi).tobytes())
        f.write(np.concatenate(epj).tobytes())
        f.write(bytes(q))
    exe = os.path.join(ROOT, "tests", "cpp", f"bridge_driver_{dim}d")
    subprocess.run([exe, str(fin), str(tmp_path / "out.bin"), "5"], check=True, timeout=60)
    got = np.fromfile(tmp_path / "out.bin", W.res_dtype(dim))
    fp = O.force_params(dim, alpha=q.alpha, beta=q.beta, eps=q.eps, c_cfl=q.c_cfl, vsig_mode=q.vsig_mode)
    ref = oracle_walks(O, fp, epi, epj)
    assert rel(got["acc"], ref["acc"]) <= TOL
    assert rel(got["div_v"], ref["div_v"]) <= TOL
