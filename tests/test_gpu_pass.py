"""GPU parity of the stateless force pass on device-resident particles
(wcsph_force_pass) against the FP64 oracle: acc and div v within 1e-5
normwise, dt within 1e-6, and the pair count equal to the oracle's count of
the same cell-local predicate, on a developed 2D dam break and on a 3D state
with varying smoothing lengths and masses (v_sig mode 1, beta != 0).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def developed(O, W, dim, dx, steps, **geom):
    p = W.dam_break_params(dim)
    s = O.State.dam_break(dim, dx, c0=p.c0, g=p.g, **geom)
    op = O.sim_params(dim, alpha=p.alpha, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu)
    s.init(op, 8)
    for _ in range(steps):
        s.step(op, 8)
    return s


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def pass_grid(desc, dim):
    # the grid the header documents: cells h_max wide (as FP32, times
    # 1 + 1e-5) from box_lo, enough of them to cover box_hi
    L = 2.0 * float(np.float32(desc.h_max)) * (1.0 + 1.0e-5) / 2
    lo = np.array([desc.box_lo[k] for k in range(3)])
    nc = np.array([int(np.ceil((desc.box_hi[k] - desc.box_lo[k]) / L)) if k < dim else 1 for k in range(3)], np.int32)
    return lo, L, nc


def run_pass(W, devmem, fn, p, s, margin=0.1):
    dim, n = s.x.shape
    d = W._abi.PassDesc()
    d.n = n
    for k in range(dim):
        d.box_lo[k] = float(s.x[k].min()) - margin
        d.box_hi[k] = float(s.x[k].max()) + margin
    d.h_max = float(s.h.max())
    ins = [devmem.upload(a) for a in (s.x, s.v, s.dens, s.mass, s.h)]
    outs = [devmem.alloc(8 * dim * n), devmem.alloc(8 * n), devmem.alloc(8 * n), devmem.alloc(8)]
    W.check(fn(C.byref(p), C.byref(d), *ins, *outs, None))
    acc = devmem.download(outs[0], (dim, n), np.float64)
    divv = devmem.download(outs[1], n, np.float64)
    dt = devmem.download(outs[2], n, np.float64)
    pairs = int(devmem.download(outs[3], 1, np.int64)[0])
    return dict(acc=acc, div_v=divv, dt=dt, pairs=pairs), d


CASES = [(2, 0.02, 40, {}, 0, 0.0), (3, 0.04, 30, dict(tank_l=2.0, col_h=1.0, tank_h=1.5), 1, 0.2)]


@pytest.mark.parametrize("dim,dx,steps,geom,vsig,beta", CASES)
def test_force_pass_matches_oracle(O, W, devmem, dim, dx, steps, geom, vsig, beta, dt_neighbours):
    s = developed(O, W, dim, dx, steps, **geom)
    if dim == 3:  # the general pair body: per-particle h and m
        rng = np.random.default_rng(3)
        s.h *= 1 + 0.15 * rng.uniform(-1, 1, s.n)
        s.mass *= 1 + 0.3 * rng.uniform(-1, 1, s.n)
    p = W.dam_break_params(dim, beta=beta, vsig_mode=vsig)
    op = O.sim_params(dim, alpha=p.alpha, beta=beta, vsig_mode=vsig, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu)
    s.eos(op)
    s.force(op, 8)
    out, d = run_pass(W, devmem, W._abi.lib().wcsph_force_pass, p, s)
    grid = pass_grid(d, dim)
    ref_pairs = int(O.neighbours_brute_cells(s.x, s.h.astype(np.float32), *grid).sum())
    print(f"{dim}D pass: acc {rel(out['acc'], s.a):.2e}, div_v {rel(out['div_v'], s.divv):.2e}, "
          f"pairs {out['pairs']} (oracle {ref_pairs})")
    assert rel(out["acc"], s.a) <= 1e-5 and rel(out["div_v"], s.divv) <= 1e-5
    assert np.allclose(out["dt"], dt_neighbours(s, op, grid), rtol=1e-6)
    assert out["pairs"] == ref_pairs


def test_force_pass_rejects_bad_arguments(O, W, devmem):
    s = developed(O, W, 2, 0.05, 2)
    p = W.dam_break_params(2)
    d = W._abi.PassDesc()
    d.n, d.h_max = s.n, float(s.h.max())
    d.box_lo[0], d.box_lo[1], d.box_hi[0], d.box_hi[1] = 0.0, 0.0, -1.0, 1.0  # inverted box
    bufs = [devmem.alloc(8 * 2 * s.n)] * 8
    rc = W._abi.lib().wcsph_force_pass(C.byref(p), C.byref(d), *bufs, None, None)
    assert rc == W._abi.WCSPH_EINVAL and "box" in W.last_error()
    d.box_hi[0] = 5.0
    rc = W._abi.lib().wcsph_force_pass(C.byref(p), C.byref(d), *bufs[:4], None, *bufs[5:], None, None)
    assert rc == W._abi.WCSPH_EINVAL and "null" in W.last_error()
