"""The force pass that allocates nothing (wcsph_force_pass_noalloc, its
workspace a fixed device array of the library) against the force pass with
a library-owned workspace (wcsph_force_pass): two calls in a row, the second
with a smaller input, each bit-identical to the allocating pass on the same
device inputs (same kernels, same order of every sum) and within 1e-5 of the
FP64 oracle; and sizes beyond the fixed array are refused before any work
is queued.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL3 = dict(tank_l=2.0, col_h=1.0, tank_h=1.5)


def developed(O, W, dim, dx, steps, **geom):
    p = W.dam_break_params(dim)
    s = O.State.dam_break(dim, dx, c0=p.c0, g=p.g, **geom)
    op = O.sim_params(dim, alpha=p.alpha, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu)
    s.init(op, 8)
    for _ in range(steps):
        s.step(op, 8)
    s.force(op, 8)
    return s, p


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def call(W, devmem, fn, p, s, ins):
    dim, n = s.x.shape
    d = W._abi.PassDesc()
    d.n = n
    for k in range(dim):
        d.box_lo[k] = float(s.x[k].min()) - 0.1
        d.box_hi[k] = float(s.x[k].max()) + 0.1
    d.h_max = float(s.h.max())
    outs = [devmem.alloc(8 * dim * n), devmem.alloc(8 * n), devmem.alloc(8 * n), devmem.alloc(8)]
    W.check(fn(C.byref(p), C.byref(d), *ins, *outs, None))
    return dict(acc=devmem.download(outs[0], (dim, n), np.float64), div_v=devmem.download(outs[1], n, np.float64),
                dt=devmem.download(outs[2], n, np.float64), pairs=int(devmem.download(outs[3], 1, np.int64)[0]))


def test_noalloc_pass_equals_allocating_pass(O, W, devmem):
    lib = W._abi.lib()
    sizes = []
    for dim, dx, steps, geom in ((3, 0.04, 20, SMALL3), (2, 0.05, 10, {})):
        s, p = developed(O, W, dim, dx, steps, **geom)
        ins = [devmem.upload(a) for a in (s.x, s.v, s.dens, s.mass, s.h)]
        ref = call(W, devmem, lib.wcsph_force_pass, p, s, ins)
        got = call(W, devmem, lib.wcsph_force_pass_noalloc, p, s, ins)
        print(f"{dim}D, {s.n} particles: pairs {got['pairs']} (allocating pass {ref['pairs']}), "
              f"acc {rel(got['acc'], s.a):.2e} from the oracle", flush=True)
        for k in ("acc", "div_v", "dt"):
            assert np.array_equal(got[k], ref[k]), k
        assert got["pairs"] == ref["pairs"] > 0
        assert rel(got["acc"], s.a) <= 1e-5 and rel(got["div_v"], s.divv) <= 1e-5
        sizes.append(s.n)
    assert sizes[1] < sizes[0]


def test_noalloc_refuses_sizes_beyond_its_array(W, devmem):
    lib = W._abi.lib()
    p = W.dam_break_params(3)
    d = W._abi.PassDesc()
    d.n, d.h_max = (1 << 19) + 1, 0.01
    for k in range(3):
        d.box_lo[k], d.box_hi[k] = 0.0, 1.0
    buf = devmem.alloc(64)
    rc = lib.wcsph_force_pass_noalloc(C.byref(p), C.byref(d), *([buf] * 8), None, None)
    assert rc == W._abi.WCSPH_EOVERFLOW and "524288 particles" in W.last_error()
    d.n = 1000
    d.box_hi[0] = 200.0  # 20000 x 100 x 100 cells of 0.01
    rc = lib.wcsph_force_pass_noalloc(C.byref(p), C.byref(d), *([buf] * 8), None, None)
    assert rc == W._abi.WCSPH_EOVERFLOW and "2097152 cells" in W.last_error()
