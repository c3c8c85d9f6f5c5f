"""Frozen walls (wcsph_params.wall_mode = 1) on the device-resident step.

A wall particle never moves (initialKick, fullDrift's drift, predict and
finalKick return early for FREEZE, src/class.h:146-163) and is left out of
the global time step (src/integral.h:6), so the only result of the force
pass that the step uses for a wall is div v (its density update,
src/class.h:150). A wall-wall pair has dv = 0, so it adds exactly zero to
div v and to v_sig. In wall_mode 1 a wave of 64 walls takes fluid
neighbours only on the FP32 and mixed paths, and every path writes wall
accelerations as zero (the FP64 path, the accuracy reference, still
evaluates every pair).

Checked against wall_mode 0 on the same developed state: every fluid result
and every dt_i the same bits; wall accelerations zero; wall div v the same
bits on the FP32 and FP64 paths and within FP32 rounding on the mixed path (its stack
order changes when entries are left out, so the kept terms may be added in
another order); fewer pairs evaluated. Over many steps the FP32 state is the
same bits and the mixed state stays within the rounding the reordered sums
seed.
"""
import numpy as np
import pytest

from test_gpu_sim import SMALL3, developed, state_dict

pytestmark = pytest.mark.gpu


def one_pass(W, p, st, precision, wall_mode, **kw):
    p.wall_mode = wall_mode
    with W.Simulation.from_state(p, st, precision=precision, **kw) as sim:
        sim.force()
        out = sim.download()
        out["pairs"] = sim.stats()["pairs_last"]
    return out


@pytest.mark.parametrize("precision", ["fp32", "mixed", "fp64"])
@pytest.mark.parametrize("vsig", [0, 1])
@pytest.mark.parametrize("dim,dx,steps,geom", [(2, 0.02, 400, {}), (3, 0.04, 120, SMALL3)])
def test_frozen_walls_one_pass(O, W, dim, dx, steps, geom, precision, vsig):
    s, p, _ = developed(O, W, dim, dx, steps, **geom)
    p.vsig_mode = vsig
    st = state_dict(s)
    a = one_pass(W, p, st, precision, 0)
    b = one_pass(W, p, st, precision, 1)
    fluid = st["type"] == 0
    wall = ~fluid
    assert np.array_equal(a["acc"][:, fluid], b["acc"][:, fluid])
    assert np.array_equal(a["div_v"][fluid], b["div_v"][fluid])
    assert np.array_equal(a["dt"], b["dt"])
    assert not np.any(b["acc"][:, wall])
    assert np.any(a["acc"][:, wall])
    dw0, dw1 = a["div_v"][wall], b["div_v"][wall]
    assert np.count_nonzero(dw0) > 0  # walls next to moving fluid
    if precision != "mixed":
        assert np.array_equal(dw0, dw1)
    else:
        err = np.abs(dw1 - dw0).max() / np.abs(dw0).max()
        print(f"mixed wall div v, reordered sums: max |diff| / max |div v| = {err:.2e}")
        assert err <= 1e-5
    # the skipped wall-wall pairs are no longer counted (the FP64 path, the
    # accuracy reference, evaluates every pair and only zeroes the output)
    print(f"pairs evaluated: {a['pairs']} -> {b['pairs']} ({b['pairs'] / a['pairs']:.3f})")
    assert b["pairs"] < a["pairs"] if precision != "fp64" else b["pairs"] == a["pairs"]


@pytest.mark.parametrize("precision,tol", [("fp32", 0.0), ("mixed", 1e-6)])
def test_frozen_walls_steps(W, precision, tol):
    # 300 steps of the 2D dam break (5,000 fluid, 1,518 wall particles)
    p = W.dam_break_params(2)
    st = W.dam_break_state(W.dam_break_desc(2, 0.02), p)
    outs, pairs = [], []
    for mode in (0, 1):
        p.wall_mode = mode
        with W.Simulation.from_state(p, st, precision=precision) as sim:
            sim.init()
            sim.step(300)
            outs.append(sim.download())
            pairs.append(sim.stats()["pairs_total"])
    a, b = outs
    wall = st["type"] == 1
    assert np.array_equal(b["pos"][:, wall], st["pos"][:, wall])
    if tol == 0.0:
        for k in ("pos", "vel", "dens", "div_v", "dt"):
            assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a["acc"][:, ~wall], b["acc"][:, ~wall])
    else:
        # positions in m; the wall densities' last-bit differences feed
        # back into the fluid step by step
        dpos = np.abs(a["pos"] - b["pos"]).max()
        drho = np.abs(a["dens"] / b["dens"] - 1).max()
        print(f"mixed after 300 steps: max |dpos| = {dpos:.2e} m, max |drho/rho| = {drho:.2e}")
        assert dpos <= tol and drho <= 1e-5
    assert pairs[1] < pairs[0]


def test_frozen_walls_need_walls_at_rest(W):
    p = W.dam_break_params(2, wall_mode=1)
    st = W.dam_break_state(W.dam_break_desc(2, 0.05), p)
    bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in st.items()}
    w = np.nonzero(st["type"] == 1)[0][3]
    bad["vel"][0, w] = 1e-3
    with pytest.raises(W.WcsphError, match="wall particle at rest"):
        W.Simulation.from_state(p, bad)
    # a moving fluid particle is fine
    ok = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in st.items()}
    ok["vel"][0, np.nonzero(st["type"] == 0)[0][3]] = 1e-3
    W.Simulation.from_state(p, ok).close()
