"""GPU checks of the Shepard density filter (src/force.h:10-30) in the
device-resident step: applied after the first force pass and on steps with
step % shepard_interval == 0 (src/main.cpp:75,105), in FP64 from the master
state on the FP64 path and in FP32 from the FP32 force records on the FP32
and mixed paths. Compared with the oracle's filtered run (tests/test_oracle.py
pins the oracle's filter)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def state_dict(s):
    return dict(pos=s.x.copy(), vel=s.v.copy(), dens=s.dens.copy(), mass=s.mass.copy(), h=s.h.copy(),
                type=s.type.copy())


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("dim,dx,geom", [(2, 0.02, {}), (3, 0.05, dict(tank_l=2.0, col_h=1.0, tank_h=1.5))])
def test_filter_after_init_matches_oracle(O, W, dim, dx, geom):
    p = W.dam_break_params(dim, shepard_interval=30)
    s = O.State.dam_break(dim, dx, c0=p.c0, g=p.g, **geom)
    op = O.sim_params(dim, alpha=p.alpha, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu, shepard_interval=30)
    st0 = state_dict(s)
    s.init(op, 8)
    # FP64: the oracle's arithmetic to rounding; FP32 (FP32 and mixed paths):
    # distance, W and the sums in FP32 on cell-local positions, the density
    # taken as rho0 / (1 + T / S) (rho - rho0 carries the information)
    tol = {"fp64": 1e-12, "fp32": 1e-8, "mixed": 1e-8}
    for prec in ("fp64", "fp32", "mixed"):
        with W.Simulation.from_state(p, st0, precision=prec) as sim:
            sim.init()
            out = sim.download()
        err = np.max(np.abs(out["dens"] - s.dens) / s.dens)
        print(f"{dim}D {prec}: max relative density error after the filter {err:.2e}")
        assert err <= tol[prec], prec
        assert not np.allclose(out["dens"], st0["dens"], rtol=1e-9)


def test_filtered_steps_match_oracle_fp64(O, W):
    # 2D dam break, 20 steps with the filter every 3 steps (and at init)
    k = 3
    p = W.dam_break_params(2, shepard_interval=k)
    s = O.State.dam_break(2, 0.02, c0=p.c0, g=p.g)
    op = O.sim_params(2, alpha=p.alpha, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu, shepard_interval=k)
    st0 = state_dict(s)
    s.init(op, 8)
    with W.Simulation.from_state(p, st0, precision="fp64") as sim:
        sim.init()
        sim.step(7)
        sim.step(13)
        out = sim.download()
    for _ in range(20):
        s.step(op, 8)
    assert np.allclose(out["dens"], s.dens, rtol=1e-11)
    assert rel(out["pos"] - st0["pos"], s.x - st0["pos"]) < 1e-9
    assert rel(out["vel"], s.v) < 1e-9


def test_filtered_steps_match_oracle_fp32(O, W):
    k = 5
    p = W.dam_break_params(2, shepard_interval=k)
    s = O.State.dam_break(2, 0.02, c0=p.c0, g=p.g)
    op = O.sim_params(2, alpha=p.alpha, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu, shepard_interval=k)
    st0 = state_dict(s)
    s.init(op, 8)
    with W.Simulation.from_state(p, st0, precision="fp32") as sim:
        sim.init()
        sim.step(30)
        out = sim.download()
    for _ in range(30):
        s.step(op, 8)
    assert rel(out["pos"] - st0["pos"], s.x - st0["pos"]) < 1e-4
    assert rel(out["dens"] - st0["dens"], s.dens - st0["dens"]) < 1e-4


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_filter_schedule_graph_equals_eager(W, precision):
    # steps replayed from the two-step graph around the filter steps give the
    # same bits as steps launched one by one
    p = W.dam_break_params(3, shepard_interval=4)
    st = W.dam_break_state(W.dam_break_desc(3, 0.05, tank_l=2.0, col_h=1.0, tank_h=1.5), p)
    outs = []
    for graphs in (True, False):
        with W.Simulation.from_state(p, st, precision=precision) as sim:
            sim.set_graphs(graphs)
            sim.init()
            sim.step(9)
            sim.step(6)
            outs.append(sim.download())
    for key in outs[0]:
        assert np.array_equal(outs[0][key], outs[1][key]), key
