"""GPU checks of the mixed-precision path (h-scaled FP16 pair differences
and packed FP16 geometry; FP32 density, pressure and pair coefficients;
FP32 accumulation; FP64 integration).

The mixed path is not held to the FP32 path's 1e-5 bar: its error against
the FP64 oracle is measured and bounded here. Tolerances: one force pass on
a developed dam-break state, normwise relative error <= 1e-2 on acc and
5e-3 on div_v. Measured: 2D dx = 0.01, acc 4.9e-3, div_v 1.25e-3; 3D
dx = 0.03, acc 2.2e-3, div_v 1.7e-3. A numpy emulation of the same pass
(each FP16 rounding applied alone) attributes the 2D acc error to the FP16
cell-local positions (1.0e-3), the FP16 mass (1.1e-3) and the FP16 pair
arithmetic (5.7e-3; no single rounding above 1.3e-3), amplified because
the pressure gradient at depth is a sum of O(2(H-z)/h) larger terms that
cancel. Pair count within 0.5% of the FP32 path's exact count (the FP16
predicate blurs the support edge by ~1e-3 h).
"""
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
    return s, p, op


def state_dict(s):
    return dict(pos=s.x, vel=s.v, dens=s.dens, mass=s.mass, h=s.h, type=s.type)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


SMALL3 = dict(tank_l=2.0, col_h=1.0, tank_h=1.5)


@pytest.mark.parametrize("vsig_mode", [0, 1])
@pytest.mark.parametrize("dim,dx,geom", [(2, 0.01, {}), (3, 0.03, SMALL3)])
def test_mixed_force_pass_error(O, W, dt_neighbours, dim, dx, geom, vsig_mode):
    s, p, op = developed(O, W, dim, dx, 150, **geom)
    op.vsig_mode = vsig_mode
    p.vsig_mode = vsig_mode
    s.force(op, 8)
    out = {}
    for prec in ("mixed", "fp32"):
        with W.Simulation.from_state(p, state_dict(s), precision=prec) as sim:
            sim.force()
            out[prec] = (sim.download(), sim.stats())
            grid = sim.grid()
    m, st = out["mixed"]
    e_acc, e_div = rel(m["acc"], s.a), rel(m["div_v"], s.divv)
    print(f"mixed vs FP64 oracle: acc {e_acc:.2e}, div_v {e_div:.2e}")
    assert e_acc <= 1e-2 and e_div <= 5e-3
    assert np.allclose(m["dt"], dt_neighbours(s, op, grid), rtol=5e-3)
    n32 = out["fp32"][1]["pairs_last"]
    assert abs(st["pairs_last"] - n32) <= 5e-3 * n32


def test_mixed_steps_track_fp64(O, W):
    # 2D dam break, 300 steps from rest (t ~ 0.03 s): the mixed displacement
    # stays within 1% of the FP64 path's; velocities, still small this early,
    # carry the FP16 acceleration noise (~1% of g at depth), within 5%
    p = W.dam_break_params(2)
    st = W.dam_break_state(W.dam_break_desc(2, 0.02), p)
    res = {}
    for prec in ("mixed", "fp64"):
        with W.Simulation.from_state(p, st, precision=prec) as sim:
            sim.init()
            sim.step(300)
            res[prec] = sim.download()
    disp_m = res["mixed"]["pos"] - st["pos"]
    disp_r = res["fp64"]["pos"] - st["pos"]
    assert rel(disp_m, disp_r) < 1e-2
    assert rel(res["mixed"]["vel"], res["fp64"]["vel"]) < 5e-2


@pytest.mark.parametrize("dim,n", [(2, 20001), (3, 27001)])
def test_mixed_random_cloud(O, W, dim, n):
    # uniform random positions (no lattice, no hydrostatic cancellation), an
    # odd particle count (the last pair record has an empty half), masses
    # +-20%, densities rho0 (1 + N(0, 0.01^2)), velocities N(0, 0.5^2)
    rng = np.random.default_rng(11)
    dx = 0.02
    side = dx * n ** (1.0 / dim)
    p = W.dam_break_params(dim)
    op = O.sim_params(dim, alpha=p.alpha, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu)
    s = O.State(dim, n)
    s.x[:] = rng.uniform(0.0, side, (dim, n))
    s.v[:] = rng.normal(0.0, 0.5, (dim, n))
    s.dens[:] = p.rho0 * (1.0 + 0.01 * rng.normal(size=n))
    s.mass[:] = p.rho0 * dx ** dim * (1.0 + 0.2 * rng.uniform(-1.0, 1.0, n))
    s.h[:] = 1.3 * dx
    s.eos(op)  # Tait pressure and sound speed from the densities, as the device's reorder pass does
    s.force(op, 8)
    out = {}
    for prec in ("mixed", "fp32"):
        with W.Simulation.from_state(p, state_dict(s), precision=prec) as sim:
            sim.force()
            out[prec] = (sim.download(), sim.stats())
    m, st = out["mixed"]
    e_acc, e_div = rel(m["acc"], s.a), rel(m["div_v"], s.divv)
    print(f"mixed vs FP64 oracle, random cloud: acc {e_acc:.2e}, div_v {e_div:.2e}")
    assert e_acc <= 5e-3 and e_div <= 5e-3
    assert np.allclose(m["dt"], s.dt_i, rtol=5e-3)
    n32 = out["fp32"][1]["pairs_last"]
    assert abs(st["pairs_last"] - n32) <= 5e-3 * n32
    f32 = out["fp32"][0]
    assert rel(f32["acc"], s.a) <= 1e-5 and rel(f32["div_v"], s.divv) <= 1e-5


def test_mixed_deterministic(W):
    p = W.dam_break_params(3)
    st = W.dam_break_state(W.dam_break_desc(3, 0.04, tank_l=2.0, col_h=1.0, tank_h=1.5), p)
    outs = []
    for _ in range(2):
        with W.Simulation.from_state(p, st, precision="mixed") as sim:
            sim.init()
            sim.step(20)
            outs.append(sim.download())
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_mixed_requires_uniform_h(W):
    p = W.dam_break_params(2)
    st = W.dam_break_state(W.dam_break_desc(2, 0.05), p)
    st["h"] = st["h"].copy()
    st["h"][0] *= 1.1
    with pytest.raises(W.WcsphError, match="smoothing length"):
        W.Simulation.from_state(p, st, precision="mixed")
