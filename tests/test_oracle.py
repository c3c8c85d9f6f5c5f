"""Pins the FP64 oracle (oracle/wcsph_oracle.c) against closed forms,
invariants and brute force. The reference's own build needs FDPS, which is
absent, so these checks are what make the oracle trustworthy; each is chosen
so a plausible mistake (wrong constant, dropped term, sign or index error)
fails it. None of them retypes the oracle's formula.
"""
import numpy as np
import pytest
from scipy import integrate

RNG = np.random.default_rng(20261203)


# --------------------------------------------------------------------------
# cubic spline, src/kernel.h:27-55
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("h", [0.013, 1.0, 2.7])
def test_kernel_normalisation(O, dim, h):
    # int W dV = 1 pins C_D and the support H = 2h
    f = (lambda r: 2 * np.pi * r * O.kernel_W(dim, r, h)) if dim == 2 else \
        (lambda r: 4 * np.pi * r * r * O.kernel_W(dim, r, h))
    val, _ = integrate.quad(f, 0, 2 * h, points=[h], epsabs=1e-13, epsrel=1e-12)
    assert abs(val - 1.0) < 1e-10


@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_support_and_smoothness(O, dim):
    h = 0.7
    assert O.kernel_W(dim, 2 * h, h) == 0.0
    assert O.kernel_W(dim, 2.3 * h, h) == 0.0
    assert O.kernel_gradW_factor(dim, 2.01 * h, h) == 0.0
    # C1 at s = 1/2 (r = h): one-sided derivatives agree
    e = 1e-6 * h
    dl = (O.kernel_W(dim, h, h) - O.kernel_W(dim, h - e, h)) / e
    dr = (O.kernel_W(dim, h + e, h) - O.kernel_W(dim, h, h)) / e
    assert abs(dl - dr) < 1e-4 * abs(dl)
    # W decreasing on (0, 2h)
    rs = np.linspace(0.01, 1.99, 200) * h
    w = [O.kernel_W(dim, r, h) for r in rs]
    assert np.all(np.diff(w) < 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_gradW_is_derivative_of_W(O, dim):
    # gradW = F dr with F r = dW/dr: pins the gradient against the normalised W
    h = 0.9
    for r in np.linspace(0.05, 1.98, 40) * h:
        e = 1e-6 * h
        dWdr = (O.kernel_W(dim, r + e, h) - O.kernel_W(dim, r - e, h)) / (2 * e)
        F = O.kernel_gradW_factor(dim, r, h)
        assert abs(F * r - dWdr) < 1e-6 * max(1.0, abs(dWdr)) * O.kernel_W(dim, 0.0, h)


# --------------------------------------------------------------------------
# Tait / Murnaghan EoS, src/EoS.h:75-90
# --------------------------------------------------------------------------
def test_eos_special_values(O):
    rho0, c0, g = 1000.0, 62.6, 7.0
    B = c0 * c0 * rho0 / g
    assert O.eos_pressure(rho0, rho0, c0, g) == 0.0
    assert abs(O.eos_soundspeed(rho0, rho0, c0, g) - c0) < 1e-12 * c0
    assert O.eos_pressure(0.98 * rho0, rho0, c0, g) == 0.0  # clamped (tension removed)
    assert abs(O.eos_pressure(1.01 * rho0, rho0, c0, g) - B * (1.01 ** 7 - 1)) < 1e-9 * B
    # c^2 = dP/drho of the unclamped Tait law, checked at rho > rho0
    for rho in (1000.5, 1003.0, 1020.0):
        e = 1e-4
        dP = (O.eos_pressure(rho + e, rho0, c0, g) - O.eos_pressure(rho - e, rho0, c0, g)) / (2 * e)
        assert abs(O.eos_soundspeed(rho, rho0, c0, g) ** 2 - dP) < 1e-6 * dP
    assert O.eos_soundspeed(1e-30, rho0, c0, g) == pytest.approx(1e-8)  # floor sqrt(1e-16)


# --------------------------------------------------------------------------
# the pair functor, src/force.h:31-60
# --------------------------------------------------------------------------
def make_cloud(O, dim, n, *, spread=1.0, hvar=0.0, mvar=0.0, vscale=1.0, P=True, nu=1e-6, seed=0):
    rng = np.random.default_rng(seed)
    dx = spread / n ** (1.0 / dim)
    pos = rng.uniform(0, spread, (n, dim))
    vel = rng.normal(0, vscale, (n, dim))
    h = 1.3 * dx * (1 + hvar * rng.uniform(-1, 1, n))
    dens = 1000.0 * (1 + 0.01 * rng.normal(size=n))
    mass = dens * dx ** dim * (1 + mvar * rng.uniform(-1, 1, n))
    c0 = 30.0
    pres = np.array([O.eos_pressure(d, 1000.0, c0, 7.0) for d in dens]) if P else np.zeros(n)
    snds = np.array([O.eos_soundspeed(d, 1000.0, c0, 7.0) for d in dens])
    epi = np.zeros(n, O.epi_dtype(dim))
    epj = np.zeros(n, O.epj_dtype(dim))
    for a in (epi, epj):
        a["pos"], a["vel"], a["smth"], a["dens"], a["pres"], a["snds"], a["visc"] = pos, vel, h, dens, pres, snds, nu
        a["tag"] = np.arange(n)
    epj["mass"] = mass
    return epi, epj


@pytest.mark.parametrize("dim", [2, 3])
def test_two_particles_pressure(O, dim):
    # two particles at rest: acc_i = -m_j (P_i/rho_i^2 + P_j/rho_j^2) dW/dr(hbar) e_ij
    epi, epj = make_cloud(O, dim, 2, seed=1, P=True, nu=0.0)
    epi["vel"] = 0
    epj["vel"] = 0
    epi["pos"][0] = 0
    epi["pos"][1] = 0
    epi["pos"][1][0] = 0.7
    epj["pos"] = epi["pos"]
    for a in (epi, epj):
        a["smth"] = [0.5, 0.6]
        a["dens"] = [1010.0, 1004.0]
        a["pres"] = [3.0e4, 1.5e4]
    res = O.hydro_force(O.force_params(dim), epi, epj)
    hbar, r, e = 0.55, 0.7, 1e-7
    dWdr = (O.kernel_W(dim, r + e, hbar) - O.kernel_W(dim, r - e, hbar)) / (2 * e)
    coef = 3.0e4 / 1010.0 ** 2 + 1.5e4 / 1004.0 ** 2
    # i=0 sits at x=0, j=1 at x=0.7: dr = -0.7 e_x, so acc_0 = +m_1 coef dWdr (dWdr < 0): pushed to -x
    expect0 = -epj["mass"][1] * coef * dWdr * (-1.0)
    assert res["acc"][0][0] == pytest.approx(expect0, rel=1e-6)
    assert res["acc"][0][0] < 0  # repulsion
    assert np.all(res["acc"][0][1:] == 0)
    assert res["div_v"][0] == 0.0
    # dt quirk: v_sig stays 0, so dt = C_CFL h / c (src/force.h:47,56)
    assert res["dt"][0] == pytest.approx(0.3 * 0.5 / epi["snds"][0], rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_two_particles_artificial_viscosity(O, dim):
    # approaching pair with P = 0, nu = 0: acc_i = -m_j Pi dW/dr e_ij with the
    # Monaghan Pi = -alpha cbar mu / rhobar, mu = hbar v.r / (r^2 + eps h_i h_j)
    epi, epj = make_cloud(O, dim, 2, seed=2, P=False, nu=0.0)
    for a in (epi, epj):
        a["pos"] = 0
        a["pos"][1][0] = 0.5
        a["vel"] = 0
        a["vel"][0][0] = 2.0   # i moves towards j
        a["smth"] = [0.4, 0.4]
        a["dens"] = [1000.0, 1000.0]
        a["snds"] = [30.0, 30.0]
    alpha, eps = 0.25, 0.01
    res = O.hydro_force(O.force_params(dim, alpha=alpha), epi, epj)
    r, h, e = 0.5, 0.4, 1e-7
    vr = (2.0 - 0.0) * (0.0 - 0.5)  # (v_i - v_j).(x_i - x_j) < 0: approaching
    mu = h * vr / (r * r + eps * h * h)
    Pi = -alpha * 30.0 * mu / 1000.0
    dWdr = (O.kernel_W(dim, r + e, h) - O.kernel_W(dim, r - e, h)) / (2 * e)
    expect = -epj["mass"][1] * Pi * dWdr * (-1.0)
    assert res["acc"][0][0] == pytest.approx(expect, rel=1e-6)
    assert res["acc"][0][0] < 0  # decelerates the approaching particle
    # receding pair: no artificial viscosity
    for a in (epi, epj):
        a["vel"][0][0] = -2.0
    res = O.hydro_force(O.force_params(dim, alpha=alpha), epi, epj)
    assert res["acc"][0][0] == 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_momentum_antisymmetry(O, dim):
    # every pair term is antisymmetric in (i,j) (symmetric h, m-weighting),
    # so sum_i m_i a_i = 0 when each i sees all particles as j
    epi, epj = make_cloud(O, dim, 300, hvar=0.2, mvar=0.3, seed=3)
    res = O.hydro_force(O.force_params(dim, alpha=0.3, beta=0.5), epi, epj)
    ma = (epj["mass"][:, None] * res["acc"]).sum(0)
    scale = (epj["mass"][:, None] * np.abs(res["acc"])).sum(0)
    assert np.all(np.abs(ma) < 1e-12 * scale.max())


@pytest.mark.parametrize("dim", [2, 3])
def test_grad_mode_pair_is_mean_of_single_h_pairs(O, dim):
    # grad_mode 1 (src/kernel.cu:265): the pressure force of a pair at rest
    # with h_i != h_j is the mean of the forces of the same pair with both
    # smoothing lengths h_i and with both h_j (mode 0, pinned above by finite
    # differences); the second pair sits beyond h_i + h_j but within
    # 2 max(h_i, h_j), where mode 0 gives nothing and mode 1 half of one term
    hs = (0.5, 0.6)
    for r, beyond_mean in ((0.7, False), (1.15, True)):
        epi, epj = make_cloud(O, dim, 2, seed=1, P=True, nu=0.0)
        for a in (epi, epj):
            a["vel"] = 0
            a["pos"] = 0
            a["pos"][1][0] = r
            a["dens"] = [1010.0, 1004.0]
            a["pres"] = [3.0e4, 1.5e4]

        def acc(hpair, mode):
            for a in (epi, epj):
                a["smth"] = hpair
            return O.hydro_force(O.force_params(dim, grad_mode=mode), epi, epj)["acc"][0][0]

        a1 = acc(hs, 1)
        expect = 0.5 * (acc((hs[0], hs[0]), 0) + acc((hs[1], hs[1]), 0))
        assert a1 == pytest.approx(expect, rel=1e-13)
        assert (acc(hs, 0) == 0.0) == beyond_mean
        assert a1 < 0  # repulsion either way


@pytest.mark.parametrize("dim", [2, 3])
def test_grad_mode_uniform_h_and_antisymmetry(O, dim):
    # with one smoothing length the two gradient forms are the same numbers;
    # with varying h the mean of the gradients is symmetric in (i, j), so
    # sum_i m_i a_i = 0 (taking only the gradient at h_i breaks it)
    epi, epj = make_cloud(O, dim, 200, hvar=0.0, mvar=0.3, seed=5)
    r0 = O.hydro_force(O.force_params(dim, alpha=0.3), epi, epj)
    r1 = O.hydro_force(O.force_params(dim, alpha=0.3, grad_mode=1), epi, epj)
    assert np.array_equal(r0, r1)
    epi, epj = make_cloud(O, dim, 300, hvar=0.3, mvar=0.3, seed=6)
    res = O.hydro_force(O.force_params(dim, alpha=0.3, beta=0.5, grad_mode=1), epi, epj)
    ma = (epj["mass"][:, None] * res["acc"]).sum(0)
    scale = (epj["mass"][:, None] * np.abs(res["acc"])).sum(0)
    assert np.all(np.abs(ma) < 1e-12 * scale.max())
    res0 = O.hydro_force(O.force_params(dim, alpha=0.3, beta=0.5), epi, epj)
    assert not np.allclose(res["acc"], res0["acc"], rtol=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
def test_dissipation_sign(O, dim):
    # with P = 0 only artificial and laminar viscosity act: they must remove
    # kinetic energy, sum_i m_i v_i.a_i <= 0
    epi, epj = make_cloud(O, dim, 300, P=False, nu=1e-3, seed=4, hvar=0.1)
    res = O.hydro_force(O.force_params(dim, alpha=0.5, beta=1.0), epi, epj)
    power = (epj["mass"][:, None] * epi["vel"] * res["acc"]).sum()
    assert power < 0
    # the laminar term alone (alpha = 0) is dissipative too
    res = O.hydro_force(O.force_params(dim, alpha=0.0), epi, epj)
    assert (epj["mass"][:, None] * epi["vel"] * res["acc"]).sum() < 0


@pytest.mark.parametrize("dim", [2, 3])
def test_galilean_and_order_invariance(O, dim):
    epi, epj = make_cloud(O, dim, 200, seed=5, hvar=0.2)
    p = O.force_params(dim, alpha=0.2, beta=0.3, vsig_mode=1)
    r0 = O.hydro_force(p, epi, epj)
    # uniform velocity shift
    e2, j2 = epi.copy(), epj.copy()
    e2["vel"] += 3.7
    j2["vel"] += 3.7
    r1 = O.hydro_force(p, e2, j2)
    for k in ("acc", "div_v", "dt"):
        assert np.allclose(r1[k], r0[k], rtol=1e-9, atol=1e-9 * np.abs(r0[k]).max())
    # permuted j list plus far-away extra particles (FDPS superset lists)
    perm = RNG.permutation(len(epj))
    far = epj[:20].copy()
    far["pos"] += 100.0
    j3 = np.concatenate([epj[perm], far])
    r2 = O.hydro_force(p, epi, j3)
    for k in ("acc", "div_v"):
        assert np.allclose(r2[k], r0[k], rtol=1e-11, atol=1e-11 * np.abs(r0[k]).max())


def test_rotation_invariance_3d(O):
    epi, epj = make_cloud(O, 3, 200, seed=6, hvar=0.2)
    p = O.force_params(3, alpha=0.2)
    r0 = O.hydro_force(p, epi, epj)
    q, _ = np.linalg.qr(RNG.normal(size=(3, 3)))
    e2, j2 = epi.copy(), epj.copy()
    for a in (e2, j2):
        a["pos"] = a["pos"] @ q.T
        a["vel"] = a["vel"] @ q.T
    r1 = O.hydro_force(p, e2, j2)
    assert np.allclose(r1["acc"], r0["acc"] @ q.T, atol=1e-9 * np.abs(r0["acc"]).max())
    assert np.allclose(r1["div_v"], r0["div_v"], atol=1e-9 * np.abs(r0["div_v"]).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_lattice_interior_gradients(O, dim):
    # on a lattice interior, the functor's div_v of v = A x tends to trace(A)
    # and its pressure term of a linear P tends to -grad P / rho (SPH
    # consistency at h = 1.3 dx, a few percent): pins signs and scales.
    dx = 0.02
    m = 13 if dim == 3 else 31
    g = np.stack(np.meshgrid(*[np.arange(m)] * dim, indexing="ij"), -1).reshape(-1, dim) * dx
    n = len(g)
    c = np.argmin(((g - g.mean(0)) ** 2).sum(1))
    A = RNG.normal(size=(dim, dim))
    epj = np.zeros(n, O.epj_dtype(dim))
    epj["pos"] = g
    epj["vel"] = g @ A.T
    epj["smth"], epj["dens"], epj["snds"], epj["visc"] = 1.3 * dx, 1000.0, 30.0, 0.0
    epj["mass"] = 1000.0 * dx ** dim
    gradP = np.array([3.0e4, -2.0e4, 1.0e4][:dim])
    epj["pres"] = 5.0e4 + g @ gradP
    epi = np.zeros(1, O.epi_dtype(dim))
    for k in ("pos", "vel", "smth", "dens", "pres", "snds", "visc"):
        epi[k] = epj[k][c]
    res = O.hydro_force(O.force_params(dim, alpha=0.0), epi, epj)
    assert res["div_v"][0] == pytest.approx(np.trace(A), rel=0.05, abs=0.05 * np.abs(A).max())
    # pressure part only: rerun with v = 0
    epj["vel"] = 0
    epi["vel"] = 0
    res = O.hydro_force(O.force_params(dim, alpha=0.0), epi, epj)
    assert np.allclose(res["acc"][0], -gradP / 1000.0, rtol=0.05, atol=0.05 * np.abs(gradP).max() / 1000)


def shear_lattice(O, dim, field, *, nu_i, nu_j, rho_i=1000.0, rho_j=1000.0, dx=0.02):
    """Lattice block centred on its middle particle, P = 0 everywhere, with
    velocity field(g) (g: (n, dim) positions relative to the centre); returns
    (i record of the centre, j records of the whole block)."""
    m = 13 if dim == 3 else 31
    g = np.stack(np.meshgrid(*[np.arange(m)] * dim, indexing="ij"), -1).reshape(-1, dim) * dx
    c = int(np.argmin(((g - g.mean(0)) ** 2).sum(1)))
    g = g - g[c]
    epj = np.zeros(len(g), O.epj_dtype(dim))
    epj["pos"], epj["vel"] = g, field(g)
    epj["smth"], epj["dens"], epj["snds"], epj["visc"] = 1.3 * dx, rho_j, 30.0, nu_j
    epj["mass"] = rho_j * dx ** dim  # particle volume dx^D
    epi = np.zeros(1, O.epi_dtype(dim))
    for k in ("pos", "vel", "smth", "snds"):
        epi[k] = epj[k][c]
    epi["dens"], epi["visc"] = rho_i, nu_i
    return epi, epj


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("nu_i,nu_j,rho_j", [(1e-2, 1e-2, 1000.0), (1e-2, 3e-2, 1000.0), (1e-2, 1e-2, 3000.0)])
def test_laminar_viscosity_shear_closed_form(O, dim, nu_i, nu_j, rho_j):
    # Morris laminar term (src/force.h:53) in closed form. The SPH Laplacian
    #   lap v_i ~ 2 sum_j V_j (v_i - v_j) (r_ij . gradW_ij) / r_ij^2
    # (V_j = m_j / rho_j) is exact for quadratic fields on a symmetric
    # lattice up to the kernel's second moment (1.5% at h = 1.3 dx in 3D)
    # and the 0.01 h^2 regulariser. The functor's coefficient
    # m_j 2 (nu_i + nu_j) / (rho_i + rho_j) then gives, for v_x = A y^2
    # (lap v_x = 2 A) with uniform nu_j and rho_j around i:
    #   a_x = (nu_i + nu_j)/2 * 2 rho_j / (rho_i + rho_j) * 2 A
    # which is nu lap v for one fluid. A coefficient halved or doubled, nu_i
    # in place of the mean, or 2 rho_i in place of rho_i + rho_j fails.
    A = 2.5
    def field(g):
        v = np.zeros_like(g)
        v[:, 0] = A * g[:, 1] ** 2
        return v
    epi, epj = shear_lattice(O, dim, field, nu_i=nu_i, nu_j=nu_j, rho_j=rho_j)
    res = O.hydro_force(O.force_params(dim, alpha=0.0), epi, epj)
    expect = 0.5 * (nu_i + nu_j) * 2.0 * rho_j / (1000.0 + rho_j) * 2.0 * A
    assert res["acc"][0][0] == pytest.approx(expect, rel=0.03)
    assert np.all(np.abs(res["acc"][0][1:]) < 1e-9 * expect)


@pytest.mark.parametrize("dim", [2, 3])
def test_laminar_viscosity_linear_field_is_zero(O, dim):
    # a linear velocity field has no Laplacian: on the symmetric lattice the
    # laminar term cancels pair by pair (j and its mirror image), so the
    # acceleration is zero to rounding while each pair term is not
    G = np.random.default_rng(12).normal(size=(dim, dim))
    epi, epj = shear_lattice(O, dim, lambda g: g @ G.T, nu_i=1e-2, nu_j=1e-2)
    res = O.hydro_force(O.force_params(dim, alpha=0.0), epi, epj)
    one = O.hydro_force(O.force_params(dim, alpha=0.0), epi, epj[np.argsort(
        ((epj["pos"] - np.array([0.02] + [0.0] * (dim - 1))) ** 2).sum(1))[:1]])
    assert np.abs(one["acc"][0]).max() > 1e-3  # one neighbour alone does act
    assert np.abs(res["acc"][0]).max() < 1e-12 * np.abs(one["acc"][0]).max() * len(epj)


def test_internal_energy_of_tait_law(O):
    # the internal energy summed by the conservation diagnostics is the
    # work of compression of the clamped Tait law: e(rho0) = 0,
    # de/drho = P/rho^2 (P from the separately pinned EoS), and e = 0 below
    # rho0 where P is clamped to 0. A wrong constant in e fails e(rho0) = 0
    # or the derivative.
    p = O.sim_params(3)
    s = O.State(3, 1)
    s.mass[0], s.h[0] = 2.0, 0.01

    def e(rho):
        s.dens[0] = rho
        return s.conservation(p)[6] / s.mass[0]

    assert e(p.rho0) == 0.0
    assert e(0.97 * p.rho0) == 0.0 and e(0.5 * p.rho0) == 0.0
    for rho in (1000.2, 1004.0, 1030.0, 1150.0):
        d = 1e-3 * (rho - p.rho0)
        dedrho = (e(rho + d) - e(rho - d)) / (2 * d)
        P = O.eos_pressure(rho, p.rho0, p.c0, p.gamma)
        assert dedrho == pytest.approx(P / rho ** 2, rel=1e-6)
    # and e integrates P/rho^2 from rho0 (trapezoid on a fine grid)
    rr = np.linspace(p.rho0, 1050.0, 20001)
    f = np.array([O.eos_pressure(r, p.rho0, p.c0, p.gamma) for r in rr]) / rr ** 2
    assert e(1050.0) == pytest.approx(np.sum(0.5 * (f[1:] + f[:-1]) * np.diff(rr)), rel=1e-7)


def _pair_state(O, dim, xs, dens, vel=None, h=0.013):
    p = O.sim_params(dim, nu=1e-3)
    s = O.State(dim, len(xs))
    for i, x in enumerate(xs):
        s.x[:, i] = x
    s.dens[:] = dens
    s.mass[:] = 1000.0 * 0.01 ** dim
    s.h[:] = h
    if vel is not None:
        s.v[:] = np.asarray(vel, float).T
    s.eos(p)
    return p, s


@pytest.mark.parametrize("dim", [2, 3])
def test_force_scales_single_pair_is_the_magnitude(O, dim):
    # with one neighbour each sum has one term: its scale is the magnitude
    # of the force and of div v themselves (a scale that dropped a term,
    # missed the absolute value or the 1/rho_i would differ)
    x0 = np.zeros(dim)
    x1 = np.array([0.011, 0.007, -0.004][:dim])
    p, s = _pair_state(O, dim, [x0, x1], [1012.0, 1003.0], vel=[[0.3, -0.1, 0.2][:dim], [-0.2, 0.4, 0.1][:dim]])
    s.force(p)
    sa, sd = s.force_scales(p)
    assert np.abs(s.a).max() > 0 and np.abs(s.divv).max() > 0
    np.testing.assert_allclose(sa, np.linalg.norm(s.a, axis=0), rtol=1e-13)
    np.testing.assert_allclose(sd, np.abs(s.divv), rtol=1e-13)
    # the state is left as the force pass wrote it
    a0 = s.a.copy()
    s.force_scales(p)
    assert np.array_equal(s.a, a0)


@pytest.mark.parametrize("dim", [2, 3])
def test_force_scales_count_cancelled_terms(O, dim):
    # three particles on a line, equal densities, the ends beyond each
    # other's support: the middle's two pair terms cancel exactly (a = 0),
    # while its scale is the sum of their magnitudes, twice an end's
    d = 0.02
    xs = [np.eye(dim)[0] * k * d for k in (-1, 0, 1)]
    p, s = _pair_state(O, dim, xs, [1010.0] * 3)
    s.force(p)
    sa, sd = s.force_scales(p)
    end = np.linalg.norm(s.a[:, 0])
    assert end > 0
    assert np.linalg.norm(s.a[:, 1]) < 1e-12 * end
    assert sa[1] == pytest.approx(2 * end, rel=1e-12)
    assert sa[0] == pytest.approx(end, rel=1e-12) and sa[2] == pytest.approx(end, rel=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_force_scales_bound_the_sums(O, dim):
    # on a developed dam break every |a_i| and |div v_i| is at most its
    # scale (triangle inequality), and at depth far below it
    pp = O.sim_params(dim, alpha=0.1, nu=1e-6)
    s = O.State.dam_break(dim, 0.05 if dim == 2 else 0.1)
    s.init(pp)
    for _ in range(20):
        s.step(pp)
    s.force(pp)
    sa, sd = s.force_scales(pp)
    an = np.linalg.norm(s.a, axis=0)
    assert np.all(an <= sa * (1 + 1e-12))
    assert np.all(np.abs(s.divv) <= sd * (1 + 1e-12) + 1e-300)
    fl = s.type == 0
    assert np.median(an[fl] / sa[fl]) < 0.5


# --------------------------------------------------------------------------
# neighbour sets, FDPS-style walks
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_neighbour_brute_vs_distance(O, dim):
    n = 800
    pos = RNG.uniform(0, 1, (dim, n))
    h = (0.05 * (1 + 0.3 * RNG.uniform(-1, 1, n))).astype(np.float32)
    counts, lists = O.neighbours_brute(pos, h, cap=256)
    d = np.sqrt(((pos[:, :, None] - pos[:, None, :]) ** 2).sum(0))
    H = h[:, None].astype(np.float64) + h[None, :]
    exact = (d < H) & (d > 0)
    # the FP32 predicate may only differ from the FP64 distance within rounding of the boundary
    border = np.abs(d - H) < 1e-5 * H
    got = np.zeros_like(exact)
    for i in range(n):
        got[i, lists[i, :counts[i]]] = True
    assert np.all((got == exact) | border)
    assert np.array_equal(got, got.T)  # the predicate is symmetric


@pytest.mark.parametrize("dim", [2, 3])
def test_walks_are_supersets(O, dim):
    n = 3000
    pos = RNG.uniform(0, 1, (dim, n))
    h = np.full(n, 0.02, np.float32)
    go, gi, no, ni = O.make_walks(pos, r_search=2 * 0.02, group_max=128)
    assert np.array_equal(np.sort(gi), np.arange(n))  # the groups partition the particles
    assert np.all(np.diff(go) <= 128)
    counts, lists = O.neighbours_brute(pos, h, cap=128)
    for w in range(len(go) - 1):
        js = set(ni[no[w]:no[w + 1]].tolist())
        for i in gi[go[w]:go[w + 1]]:
            assert set(lists[i, :counts[i]].tolist()) <= js
            assert i in js


# --------------------------------------------------------------------------
# initial condition, integrator, time step
# --------------------------------------------------------------------------
def test_dam_break_counts(O):
    s2 = O.State.dam_break(2, 0.02)
    assert (s2.type == 0).sum() == 5000 and (s2.type == 1).sum() == 1518
    s3 = O.State.dam_break(3, 0.0126)
    assert (s3.type == 0).sum() == 502440 and (s3.type == 1).sum() == 562938


def test_free_fall_is_exact(O):
    # one isolated fluid particle: no hydro force, dt = min(C h / c, 0.25 sqrt(h/g)),
    # and KDK integrates constant gravity exactly: z = z0 - g t^2 / 2, v = -g t
    s = O.State(3, 1)
    s.x[:, 0] = [0.3, 0.2, 5.0]
    s.dens[0], s.mass[0], s.h[0] = 1000.0, 1e-3, 0.01
    p = O.sim_params(3)
    s.init(p)
    dt_expect = min(0.25 * 0.01 / p.c0, 0.25 * np.sqrt(0.01 / p.g))
    assert s.global_dt(p) == pytest.approx(dt_expect, rel=1e-14)
    for _ in range(50):
        s.step(p)
    t = s.time
    assert t == pytest.approx(50 * dt_expect, rel=1e-12)
    assert s.x[2, 0] == pytest.approx(5.0 - 0.5 * p.g * t * t, rel=1e-12)
    assert s.v[2, 0] == pytest.approx(-p.g * t, rel=1e-12)
    assert s.x[0, 0] == 0.3 and s.v[0, 0] == 0.0


def test_viscous_time_step_bound(O):
    # getTimeStepGlobal (src/integral.h:2-12) with a viscosity large enough
    # that the explicit-diffusion bound dt <= 0.125 h^2 / nu (src/integral.h:8)
    # is the binding one: one isolated particle, h = 0.01 m, nu = 1 m^2/s gives
    # 1.25e-5 s against C h / c = 4.0e-5 s and 0.25 sqrt(h / g) = 7.98e-3 s
    s = O.State(3, 1)
    s.x[:, 0] = [0.3, 0.2, 5.0]
    s.dens[0], s.mass[0], s.h[0] = 1000.0, 1e-3, 0.01
    p = O.sim_params(3, nu=1.0)
    s.init(p)
    assert s.global_dt(p) == pytest.approx(1.25e-5, rel=1e-14)
    # below that viscosity the CFL bound of the particle takes over again
    q = O.sim_params(3, nu=0.1)
    s.init(q)
    assert s.global_dt(q) == pytest.approx(0.25 * 0.01 / q.c0, rel=1e-14)
    # walls never bound the step
    s.type[0] = 1
    assert s.global_dt(p) == 1.0e30


def test_walls_do_not_move(O):
    s = O.State.dam_break(2, 0.05)
    p = O.sim_params(2)
    s.init(p)
    x0 = s.x.copy()
    for _ in range(20):
        s.step(p)
    w = s.type == 1
    assert np.array_equal(s.x[:, w], x0[:, w])
    assert np.all(s.v[:, w] == 0)
    assert np.any(s.x[:, ~w] != x0[:, ~w])


def test_hydrostatic_start_is_near_balance(O):
    # the hydrostatic initial density makes the pressure force hold the column
    # against gravity: the interior of the 2D column barely accelerates
    s = O.State.dam_break(2, 0.02)
    p = O.sim_params(2)
    s.init(p)
    x, y = s.x
    inner = (s.type == 0) & (x > 0.2) & (x < 0.8) & (y > 0.2) & (y < 1.6)
    a_net = s.a[:, inner]  # includes gravity
    assert np.abs(a_net[1]).mean() < 0.1 * p.g


def test_cell_force_equals_all_pairs(O):
    # the whole-step oracle's cell-list force equals the plain functor over all particles
    s = O.State.dam_break(2, 0.05)
    p = O.sim_params(2)
    s.init(p)
    for _ in range(5):
        s.step(p)
    n = s.n
    epi = np.zeros(n, O.epi_dtype(2))
    epj = np.zeros(n, O.epj_dtype(2))
    for a in (epi, epj):
        a["pos"], a["vel"], a["smth"], a["dens"], a["pres"], a["snds"], a["visc"] = \
            s.x.T, s.v.T, s.h, s.dens, s.pres, s.snds, p.nu
    epj["mass"] = s.mass
    fp = O.force_params(2, alpha=p.alpha, c_cfl=p.c_cfl)
    ref = O.hydro_force(fp, epi, epj)
    s.force(p)  # hydro only
    assert np.allclose(s.a.T, ref["acc"], rtol=1e-10, atol=1e-10 * np.abs(ref["acc"]).max())
    assert np.allclose(s.divv, ref["div_v"], rtol=1e-10, atol=1e-10 * np.abs(ref["div_v"]).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_signal_velocity_modes(O, dim):
    # approaching pair: mu < 0. Mode 0 (src/force.h:47) keeps v_sig = max(0, mu) = 0;
    # mode 1 (src/kernel.cu:263) uses -mu, so dt = C h / (c + |mu|)
    epi, epj = make_cloud(O, dim, 2, seed=7, P=False, nu=0.0)
    for a in (epi, epj):
        a["pos"] = 0
        a["pos"][1][0] = 0.5
        a["vel"] = 0
        a["vel"][0][0] = 2.0
        a["smth"], a["dens"], a["snds"] = 0.4, 1000.0, 30.0
    mu = 0.4 * (2.0 * -0.5) / (0.25 + 0.01 * 0.16)
    r0 = O.hydro_force(O.force_params(dim, vsig_mode=0), epi, epj)
    r1 = O.hydro_force(O.force_params(dim, vsig_mode=1), epi, epj)
    assert r0["dt"][0] == pytest.approx(0.3 * 0.4 / 30.0, rel=1e-14)
    assert r1["dt"][0] == pytest.approx(0.3 * 0.4 / (30.0 - mu), rel=1e-12)


def test_step_composition(O):
    # one step of src/main.cpp:89-133 composed by hand from the (separately
    # pinned) functor: kick-drift with the old acceleration, force on the
    # predicted velocity v + dt a, gravity, final kick v = v_half + dt/2 a
    s = O.State.dam_break(2, 0.1, col_w=0.5, col_h=0.6, tank_l=1.0, tank_h=1.0)
    p = O.sim_params(2, alpha=0.3)
    s.init(p)
    for _ in range(3):
        s.step(p)
    x0, v0, a0, rho0, dv0 = s.x.copy(), s.v.copy(), s.a.copy(), s.dens.copy(), s.divv.copy()
    dt = s.global_dt(p)
    s.step(p)
    fl = s.type == 0
    vh = np.where(fl, v0 + 0.5 * dt * a0, 0.0)
    x1 = np.where(fl, x0 + dt * vh, x0)
    vp = np.where(fl, v0 + dt * a0, v0)
    rho1 = rho0 - rho0 * dt * dv0
    n = s.n
    epi = np.zeros(n, O.epi_dtype(2))
    epj = np.zeros(n, O.epj_dtype(2))
    P = np.array([O.eos_pressure(r, p.rho0, p.c0, p.gamma) for r in rho1])
    c = np.array([O.eos_soundspeed(r, p.rho0, p.c0, p.gamma) for r in rho1])
    for a in (epi, epj):
        a["pos"], a["vel"], a["smth"], a["dens"], a["pres"], a["snds"], a["visc"] = x1.T, vp.T, s.h, rho1, P, c, p.nu
    epj["mass"] = s.mass
    ref = O.hydro_force(O.force_params(2, alpha=p.alpha, c_cfl=p.c_cfl), epi, epj)
    acc = ref["acc"].T.copy()
    acc[1] -= p.g
    v1 = np.where(fl, vh + 0.5 * dt * acc, vp)
    assert np.allclose(s.x, x1, rtol=0, atol=1e-14)
    assert np.allclose(s.a, acc, rtol=1e-10, atol=1e-10 * np.abs(acc).max())
    assert np.allclose(s.v, v1, rtol=1e-10, atol=1e-12)
    assert np.allclose(s.dens, rho1, rtol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_cell_local_predicate(O, dim):
    # the device path's predicate on cell-local FP32 coordinates agrees with
    # the FP64 distance up to rounding at the support boundary, is symmetric,
    # and does not depend on where the grid origin sits
    n = 700
    pos = RNG.uniform(0, 1, (dim, n))
    h = (0.05 * (1 + 0.3 * RNG.uniform(-1, 1, n))).astype(np.float32)
    L = 2 * float(h.max()) * (1 + 1e-5) / 2
    lo = np.full(3, -0.1)
    nc = np.array([int(np.ceil(1.2 / L))] * 3, np.int32)
    cnt, lists = O.neighbours_brute_cells(pos, h, lo, L, nc, cap=256)
    d = np.sqrt(((pos[:, :, None] - pos[:, None, :]) ** 2).sum(0))
    H = h[:, None].astype(np.float64) + h[None, :]
    exact = (d < H) & (d > 0)
    border = np.abs(d - H) < 1e-5 * H
    got = np.zeros_like(exact)
    for i in range(n):
        got[i, lists[i, :cnt[i]]] = True
    assert np.all((got == exact) | border)
    assert np.array_equal(got, got.T)
    cnt2 = O.neighbours_brute_cells(pos, h, lo - 0.37 * L, L, nc + 1)
    assert np.abs(cnt2 - cnt).sum() <= max(2, int(border.sum()))


# ---------------------------------------------------------------------------
# Shepard filter (src/force.h:10-30)
# ---------------------------------------------------------------------------
def test_shepard_two_particles_closed_form(O):
    # rho_i' = (m_i W(0,h_i) + m_j W(r,h_i)) / (m_i/rho_i W(0,h_i) + m_j/rho_j W(r,h_i))
    for dim in (2, 3):
        r, hi, hj = 0.37, 0.3, 0.45
        pos = np.zeros((dim, 2)); pos[0, 1] = r
        h = np.array([hi, hj]); m = np.array([1.3, 0.7]); rho = np.array([1010.0, 990.0])
        out = O.shepard(pos, h, m, rho)
        for i, j in ((0, 1), (1, 0)):
            w0, wr = O.kernel_W(dim, 0.0, h[i]), O.kernel_W(dim, r, h[i])
            want = (m[i] * w0 + m[j] * wr) / (m[i] / rho[i] * w0 + m[j] / rho[j] * wr)
            assert out[i] == pytest.approx(want, rel=1e-14)
        # the kernel is evaluated at h_i, not the mean: j beyond 2 h_i is ignored by i
        pos[0, 1] = 2.0 * hi * 1.01
        out = O.shepard(pos, h, m, rho)
        assert out[0] == pytest.approx(rho[0], rel=1e-14)
        assert out[1] != pytest.approx(rho[1], rel=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
def test_shepard_linear_field_on_lattice(O, dim):
    # on a uniform lattice with volumes V = dx^D and m_j = rho_j V, the filter is
    # the W-weighted mean of rho_j; W is symmetric on the lattice, so a linear
    # field comes back exactly in the interior (a wrong volume weighting, e.g.
    # m_j instead of m_j / rho_j in the denominator, breaks this)
    dx, h = 0.1, 0.13
    g = np.arange(10 if dim == 3 else 16) * dx
    grids = np.meshgrid(*([g] * dim), indexing="ij")
    pos = np.vstack([a.ravel() for a in grids])
    n = pos.shape[1]
    rho = 1000.0 + 3.0 * pos[0] - 2.0 * pos[1]
    m = rho * dx ** dim
    out = O.shepard(pos, np.full(n, h), m, rho, 4)
    lo, hi = g[0] + 2 * h, g[-1] - 2 * h
    interior = np.all((pos > lo) & (pos < hi), axis=0)
    assert interior.sum() > 0
    assert np.allclose(out[interior], rho[interior], rtol=1e-13)


def test_shepard_uniform_density_is_fixed_point(O):
    # with one density everywhere the filter returns it exactly, for any positions
    rng = np.random.default_rng(7)
    for dim in (2, 3):
        n = 400
        pos = rng.uniform(0, 1, (dim, n))
        h = rng.uniform(0.08, 0.12, n)
        m = rng.uniform(0.5, 2.0, n)
        out = O.shepard(pos, h, m, np.full(n, 1000.0))
        assert np.allclose(out, 1000.0, rtol=1e-13)


def test_shepard_smooths_noise_brute_force(O):
    # against a direct O(n^2) restatement: noisy density on a jittered lattice
    rng = np.random.default_rng(3)
    g = np.arange(12) * 0.1
    X, Y = np.meshgrid(g, g, indexing="ij")
    pos = np.vstack([X.ravel(), Y.ravel()]) + rng.uniform(-0.01, 0.01, (2, g.size ** 2))
    n = pos.shape[1]
    h = np.full(n, 0.13); m = np.full(n, 10.0)
    rho = 1000.0 * (1 + 0.01 * rng.standard_normal(n))
    out = O.shepard(pos, h, m, rho, 4)
    ref = np.empty(n)
    for i in range(n):
        r = np.linalg.norm(pos - pos[:, i:i + 1], axis=0)
        W = np.array([O.kernel_W(2, rr, h[i]) for rr in r])
        ref[i] = np.sum(m * W) / np.sum(m / rho * W)
    assert np.allclose(out, ref, rtol=1e-13)
    interior = np.all((pos > 0.3) & (pos < 0.8), axis=0)
    assert np.std(out[interior]) < 0.5 * np.std(rho[interior])


def test_init_applies_shepard(O):
    # src/main.cpp:69-79: the first filter runs after the first force pass
    p = O.sim_params(2, shepard_interval=30)
    q = O.sim_params(2, shepard_interval=0)
    a = O.State.dam_break(2, 0.05, col_w=0.5, col_h=0.6, tank_l=1.0, tank_h=1.0)
    b = O.State.dam_break(2, 0.05, col_w=0.5, col_h=0.6, tank_l=1.0, tank_h=1.0)
    a.init(p)
    b.init(q)
    assert np.allclose(a.dens, O.shepard(b.x, b.h, b.mass, b.dens), rtol=1e-14)
    assert np.array_equal(a.a, b.a)  # the force pass saw the unfiltered density


def test_step_composition_with_shepard(O):
    # a filtering step composed by hand: kick-drift, the filter on the drifted
    # state, pressure from the filtered density, force, gravity, final kick
    s = O.State.dam_break(2, 0.1, col_w=0.5, col_h=0.6, tank_l=1.0, tank_h=1.0)
    p = O.sim_params(2, alpha=0.3, shepard_interval=2)
    s.init(p)
    s.step(p)
    s.step(p)
    assert s.step_count == 2  # 2 % 2 == 0: the next step filters
    x0, v0, a0, rho0, dv0 = s.x.copy(), s.v.copy(), s.a.copy(), s.dens.copy(), s.divv.copy()
    dt = s.global_dt(p)
    s.step(p)
    fl = s.type == 0
    vh = np.where(fl, v0 + 0.5 * dt * a0, 0.0)
    x1 = np.where(fl, x0 + dt * vh, x0)
    vp = np.where(fl, v0 + dt * a0, v0)
    rho1 = O.shepard(x1, s.h, s.mass, rho0 - rho0 * dt * dv0)
    n = s.n
    epi = np.zeros(n, O.epi_dtype(2))
    epj = np.zeros(n, O.epj_dtype(2))
    P = np.array([O.eos_pressure(r, p.rho0, p.c0, p.gamma) for r in rho1])
    c = np.array([O.eos_soundspeed(r, p.rho0, p.c0, p.gamma) for r in rho1])
    for a in (epi, epj):
        a["pos"], a["vel"], a["smth"], a["dens"], a["pres"], a["snds"], a["visc"] = x1.T, vp.T, s.h, rho1, P, c, p.nu
    epj["mass"] = s.mass
    ref = O.hydro_force(O.force_params(2, alpha=p.alpha, c_cfl=p.c_cfl), epi, epj)
    acc = ref["acc"].T.copy()
    acc[1] -= p.g
    assert np.allclose(s.dens, rho1, rtol=1e-14)
    assert np.allclose(s.a, acc, rtol=1e-10, atol=1e-10 * np.abs(acc).max())
    # and the following step (index 3) does not filter
    rho2 = s.dens.copy()
    dv2 = s.divv.copy()
    dt2 = s.global_dt(p)
    s.step(p)
    assert np.allclose(s.dens, rho2 - rho2 * dt2 * dv2, rtol=1e-14)


@pytest.mark.parametrize("dim,R", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_neighbour_counts_cells_equal_brute_force(O, dim, R):
    # the cell-list count of the cell-local predicate (for large states)
    # equals the O(N^2) pass over the same grid, particle by particle, with
    # varying h and cells of 2 h_max / R
    n = 1500 if dim == 2 else 2500
    pos = RNG.uniform(0, 1, (dim, n))
    h = ((0.03 if dim == 2 else 0.07) * (1 + 0.3 * RNG.uniform(-1, 1, n))).astype(np.float32)
    L = 2 * float(h.max()) * (1 + 1e-5) / R
    lo = np.full(3, -0.05)
    nc = np.array([int(np.ceil(1.1 / L))] * 3, np.int32)
    ref = O.neighbours_brute_cells(pos, h, lo, L, nc)
    got = O.neighbour_counts_cells(pos, h, lo, L, nc, 4)
    assert np.array_equal(got, ref) and ref.sum() > 10 * n
    with pytest.raises(ValueError):
        O.neighbour_counts_cells(pos + 2.0, h, lo, L, nc)
