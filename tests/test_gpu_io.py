"""GPU checks of snapshots and restart files: a restored simulation continues
bit for bit; the device snapshot holds the downloaded state in the
reference's ASCII format; a run resumed the reference's way (read the
snapshot, init, set the clock) follows the oracle resumed the same way."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL3 = dict(tank_l=2.0, col_h=1.0, tank_h=1.5)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision", ["fp32", "mixed", "fp64"])
@pytest.mark.parametrize("dim,dx,geom", [(2, 0.02, {}), (3, 0.05, SMALL3)])
def test_restore_continues_bit_for_bit(W, tmp_path, dim, dx, geom, precision):
    # filter every 4 steps: steps 20 and 24 fall after the restore at 17;
    # odd counts mix graph replays with single steps on both sides
    print(f"case restore dim={dim} precision={precision}", flush=True)
    p = W.dam_break_params(dim, shepard_interval=4)
    st = W.dam_break_state(W.dam_break_desc(dim, dx, **geom), p)
    f = str(tmp_path / "r.bin")
    with W.Simulation.from_state(p, st, precision=precision) as a:
        a.init()
        a.step(17)
        a.save(f)
        a.step(11)
        want, want_st = a.download(), a.stats()
    with W.Simulation.restore(f) as b:
        assert b.n == len(st["dens"]) and b.dim == dim and b.params.shepard_interval == 4
        assert b.stats()["step"] == 17
        b.step(11)
        got, got_st = b.download(), b.stats()
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    for k in ("time", "dt", "step", "pairs_last", "pairs_total"):
        assert got_st[k] == want_st[k], k


def test_save_restore_save_is_identical(W, tmp_path):
    p = W.dam_break_params(2)
    st = W.dam_break_state(W.dam_break_desc(2, 0.03), p)
    f1, f2 = str(tmp_path / "a.bin"), str(tmp_path / "b.bin")
    with W.Simulation.from_state(p, st, precision="fp32") as a:
        a.init()
        a.step(5)
        a.save(f1)
    with W.Simulation.restore(f1) as b:
        b.save(f2)
    assert open(f1, "rb").read() == open(f2, "rb").read()


def test_save_needs_a_steppable_state(W, tmp_path):
    p = W.dam_break_params(2)
    st = W.dam_break_state(W.dam_break_desc(2, 0.05), p)
    with W.Simulation.from_state(p, st) as a:
        with pytest.raises(W.WcsphError):
            a.save(str(tmp_path / "x.bin"))  # nothing computed yet
        a.init()
        a.force()
        with pytest.raises(W.WcsphError):
            a.save(str(tmp_path / "x.bin"))  # hydro-only acc after force()
        with pytest.raises(W.WcsphError):
            a.step(1)
        a.init()
        a.step(2)
        a.save(str(tmp_path / "x.bin"))


def test_restore_rejects_corruption(W, tmp_path):
    p = W.dam_break_params(2)
    st = W.dam_break_state(W.dam_break_desc(2, 0.05), p)
    f = tmp_path / "r.bin"
    with W.Simulation.from_state(p, st) as a:
        a.init()
        a.save(str(f))
    raw = bytearray(f.read_bytes())
    raw[-100] ^= 0x01
    f.write_bytes(bytes(raw))
    with pytest.raises(W.WcsphError, match="checksum"):
        W.Simulation.restore(str(f))
    f.write_bytes(bytes(raw[:-8]))
    with pytest.raises(W.WcsphError, match="truncated"):
        W.Simulation.restore(str(f))


@pytest.mark.parametrize("dim,dx,geom", [(2, 0.02, {}), (3, 0.05, SMALL3)])
def test_device_snapshot_holds_the_state(W, tmp_path, dim, dx, geom):
    print(f"case snapshot dim={dim}", flush=True)
    p = W.dam_break_params(dim)
    st = W.dam_break_state(W.dam_break_desc(dim, dx, **geom), p)
    f = str(tmp_path / "s.dat")
    with W.Simulation.from_state(p, st, precision="fp32") as sim:
        sim.init()
        sim.step(9)
        sim.write_ascii(f)
        d, t = sim.download(), sim.stats()["time"]
    r = W.read_snapshot_ascii(f, dim)
    rt = np.vectorize(lambda x: float("%.16f" % x))
    n = len(st["dens"])
    assert r["time"] == float("%e" % t)
    assert np.array_equal(r["id"], np.arange(n)) and np.array_equal(r["tag"], st["type"])
    assert np.array_equal(r["pos"], rt(d["pos"])) and np.array_equal(r["vel"], rt(d["vel"]))
    assert np.array_equal(r["dens"], rt(d["dens"])) and np.array_equal(r["mass"], rt(st["mass"]))
    B = p.c0 ** 2 * p.rho0 / p.gamma
    x = d["dens"] / p.rho0
    pres = np.maximum(B * (x ** p.gamma - 1.0), 0.0)
    assert np.allclose(r["pres"], pres, rtol=1e-14, atol=1e-9)
    if dim == 2:
        snds = np.sqrt(np.maximum(B * p.gamma / p.rho0 * x ** (p.gamma - 1.0), 1e-16))
        assert np.allclose(r["snds"], snds, rtol=1e-14)
    else:
        assert np.all(r["snds"] == 0.0)


def test_reference_style_resume_follows_oracle(O, W, tmp_path):
    # write a snapshot at step 8, read it back, init (force + filter), set the
    # clock to the snapshot's time and the next step index; the filter (every
    # 5 steps) then falls at steps 10 and 15 on both sides
    k = 5
    p = W.dam_break_params(2, shepard_interval=k)
    op = O.sim_params(2, alpha=p.alpha, c_cfl=p.c_cfl, c0=p.c0, g=p.g, nu=p.nu, shepard_interval=k)
    st = W.dam_break_state(W.dam_break_desc(2, 0.02), p)
    f = str(tmp_path / "s.dat")
    with W.Simulation.from_state(p, st, precision="fp64") as sim:
        sim.init()
        sim.step(8)
        sim.write_ascii(f)
    r = W.read_snapshot_ascii(f, 2)
    n = len(r["dens"])
    state = dict(pos=r["pos"], vel=r["vel"], dens=r["dens"], mass=r["mass"], h=st["h"],
                 type=r["tag"].astype(np.int32))
    t0 = r["time"]
    with W.Simulation.from_state(p, state, precision="fp64") as sim:
        sim.init()
        sim.set_clock(t0, 9)
        s0 = sim.stats()
        assert s0["time"] == t0 and s0["step"] == 9
        sim.step(10)
        got, gs = sim.download(), sim.stats()
    s = O.State(2, n)
    s.x[:], s.v[:], s.dens[:], s.mass[:], s.h[:], s.type[:] = (state["pos"], state["vel"], state["dens"],
                                                              state["mass"], state["h"], state["type"])
    s.init(op, 8)
    s.c.time, s.c.step = t0, 9
    for _ in range(10):
        s.step(op, 8)
    assert gs["step"] == s.step_count == 19
    assert abs(gs["time"] - s.time) <= 1e-12 * s.time
    assert np.allclose(got["dens"], s.dens, rtol=1e-11)
    assert rel(got["pos"] - state["pos"], s.x - state["pos"]) < 1e-9


def test_output_schedule_follows_reference(W, tmp_path):
    # OutputFileWithTimeInterval: a snapshot at the start and after each step
    # whose time reaches the next output time; the times come from a run
    # stepped one step at a time and read back after every step
    p = W.dam_break_params(2)
    st = W.dam_break_state(W.dam_break_desc(2, 0.05), p)
    end, nsnap = 0.02, 7
    times = [0.0]
    with W.Simulation.from_state(p, st, precision="fp32") as sim:
        sim.init()
        while times[-1] < end:
            sim.step(1)
            times.append(sim.stats()["time"])
    want, t_next = [], 0.0
    for t in times:
        if t >= t_next:
            want.append(t)
            t_next += end / nsnap
    with W.Simulation.from_state(p, st, precision="fp32") as sim:
        sim.init()
        k = sim.run_with_output(end, nsnap, str(tmp_path))
        assert sim.stats()["time"] == times[-1]
    assert k == len(want)
    got = [W.read_snapshot_ascii(W.snapshot_path(str(tmp_path), i), 2)["time"] for i in range(k)]
    assert got == [float("%e" % t) for t in want]
    assert (tmp_path / "00000_00001_00000.bin").exists()


def test_cpp_driver_matches_python_and_resumes(W, tmp_path):
    # the C++ program of tests/cpp/sim_driver.cpp (C-ABI only) writes the same
    # snapshot files as Simulation.run_with_output, and a run resumed from
    # its restart file of snapshot 3 writes the same later snapshots
    import os
    import subprocess
    drv = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "sim_driver")
    end, nsnap, k = 0.02, 6, 5
    p = W.dam_break_params(2, shepard_interval=k)
    st = W.dam_break_state(W.dam_break_desc(2, 0.05), p)
    dpy, d1, d2 = (tmp_path / x for x in ("py", "cpp", "resumed"))
    for d in (dpy, d1, d2):
        d.mkdir()
    with W.Simulation.from_state(p, st, precision="fp32") as sim:
        sim.init()
        n = sim.run_with_output(end, nsnap, str(dpy))
    args = ["2", "0.05", repr(end), str(nsnap)]
    subprocess.run([drv, *args, str(d1), "0", str(k)], check=True, timeout=120)
    subprocess.run([drv, *args, str(d2), "0", str(k), str(d1 / "00003_00001_00000.bin"), "3"], check=True,
                   timeout=120)
    names = sorted(f.name for f in dpy.glob("*.dat"))
    assert len(names) == n and names == sorted(f.name for f in d1.glob("*.dat"))
    for f in names:
        assert (d1 / f).read_bytes() == (dpy / f).read_bytes(), f
    resumed = sorted(f.name for f in d2.glob("*.dat"))
    assert resumed == names[3:]
    for f in resumed:
        assert (d2 / f).read_bytes() == (d1 / f).read_bytes(), f
