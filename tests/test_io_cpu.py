"""Snapshot and restart files on the host (no GPU): the reference's ASCII
column format byte for byte against the fixtures, read-back, error paths,
and restart-file validation that fails before any device call."""
import os
import struct

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def golden_inputs(dim):
    import importlib.util
    spec = importlib.util.spec_from_file_location("snapgold", os.path.join(GOLDEN, "make_snapshot_golden.py"))
    g = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(g)
    P = g.PARTICLES
    pos = np.array([[p[3][k] for p in P] for k in range(dim)])
    vel = np.array([[p[4][k] for p in P] for k in range(dim)])
    return g, dict(time=g.TIME, id=np.array([p[0] for p in P]), tag=np.array([p[1] for p in P]),
                   mass=np.array([p[2] for p in P]), pos=pos, vel=vel, dens=np.array([p[5] for p in P]),
                   snds=np.array([p[6] for p in P]), pres=np.array([p[7] for p in P]))


@pytest.mark.parametrize("dim", [2, 3])
def test_ascii_snapshot_matches_fixture(W, tmp_path, dim):
    g, a = golden_inputs(dim)
    out = tmp_path / "snap.dat"
    W.write_snapshot_ascii(out, a["time"], a["mass"], a["pos"], a["vel"], a["dens"], snds=a["snds"],
                           pres=a["pres"], id=a["id"], tag=a["tag"])
    with open(os.path.join(GOLDEN, f"snapshot_{dim}d.dat"), "rb") as f:
        want = f.read()
    assert out.read_bytes() == want
    # the fixture is what the restated format prints for these inputs
    assert want.decode() == g.text(dim)


@pytest.mark.parametrize("dim", [2, 3])
def test_ascii_snapshot_reads_fixture(W, dim):
    _, a = golden_inputs(dim)
    r = W.read_snapshot_ascii(os.path.join(GOLDEN, f"snapshot_{dim}d.dat"), dim)
    assert r["time"] == float("%e" % a["time"])
    assert np.array_equal(r["id"], a["id"]) and np.array_equal(r["tag"], a["tag"])
    rt = np.vectorize(lambda x: float("%.16f" % x))
    for k in ("mass", "pos", "vel", "dens", "pres"):
        assert np.array_equal(r[k], rt(a[k])), k
    if dim == 2:
        assert np.array_equal(r["snds"], rt(a["snds"]))
    else:
        assert np.all(r["snds"] == 0.0)


def test_ascii_defaults_and_chunk_order(W, tmp_path):
    # more lines than one formatting batch (64 chunks of 8192), ids from the
    # default (0..n-1), tags 0, snds/pres 0; every line lands in order
    rng = np.random.default_rng(7)
    n = 64 * 8192 + 1000
    pos = rng.uniform(-1, 4, (3, n))
    vel = rng.normal(0, 2, (3, n))
    dens = rng.uniform(990, 1010, n)
    mass = np.full(n, 0.002)
    out = tmp_path / "big.dat"
    W.write_snapshot_ascii(out, 1.25, mass, pos, vel, dens)
    r = W.read_snapshot_ascii(out, 3)
    assert np.array_equal(r["id"], np.arange(n))
    assert np.all(r["tag"] == 0) and np.all(r["pres"] == 0) and np.all(r["snds"] == 0)
    assert np.max(np.abs(r["pos"] - pos)) <= 5e-16 * 4
    assert np.max(np.abs(r["dens"] - dens)) <= 1e-13
    with open(out) as f:
        lines = [f.readline() for _ in range(3)]
    assert lines[0] == "1.250000e+00\n" and lines[1] == f"{n}\n"
    assert lines[2] == "0\t0\t" + "\t".join("%.16f" % v for v in
                                            [0.002, *pos[:, 0], *vel[:, 0], dens[0], 0.0, 0.0]) + "\n"


def test_ascii_errors(W, tmp_path):
    E = W._abi
    with pytest.raises(W.WcsphError) as e:
        W.read_snapshot_ascii(tmp_path / "missing.dat", 3)
    assert e.value.code == E.WCSPH_EIO
    bad = tmp_path / "trunc.dat"
    bad.write_text("0.000000e+00\n3\n0\t0\t1\t2\t3\t4\t5\t6\t7\t8\t9\t10\n")
    with pytest.raises(W.WcsphError, match="truncated") as e:
        W.read_snapshot_ascii(bad, 3)
    assert e.value.code == E.WCSPH_EIO
    bad.write_text("0.000000e+00\n2\n0\t0\t1\t2\t3\t4\t5\t6\t7\t8\t9\t10\n1\t0\t1\t2\tx\t4\t5\t6\t7\t8\t9\t10\n")
    with pytest.raises(W.WcsphError, match="line 4") as e:
        W.read_snapshot_ascii(bad, 3)
    assert e.value.code == E.WCSPH_EIO
    bad.write_text("garbage\n")
    with pytest.raises(W.WcsphError) as e:
        W.read_snapshot_ascii(bad, 3)
    assert e.value.code == E.WCSPH_EIO
    with pytest.raises(W.WcsphError) as e:
        W.write_snapshot_ascii(tmp_path / "no" / "such" / "dir.dat", 0.0, [1.0], np.zeros((3, 1)),
                               np.zeros((3, 1)), [1.0])
    assert e.value.code == E.WCSPH_EIO


def test_restart_file_validation_before_device(W, tmp_path):
    # every check below fails while reading the header, before any device call
    E = W._abi
    with pytest.raises(W.WcsphError) as e:
        W.Simulation.restore(tmp_path / "missing.bin")
    assert e.value.code == E.WCSPH_EIO
    f = tmp_path / "r.bin"
    f.write_bytes(b"WCSPH")
    with pytest.raises(W.WcsphError, match="shorter than its header"):
        W.Simulation.restore(f)
    f.write_bytes(b"NOTWCSPH" + bytes(1024))
    with pytest.raises(W.WcsphError, match="bad magic"):
        W.Simulation.restore(f)
    f.write_bytes(b"WCSPHRST" + struct.pack("<II", 99, 0) + bytes(1024))
    with pytest.raises(W.WcsphError, match="another version"):
        W.Simulation.restore(f)
