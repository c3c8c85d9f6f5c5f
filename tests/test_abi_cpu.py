"""CPU-side checks of the C-ABI library: it loads, exports every declared
symbol, its records match the reference's FDPS layouts, its host-side
pieces (parameters, initial condition, error paths) behave, and the C++
bridge header compiles against the reference's record shapes. No call
here touches a GPU.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wcsph_abi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wcsph_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(W):
    lib = W._abi.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(W._abi.EXPORTS)


def test_record_layouts_match_reference(tmp_path):
    # EPI/EPJ/RESULT::Hydro of the 64-bit FDPS build (src/class.h:63-67,
    # 167-176, 212-223): 3D 96/112/40 bytes, 2D 80/96/32 bytes
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stddef.h>
#include "wcsph_abi.h"
#define A(x) _Static_assert(x, #x)
A(sizeof(wcsph_epi_hydro_3d) == 96); A(sizeof(wcsph_epj_hydro_3d) == 112); A(sizeof(wcsph_result_hydro_3d) == 40);
A(sizeof(wcsph_epi_hydro_2d) == 80); A(sizeof(wcsph_epj_hydro_2d) == 96); A(sizeof(wcsph_result_hydro_2d) == 32);
A(offsetof(wcsph_epi_hydro_3d, smth) == 48); A(offsetof(wcsph_epi_hydro_3d, visc) == 80);
A(offsetof(wcsph_epi_hydro_3d, tag) == 88);
A(offsetof(wcsph_epj_hydro_3d, dens) == 48); A(offsetof(wcsph_epj_hydro_3d, mass) == 56);
A(offsetof(wcsph_epj_hydro_3d, smth) == 64); A(offsetof(wcsph_epj_hydro_3d, type) == 104);
A(offsetof(wcsph_result_hydro_3d, dt) == 24); A(offsetof(wcsph_result_hydro_3d, div_v) == 32);
A(offsetof(wcsph_epj_hydro_2d, type) == 88);
int main(void) { return 0; }
''')
    subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)], check=True)


def test_ctypes_structs_match_header(W, tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "wcsph_abi.h"
int main(void) { printf("%zu %zu %zu %zu %zu\n", sizeof(wcsph_params), sizeof(wcsph_sim_config),
    sizeof(wcsph_sim_stats), sizeof(wcsph_dam_break_desc), offsetof(wcsph_params, device));
    printf("%zu\n", offsetof(wcsph_params, wall_mode)); return 0; }
''')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    a = W._abi
    assert got == [C.sizeof(a.Params), C.sizeof(a.SimConfig), C.sizeof(a.SimStats), C.sizeof(a.DamBreakDesc),
                   a.Params.device.offset, a.Params.wall_mode.offset]
    for dim in (2, 3):
        assert W.epi_dtype(dim).itemsize == (96 if dim == 3 else 80)
        assert W.epj_dtype(dim).itemsize == (112 if dim == 3 else 96)
        assert W.res_dtype(dim).itemsize == (40 if dim == 3 else 32)


def test_reference_params(W):
    p = W.reference_params(3)
    assert (p.alpha, p.beta, p.eps, p.c_cfl, p.vsig_mode) == (0.01, 0.0, 0.01, 0.3, 0)
    assert p.c0 == pytest.approx(20 * np.sqrt(2 * 9.8 * 0.55))
    assert (p.rho0, p.gamma, p.g, p.nu) == (1000.0, 7.0, 9.8, 1e-6)
    q = W.dam_break_params(2)
    assert q.c0 == pytest.approx(10 * np.sqrt(2 * 9.81 * 2.0)) and q.alpha == 0.1 and q.c_cfl == 0.25


@pytest.mark.parametrize("dim,dx", [(2, 0.02), (2, 0.05), (3, 0.05)])
def test_initial_condition_matches_oracle(W, O, dim, dx):
    p = W.dam_break_params(dim)
    st = W.dam_break_state(W.dam_break_desc(dim, dx), p)
    o = O.State.dam_break(dim, dx, c0=p.c0, g=p.g)
    assert st["pos"].shape == o.x.shape
    assert np.array_equal(st["pos"], o.x)
    assert np.array_equal(st["type"], o.type)
    assert np.array_equal(st["mass"], o.mass)
    assert np.array_equal(st["h"], o.h)
    assert np.allclose(st["dens"], o.dens, rtol=1e-15)


def test_config_particle_counts(W):
    # 2D: 5,000 fluid + 1,518 wall; 3D at dx = 0.0126: 502,440 + 562,938
    for dim, dx, nf, nw in [(2, 0.02, 5000, 1518), (3, 0.0126, 502440, 562938)]:
        a, b = C.c_int64(), C.c_int64()
        W.check(W._abi.lib().wcsph_dam_break_count(C.byref(W.dam_break_desc(dim, dx)), C.byref(a), C.byref(b)))
        assert (a.value, b.value) == (nf, nw)


def test_error_paths_without_gpu(W):
    lib = W._abi.lib()
    # retrieve before any init is an ordering error, reported not thrown
    assert lib.wcsph_hydro_retrieve_3d(0, 0, None, None) == W._abi.WCSPH_ESTATE
    assert "wcsph_init" in W.last_error()
    assert lib.wcsph_init(None) == W._abi.WCSPH_EINVAL
    p = W.reference_params(3)
    p.dim = 4
    h = C.c_void_p()
    assert lib.wcsph_sim_create(C.byref(p), None, 10, C.byref(h)) == W._abi.WCSPH_EINVAL
    assert not h.value
    # wall_mode: 0 or 1 on the device-resident step, 0 only for the FDPS drop-in
    p = W.reference_params(3)
    p.wall_mode = 2
    assert lib.wcsph_sim_create(C.byref(p), None, 10, C.byref(h)) == W._abi.WCSPH_EINVAL
    assert "wall_mode" in W.last_error() and not h.value
    p.wall_mode = 1
    assert lib.wcsph_init(C.byref(p)) == W._abi.WCSPH_EINVAL
    assert "wall_mode" in W.last_error()
    with pytest.raises(W.WcsphError):
        W.check(lib.wcsph_sim_step(None, 1))
    bad = W.dam_break_desc(3, -1.0)
    assert lib.wcsph_dam_break_count(C.byref(bad), None, None) == W._abi.WCSPH_EINVAL
    assert W._abi.lib().wcsph_sim_bytes(3, 1_000_000, None) > 100 * 1_000_000


@pytest.mark.parametrize("two_d", [False, True])
def test_fdps_bridge_header_compiles(two_d, tmp_path):
    flags = ["-DPARTICLE_SIMULATOR_TWO_DIMENSION"] if two_d else []
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", *flags, "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "bridge_driver.cpp")],
                   check=True)


def test_bridge_driver_binaries_built():
    for d in ("2d", "3d"):
        assert os.path.exists(os.path.join(ROOT, "tests", "cpp", f"bridge_driver_{d}"))
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "sim_driver"))


def test_sim_driver_compiles_against_the_c_header():
    # the device-resident driver uses nothing but include/wcsph_abi.h
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "sim_driver.cpp")], check=True)


def test_host_worker_pool(tmp_path):
    # parallel_for's persistent pool: each index once, nested calls inline,
    # exceptions rethrown after every part, concurrent callers (host code only)
    exe = str(tmp_path / "pool_check")
    subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-O1", "-x", "hip", "--offload-arch=gfx950",
                    "--cuda-host-only", os.path.join(ROOT, "tests", "cpp", "pool_check.cpp"), "-o", exe,
                    "-pthread"], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120)
    assert "pool checks: ok" in out.stdout


def test_cell_order_ranking_host_replay():
    # the ranked gather of k_reorder (ranked_member, compiled for the host by
    # hipcc) puts every cell of 1..256 members into ascending order of the
    # previous index, whatever order the scatter left it in; cells beyond
    # 256 are flagged overfull instead of ranked (the search is quadratic)
    exe = os.path.join(ROOT, "tests", "cpp", "rank_check")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
