import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libwcsph.so")


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def W():
    import weaklycompressiblesph_amd as W
    W._abi.lib()
    return W


@pytest.fixture(scope="session")
def dt_neighbours(O):
    """Per-particle dt of src/force.h:56 with v_sig taken over the true
    neighbours (the device predicate, via the oracle's cell-local brute
    force): dt_neighbours(s, op, (lo, L, nc)). The device-resident passes
    use this set; the oracle's force pass takes v_sig over all 27-cell
    candidates, as the reference does over the lists it is handed, which
    differs in v_sig mode 1 (max of -mu)."""
    import numpy as np

    def dt(s, op, grid):
        lo, L, nc = grid
        counts, lists = O.neighbours_brute_cells(s.x, s.h.astype(np.float32), lo, L, nc, cap=160)
        assert counts.max() < 160
        I = np.repeat(np.arange(s.n), counts)
        J = lists[np.arange(lists.shape[1])[None, :] < counts[:, None]]
        dr, dv = s.x[:, I] - s.x[:, J], s.v[:, I] - s.v[:, J]
        r2, drdv = (dr * dr).sum(0), (dr * dv).sum(0)
        hi, hj = s.h[I], s.h[J]
        mu = np.where(drdv < 0.0, 0.5 * (hi + hj) * drdv / (r2 + op.eps * hi * hj), 0.0)
        vsig = np.zeros(s.n)
        np.maximum.at(vsig, I, -mu if op.vsig_mode else mu)
        return op.c_cfl * s.h / (s.snds + vsig)
    return dt


class DeviceMemory:
    """Device buffers for the entry points that take device pointers,
    through the HIP runtime libwcsph.so itself links (libamdhip64.so.7),
    so the pointers belong to the library's runtime."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.live = []

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        self._ok(self.hip.hipMalloc(self.C.byref(p), max(int(nbytes), 1)), "hipMalloc")
        self.live.append(p.value)
        return p.value

    def upload(self, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self._ok(self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, shape, dtype):
        import numpy as np
        out = np.empty(shape, dtype)
        self.sync()
        self._ok(self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def sync(self):
        self._ok(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free_all(self):
        self.sync()
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


@pytest.fixture
def devmem(W):
    m = DeviceMemory()
    yield m
    m.free_all()
