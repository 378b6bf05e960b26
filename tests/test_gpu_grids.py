"""gasp_sweep_grids: stream-ordered sweeps of caller-owned int32 device grids
(the layout gpuInitGrid makes) with the counter kept in the RNG state's device
header.  Against the oracle's restatement on the same Philox streams
(src/mc_cpu.c:41-89 arithmetic, single-precision table of src/mc_gpu.cu:21-37):
grids and magnetisations after two calls in a row (the second continues the
first's chain), for every method; a non-Philox state is refused on the
device (grids untouched, NaN magnetisations)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.h.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self, p):
        self.h.hipFree(p)


@pytest.fixture(scope="module")
def hip():
    return Hip()


@pytest.mark.parametrize("method,L,R", [(2, 32, 200), (2, 64, 130), (1, 24, 70), (0, 16, 50), (0, 64, 9)])
def test_sweep_grids_matches_oracle(gpu, engine_lib, oracle, hip, method, L, R):
    G = engine_lib
    lib = G.lib()
    beta, h, seed = 0.5, 0.08, 4242
    rs = np.random.default_rng(L * 1000 + R)
    g0 = np.where(rs.random((R, L, L)) < 0.35, 1, -1).astype(np.int32)
    print(f"[sweep_grids] method={method} L={L} R={R}", flush=True)
    d_g = hip.alloc(g0.nbytes)
    d_m = hip.alloc(R * 4)
    hip.h2d(d_g, g0)
    rng = G.RngState(R, seed)
    ref = g0.copy()
    a = 0
    for nsw in (3, 2):
        G.check(lib.gasp_sweep_grids(L, R, method, rng.ptr, beta, h, d_g, nsw, d_m, None))
        hip.sync()
        oracle.sweep_philox(L, ref, seed, a, nsw * L * L, beta, h)
        a += nsw * L * L
        got = np.empty_like(g0)
        hip.d2h(got, d_g)
        mag = np.empty(R, dtype=np.float32)
        hip.d2h(mag, d_m)
        assert np.array_equal(got, ref)
        assert np.array_equal(mag, oracle.magnetisation_gpu(ref))
    # no magnetisation buffer: the workspace holds it
    G.check(lib.gasp_sweep_grids(L, R, method, rng.ptr, beta, h, d_g, 1, None, None))
    hip.sync()
    oracle.sweep_philox(L, ref, seed, a, L * L, beta, h)
    got = np.empty_like(g0)
    hip.d2h(got, d_g)
    assert np.array_equal(got, ref)
    rng.free()
    hip.free(d_g)
    hip.free(d_m)


def test_sweep_grids_refuses_mt_state(gpu, engine_lib, hip):
    G = engine_lib
    lib = G.lib()
    L, R = 16, 64
    g0 = -np.ones((R, L, L), dtype=np.int32)
    d_g = hip.alloc(g0.nbytes)
    d_m = hip.alloc(R * 4)
    hip.h2d(d_g, g0)
    rng = G.RngState(R, 5, mt=True)
    G.check(lib.gasp_sweep_grids(L, R, 2, rng.ptr, 0.5, 0.1, d_g, 2, d_m, None))
    hip.sync()
    got = np.empty_like(g0)
    hip.d2h(got, d_g)
    mag = np.empty(R, dtype=np.float32)
    hip.d2h(mag, d_m)
    assert np.array_equal(got, g0) and np.isnan(mag).all()
    rng.free()
    hip.free(d_g)
    hip.free(d_m)
