"""gasp_sweep_grids_noalloc: the sweeps of gasp_sweep_grids with the workspace
in the library's own device array instead of a per-call allocation.  Through
the C ABI, against gasp_sweep_grids on twin inputs (same grids, twin Philox
states of the same seed): grids and magnetisations bit for bit, for every
method, on two calls in a row of which the second is smaller (the array then
still holds the first call's lattice), with and without a caller
magnetisation buffer; a size beyond the array is refused and leaves the grids
alone."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GASP_ERR_ARG = 1


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


def _sweep(G, hip, fn, L, R, method, seed, g0, calls, with_mag):
    """Runs `calls` (sweeps per call) of entry point `fn` on a device copy of
    g0 with a fresh Philox state; returns the grids and magnetisations after
    every call."""
    d_g = hip.alloc(g0.nbytes)
    d_m = hip.alloc(R * 4) if with_mag else None
    hip.h2d(d_g, g0)
    rng = G.RngState(R, seed)
    out = []
    try:
        for nsw in calls:
            G.check(fn(L, R, method, rng.ptr, 0.52, 0.06, d_g, nsw, d_m, None))
            hip.sync()
            got = np.empty_like(g0)
            hip.d2h(got, d_g)
            mag = None
            if with_mag:
                mag = np.empty(R, dtype=np.float32)
                hip.d2h(mag, d_m)
            out.append((got, mag))
    finally:
        rng.free()
        hip.free(d_g)
        if d_m is not None:
            hip.free(d_m)
    return out


# (method, (L, R) of the first call, (L, R) of the second, smaller call)
CASES = [(2, (64, 1000), (32, 300)), (1, (48, 600), (24, 70)), (0, (64, 500), (16, 50))]


@pytest.mark.parametrize("method,first,second", CASES)
def test_noalloc_matches_allocating_entry_point(gpu, engine_lib, hip, method, first, second):
    G = engine_lib
    lib = G.lib()
    for k, ((L, R), with_mag) in enumerate([(first, True), (second, True), (second, False)]):
        print(f"[noalloc] method={method} L={L} R={R} call {k} d_mag={'yes' if with_mag else 'NULL'}", flush=True)
        rs = np.random.default_rng(1000 * method + 10 * k + L)
        g0 = np.where(rs.random((R, L, L)) < 0.35, 1, -1).astype(np.int32)
        seed = 90 + 7 * k + method
        calls = (3, 2)
        ref = _sweep(G, hip, lib.gasp_sweep_grids, L, R, method, seed, g0, calls, with_mag)
        got = _sweep(G, hip, lib.gasp_sweep_grids_noalloc, L, R, method, seed, g0, calls, with_mag)
        for (g_ref, m_ref), (g_got, m_got) in zip(ref, got):
            assert not np.array_equal(g_ref, g0)
            assert np.array_equal(g_got, g_ref)
            if with_mag:
                assert np.array_equal(m_got, m_ref)


def test_noalloc_refuses_beyond_its_array(gpu, engine_lib, hip):
    G = engine_lib
    lib = G.lib()
    # method 0 at L = 256 keeps a byte per site: 300 replicas need 19.7 MB
    L, R = 256, 300
    g0 = -np.ones((R, L, L), dtype=np.int32)
    d_g = hip.alloc(g0.nbytes)
    d_m = hip.alloc(R * 4)
    hip.h2d(d_g, g0)
    rng = G.RngState(R, 3)
    print(f"[noalloc] refusal at method=0 L={L} R={R}", flush=True)
    rc = lib.gasp_sweep_grids_noalloc(L, R, 0, rng.ptr, 0.5, 0.1, d_g, 1, d_m, None)
    assert rc == GASP_ERR_ARG and lib.gasp_last_error() == GASP_ERR_ARG
    assert b"16777216" in lib.gasp_error_string()
    hip.sync()
    got = np.empty_like(g0)
    hip.d2h(got, d_g)
    assert np.array_equal(got, g0)
    rng.free()
    hip.free(d_g)
    hip.free(d_m)
