"""mc_driver_gpu through the C ABI against the oracle's restatement of the
reference schedule (src/mc_gpu.cu:530-831) and against the reference's own
golden vectors (L = 32, 256 replicas, 2000 sweeps, computed by the compiled src/mc_cpu.c).

Observable outputs compared exactly: the result array (nucleated fraction or
pB), the fates, and everything handed to the output callback (sweep number,
host grids, magnetisation / cluster sizes)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_ref.npz")


class Capture:
    """Output callback recording what the driver hands to GridOutputFunc."""

    def __init__(self, L, R, lclus):
        self.L, self.R, self.lclus = L, R, lclus
        self.sweeps, self.grids, self.mags, self.clus = [], [], [], []

    def __call__(self, L, ngrids, grid_data, isweep, mag, lclus, cv, dn, up, fname):
        n = ngrids * L * L
        self.sweeps.append(isweep)
        self.grids.append(np.ctypeslib.as_array(grid_data, shape=(n,)).copy().reshape(ngrids, L, L))
        self.mags.append(np.ctypeslib.as_array(mag, shape=(ngrids,)).copy())
        if self.lclus:
            self.clus.append(np.ctypeslib.as_array(lclus, shape=(ngrids,)).copy())
        return 0


def _run(G, L, g0, *, mt, seed, method, beta=0.54, h=0.07, tot=300, mag=20, gout=40, itask=0,
         dn=-0.9, up=0.9, ninputs=1, cv="magnetisation", fates=None):
    R = g0.shape[0]
    rng = G.RngState(R, seed, mt=mt)
    cap = Capture(L, R, cv == "largest_cluster")
    try:
        out = G.run_driver(L, g0, beta=beta, h=h, rng=rng, method=method, tot_nsweeps=tot,
                           mag_output_int=mag, grid_output_int=gout, itask=itask, dn_thr=dn, up_thr=up,
                           ninputs=ninputs, cv=cv, fates=fates, outfunc=cap)
    finally:
        rng.free()
    return out, cap


def _oracle(O, L, g0, *, mt, seed, beta=0.54, h=0.07, tot=300, mag=20, gout=40, itask=0, dn=-0.9,
            up=0.9, ninputs=1, cv=0, fates=None):
    R = g0.shape[0]
    return O.driver(L, R, g0.copy(), rng_kind=0 if mt else 1, seed=seed, beta=beta, h=h,
                    tot_nsweeps=tot, mag_output_int=mag, grid_output_int=gout, itask=itask, cv_kind=cv,
                    dn_thr=dn, up_thr=up, ninputs=ninputs, fates=fates, max_dumps=64)


def _compare(out, cap, ref, mag, gout):
    assert np.array_equal(out["result"], ref["result"])
    assert cap.sweeps == list(ref["dump_sweeps"])
    for k in range(len(cap.sweeps)):
        assert np.array_equal(cap.grids[k], ref["dumps"][k].reshape(cap.grids[k].shape))
    # the callback's cv array is the one of the latest check at or before its sweep
    cs = list(ref["check_sweeps"])
    for k, sw in enumerate(cap.sweeps):
        j = max(i for i, c in enumerate(cs) if c <= sw)
        arr = cap.clus[k] if cap.lclus else cap.mags[k]
        assert np.array_equal(arr, ref["cv"][j])


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("mt", [True, False])
def test_itask0_matches_oracle(gpu, engine_lib, oracle, method, mt):
    print(f"[driver itask0] method={method} mt={mt}", flush=True)
    L, R = 16, 96
    g0 = -np.ones((R, L, L), dtype=np.int32)
    out, cap = _run(engine_lib, L, g0, mt=mt, seed=4242, method=method)
    ref = _oracle(oracle, L, g0, mt=mt, seed=4242)
    _compare(out, cap, ref, 20, 40)


@pytest.mark.parametrize("method", [1, 2])
def test_itask1_fates_and_pB(gpu, engine_lib, oracle, method):
    print(f"[driver itask1] method={method}", flush=True)
    L, R, nin = 16, 128, 4
    rs = np.random.default_rng(9)
    inputs = []
    for k in range(nin):  # -1 sea with a +1 square droplet of growing size
        g = -np.ones((L, L), dtype=np.int32)
        a = 2 + 2 * k
        g[4:4 + a, 4:4 + a] = 1
        inputs.append(g)
    g0 = np.repeat(np.stack(inputs), R // nin, axis=0)
    fates = np.full(R, -1, dtype=np.int32)
    out, cap = _run(engine_lib, L, g0, mt=False, seed=77, method=method, itask=1, dn=-0.8, up=0.5,
                    ninputs=nin, fates=fates.copy(), tot=2000, mag=10, gout=50)
    ref = _oracle(oracle, L, g0, mt=False, seed=77, itask=1, dn=-0.8, up=0.5, ninputs=nin,
                  fates=fates.copy(), tot=2000, mag=10, gout=50)
    assert np.array_equal(out["fates"], ref["fates"])
    _compare(out, cap, ref, 10, 50)


def test_cluster_cv_matches_oracle(gpu, engine_lib, oracle):
    print("[driver cluster cv]", flush=True)
    L, R = 16, 64
    g0 = -np.ones((R, L, L), dtype=np.int32)
    out, cap = _run(engine_lib, L, g0, mt=False, seed=5, method=2, cv="largest_cluster", up=60.0, dn=5.0,
                    tot=400, mag=20, gout=20)
    ref = _oracle(oracle, L, g0, mt=False, seed=5, cv=1, up=60.0, dn=5.0, tot=400, mag=20, gout=20)
    _compare(out, cap, ref, 20, 20)


@pytest.mark.parametrize("method", [1, 2])
def test_plain_run_against_reference_goldens(gpu, engine_lib, oracle, method):
    """The plain reference run (L = 32, 256 replicas, 2000 sweeps, seeds
    base + r) through mc_driver_gpu with per-replica MT19937 streams:
    the magnetisation at every check (handed to the output callback) and the final
    grids equal the reference's compiled mc_cpu.c, bit for bit."""
    gold = np.load(GOLD)
    L, nrep, chunks, spc, seed = (int(x) for x in gold["c1_meta"])
    print(f"[plain run L=32] method={method}", flush=True)
    g0 = -np.ones((nrep, L, L), dtype=np.int32)
    out, cap = _run(engine_lib, L, g0, mt=True, seed=seed, method=method, tot=chunks * spc, mag=spc,
                    gout=spc)
    assert cap.sweeps == [spc * c for c in range(chunks)]
    got_m = np.stack(cap.mags, axis=1)
    assert np.array_equal(got_m, gold["c1_mags"])
    final = np.where(np.unpackbits(gold["c1_final"], bitorder="little")[:nrep * L * L] == 1, 1, -1)
    assert np.array_equal(out["final_grids"].reshape(-1), final)
    frac = (gold["c1_mags"] > 0.9).mean(axis=0).astype(np.float32)
    assert np.array_equal(out["result"], frac)


def _max_l(shmem: int, tpb: int) -> int:
    """src/gpu_tools.cu:212-214 evaluated step by step."""
    m = int(math.sqrt(8.0 * shmem / tpb))
    return int(math.sqrt(32 * (m * m // 32)))


def test_select_gpu_method(gpu, engine_lib):
    """The reference's rule (src/gpu_tools.cu:205-232): 2 up to maxL, 0 beyond, at the
    device's shared memory per block (160 KB on MI355X: maxL 201 at 32 threads per
    block, 142 at 64)."""
    lib = engine_lib.lib()
    hip = C.CDLL("libamdhip64.so")
    v = C.c_int()
    # hipDeviceAttributeMaxSharedMemoryPerBlock
    attr = 74
    assert hip.hipDeviceGetAttribute(C.byref(v), attr, 0) == 0
    shmem = v.value
    assert shmem == 160 * 1024
    assert (_max_l(shmem, 32), _max_l(shmem, 64)) == (201, 142)
    for tpb, edge in ((32, 201), (64, 142), (128, _max_l(shmem, 128))):
        assert lib.select_gpu_method(edge, 1024, tpb, 0) == 2
        assert lib.select_gpu_method(edge + 1, 1024, tpb, 0) == 0
        assert lib.select_gpu_method(edge, 1024, tpb, -1) == 2  # current device
    assert lib.select_gpu_method(4000, 16, 32, 0) == 0
    assert lib.select_gpu_method(2, 16, 64, 0) == 2


def test_driver_argument_errors(gpu, engine_lib):
    G = engine_lib
    L, R = 8, 64
    g0 = -np.ones((R, L, L), dtype=np.int32)
    rng = G.RngState(R, 1)
    try:
        for kw, msg in [(dict(method=3), "Unknown gpu_method"),
                        (dict(method=2, threads_per_block=128), "Invalid threadsPerBlock"),
                        (dict(method=2, itask=2), "itask must be 0 or 1"),
                        (dict(method=2, itask=1, ninputs=3), "divisible by ninputs")]:
            with pytest.raises(G.GaspError, match=msg):
                G.run_driver(L, g0, beta=0.5, h=0.0, rng=rng, tot_nsweeps=10, mag_output_int=5,
                             grid_output_int=5, **kw)
    finally:
        rng.free()


def test_gridstates_dump_fast_path(gpu, engine_lib, tmp_path, monkeypatch):
    """With write_ising_grids itself as the output function the driver packs
    dumps on the device; the file and the caller's host grids must equal those
    of the generic path (a callback that calls write_ising_grids)."""
    import ctypes as C

    G = engine_lib
    monkeypatch.chdir(tmp_path)
    L, R, seed = 16, 128, 4242
    lib = G.lib()
    print("[driver fast dump]", flush=True)

    def generic(Lc, n, grid, isweep, mag, lclus, cv, dn, up, fn):
        return lib.write_ising_grids(Lc, n, grid, isweep, mag, lclus, cv, dn, up, fn)

    outs = []
    for outfunc in (generic, "write_ising_grids"):
        rng = G.RngState(R, seed)
        res = G.run_driver(L, -np.ones((R, L, L), dtype=np.int32), beta=0.54, h=0.07, rng=rng, method=2,
                           tot_nsweeps=310, mag_output_int=20, grid_output_int=40, outfunc=outfunc)
        rng.free()
        outs.append((open("gridstates.bin", "rb").read(), res["host_grids"].copy(), res["result"]))
        os.remove("gridstates.bin")
    (fa, ha, ra), (fb, hb, rb) = outs
    rec = 12 + R * L * L // 8
    assert len(fa) > 0 and len(fa) % rec == 0  # dumps every 40 sweeps until the run ends
    assert fa == fb
    assert np.array_equal(ha, hb) and np.array_equal(ra, rb)
