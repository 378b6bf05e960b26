"""The two drop-in front-ends on the GPU, against the oracle's restatement of
the reference schedule on the same per-replica MT19937 streams:
  - the `gasp` command line program (src/ising.cu): stdout lines,
    gridstates.bin records, the committor line;
  - the `gasp` extension module (src/ising_ext.cu): return values,
    module lists `magnetisation` / `grids`, GridSnapObject contents, pB and
    its bootstrap error (libc rand(), bit for bit)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BETA, H = 0.54, 0.07


def _droplet(L, a, r0=3):
    g = -np.ones((L, L), dtype=np.int32)
    g[r0:r0 + a, r0:r0 + a] = 1
    return g


def _records(path, L, R):
    raw = open(path, "rb").read()
    nb = (L * L * R + 7) // 8
    out = []
    for off in range(0, len(raw), 12 + nb):
        hdr = np.frombuffer(raw[off:off + 12], dtype=np.int32)
        bits = np.frombuffer(raw[off + 12:off + 12 + nb], dtype=np.uint8)
        g = np.where(np.unpackbits(bits, bitorder="little")[:L * L * R] == 1, 1, -1).reshape(R, L, L)
        out.append((tuple(hdr), g))
    return out


@pytest.mark.parametrize("method", [0, 1, 2])
def test_cli_itask0(gpu, engine_lib, oracle, tmp_path, method):
    L, R, seed = 16, 64, 123
    print(f"[cli itask0] method={method}", flush=True)
    env = dict(os.environ, GASP_SEED=str(seed), GASP_RNG="mt19937", GASP_L=str(L))
    r = subprocess.run([engine_lib.CLI_PATH, "300", str(R), "20", "40", "64", "0", str(method), str(BETA), str(H),
                        "0"], capture_output=True, text=True, cwd=tmp_path, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "# isweep    nucleated fraction"
    rows = [ln for ln in lines if ln[:10].strip().isdigit() and len(ln.split()) == 2]
    ref = oracle.driver(L, R, -np.ones((R, L, L), dtype=np.int32), rng_kind=0, seed=seed, beta=BETA, h=H,
                        tot_nsweeps=300, mag_output_int=20, grid_output_int=40, max_dumps=64)
    exp_rows = ["%10d  %12.6f" % (s, f) for s, f in zip(range(0, 300, 20), ref["result"])]
    assert rows == exp_rows
    assert "# Time taken on GPU =" in r.stdout
    recs = _records(tmp_path / "gridstates.bin", L, R)
    assert [h[2] for h, _ in recs] == list(ref["dump_sweeps"])
    for (hdr, g), d in zip(recs, ref["dumps"]):
        assert hdr[:2] == (L, R)
        assert np.array_equal(g.reshape(R, -1), d.reshape(R, -1))


def test_cli_itask1(gpu, engine_lib, oracle, tmp_path):
    L, R, seed = 16, 64, 99
    print("[cli itask1]", flush=True)
    g = _droplet(L, 6)
    bits = np.packbits((g.reshape(-1) == 1).astype(np.uint8), bitorder="little")
    (tmp_path / "gridinput.bin").write_bytes(np.int32(L).tobytes() + bits.tobytes())
    env = dict(os.environ, GASP_SEED=str(seed), GASP_RNG="mt19937", GASP_L=str(L))
    r = subprocess.run([engine_lib.CLI_PATH, "3000", str(R), "10", "100", "64", "0", "2", str(BETA), str(H), "1"],
                       capture_output=True, text=True, cwd=tmp_path, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    fates = np.full(R, -1, dtype=np.int32)
    ref = oracle.driver(L, R, np.repeat(g[None], R, axis=0), rng_kind=0, seed=seed, beta=BETA, h=H,
                        tot_nsweeps=3000, mag_output_int=10, grid_output_int=100, itask=1, dn_thr=-0.9,
                        up_thr=0.9, fates=fates)
    assert ("pB estimate : %10.6f" % ref["result"][0]) in r.stdout
    assert "Reached cv =  -0.90 :" in r.stdout


def _import_gasp(engine_lib):
    if engine_lib.PKG_DIR not in sys.path:
        sys.path.insert(0, engine_lib.PKG_DIR)
    import gasp

    return gasp


class _H5:
    """Reads int datasets through the HDF5 C library (hid_t is int64 in 1.10)."""

    def __init__(self, path="/opt/conda/lib/libhdf5.so"):
        self.lib = C.CDLL(path)
        self.lib.H5open()
        for f in ("H5Fopen", "H5Dopen2"):
            getattr(self.lib, f).restype = C.c_int64
        self.lib.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
        self.lib.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        self.lib.H5Dread.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        self.lib.H5Dclose.argtypes = [C.c_int64]
        self.lib.H5Fclose.argtypes = [C.c_int64]
        self.native_int = C.c_int64.in_dll(self.lib, "H5T_NATIVE_INT_g").value

    def read_int(self, path, name, shape):
        f = self.lib.H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
        assert f >= 0, path
        d = self.lib.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        out = np.zeros(shape, dtype=np.int32)
        assert self.lib.H5Dread(d, self.native_int, 0, 0, 0, out.ctypes.data) >= 0  # H5S_ALL = 0
        self.lib.H5Dclose(d)
        self.lib.H5Fclose(f)
        return int(out) if shape == () else out


def test_python_swarm(gpu, engine_lib, oracle, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    gasp = _import_gasp(engine_lib)
    assert gasp.gpu_nsms == gpu
    L, R, seed = 16, 64, 321
    print("[python swarm]", flush=True)
    frac = gasp.run_nucleation_swarm(L, R, 300, BETA, H, mag_output_int=20, grid_output_int=40, seed=seed,
                                     rng="mt19937", gpu_method=2, threadsPerBlock=64)
    ref = oracle.driver(L, R, -np.ones((R, L, L), dtype=np.int32), rng_kind=0, seed=seed, beta=BETA, h=H,
                        tot_nsweeps=300, mag_output_int=20, grid_output_int=40, up_thr=0.9, dn_thr=-0.95,
                        max_dumps=64)
    assert frac.dtype == np.float32 and frac.shape == (15,)
    assert np.array_equal(frac, ref["result"])
    cs = list(ref["check_sweeps"])
    assert len(gasp.magnetisation) == len(ref["dump_sweeps"])
    for k, sw in enumerate(ref["dump_sweeps"]):
        assert np.array_equal(np.array(gasp.magnetisation[k], dtype=np.float32), ref["cv"][cs.index(sw)])
    # snapshots: replicas with -0.95 < m < 0.9 at each dump, in replica order
    flat = [s for snaps in gasp.grids for s in snaps]
    exp = []
    for k, sw in enumerate(ref["dump_sweeps"]):
        m = ref["cv"][cs.index(sw)]
        for i in np.nonzero((m > -0.95) & (m < 0.9))[0]:
            exp.append((int(sw), int(i), float(m[i]), ref["dumps"][k][i].reshape(L, L)))
    assert len(flat) == len(exp)
    for s, (sw, i, m, grid) in zip(flat, exp):
        assert (s.isweep, s.igrid) == (sw, i) and s.magnetisation == np.float32(m) and s.committor == -1.0
        assert s.grid.dtype == np.int8 and np.array_equal(s.grid, grid)
    # HDF5 grid file (src/io.c layout): every replica at every dump
    h5 = _H5()
    dumps = list(ref["dump_sweeps"])
    assert h5.read_int("gridstates.hdf5", "/total_saved_grids", ()) == R * len(dumps)
    assert h5.read_int("gridstates.hdf5", "/L", ()) == L
    for k in (0, len(dumps) - 1):
        for i in (0, R - 1):
            g = h5.read_int("gridstates.hdf5", f"/sweep_{int(dumps[k])}/grid_{i}", (L, L))
            assert np.array_equal(g, ref["dumps"][k][i].reshape(L, L))
    # history limit; no HDF5 file on request
    os.remove("gridstates.hdf5")
    gasp.run_nucleation_swarm(L, R, 300, BETA, H, mag_output_int=20, grid_output_int=40, seed=seed,
                              rng="mt19937", max_keep_grids=5, write_hdf5=False)
    assert sum(len(s) for s in gasp.grids) == 5
    assert not os.path.exists("gridstates.hdf5")


def test_python_committor(gpu, engine_lib, oracle, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    gasp = _import_gasp(engine_lib)
    L, R, seed = 16, 128, 2024
    print("[python committor]", flush=True)
    inputs = [_droplet(L, 4).astype(np.int8), _droplet(L, 7).astype(np.int8)]
    libc = C.CDLL(None)
    libc.srand.argtypes = [C.c_uint]
    libc.srand(7)
    res = gasp.run_committor_calc(L, R, 3000, BETA, H, grid_input="NumPy", grid_array=inputs, nsms=2,
                                  mag_output_int=10, grid_output_int=100, seed=seed, rng="mt19937",
                                  up_threshold=0.9, dn_threshold=-0.93)
    fates = np.full(R, -1, dtype=np.int32)
    g0 = np.concatenate([np.repeat(x.astype(np.int32)[None], R // 2, axis=0) for x in inputs])
    ref = oracle.driver(L, R, g0, rng_kind=0, seed=seed, beta=BETA, h=H, tot_nsweeps=3000, mag_output_int=10,
                        grid_output_int=100, itask=1, dn_thr=-0.93, up_thr=0.9, ninputs=2, fates=fates)
    assert [p for p, _ in res] == [float(x) for x in ref["result"]]
    oracle.lib().srand(7)
    errs = [oracle.lib().orc_bootstrap_errbar(R // 2, R // 2, np.ascontiguousarray(ref["fates"][k * R // 2:(k + 1) * R // 2]))
            for k in range(2)]
    assert [e for _, e in res] == errs
    # the same run with GPU bootstrap error bars: same pB, statistically equal errors
    # (64 resamples of 64 fates each: a few tens of percent of noise per estimate)
    res_g = gasp.run_committor_calc(L, R, 3000, BETA, H, grid_input="NumPy", grid_array=inputs, nsms=2,
                                    mag_output_int=10, grid_output_int=100, seed=seed, rng="mt19937",
                                    up_threshold=0.9, dn_threshold=-0.93, bootstrap="gpu", write_hdf5=False)
    assert [p for p, _ in res_g] == [p for p, _ in res]
    for (_, eg), (p, e) in zip(res_g, res):
        if 0.0 < p < 1.0:
            assert 0.5 * e < eg < 2.0 * e
        else:
            assert eg == 0.0 and e == 0.0
    with pytest.raises(ValueError):
        gasp.run_committor_calc(L, 102, 10, BETA, H, grid_input="NumPy", grid_array=inputs, nsms=4)
    with pytest.raises(ValueError):
        gasp.run_committor_calc(L, R, 10, BETA, H, grid_input="NumPy", grid_array=inputs, nsms=2,
                                bootstrap="numpy")
