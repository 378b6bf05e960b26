"""CPU tests of the drop-in boundary: the C-ABI library loads and exports every
function include/gasp_amd.h declares; the host-side pieces (initial grids,
fates, gridinput.bin / gridstates.bin, bootstrap error bar) match the
reference's formats and golden values; the CLI and the `gasp` extension
behave like the reference's when no GPU is visible (the CPU fallback itself is
tests/test_cpu_fallback.py)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gasp_amd.h")
GOLD = os.path.join(ROOT, "tests", "golden", "golden_ref.npz")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", "", text)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


def _gpu_visible():
    try:
        hip = C.CDLL("libamdhip64.so")
        n = C.c_int(0)
        return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def test_header_symbols_exported(engine_lib):
    names = declared_functions()
    assert len(names) >= 30
    lib = engine_lib.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(engine_lib.EXPORTS)


def test_init_grids(engine_lib):
    lib = engine_lib.lib()
    L, R = 5, 6
    p = lib.init_grids_uniform(L, R, -1)
    g = np.ctypeslib.as_array(p, shape=(R * L * L,)).copy()
    assert np.all(g == -1)
    arr = np.arange(2 * L * L, dtype=np.int32) % 3 - 1
    p = lib.init_grids_from_array(L, R, 2, arr.ctypes.data_as(C.POINTER(C.c_int)))
    g = np.ctypeslib.as_array(p, shape=(R, L * L))
    for r in range(R):  # src/grid.c:59-67: input m fills replicas m*R/n .. (m+1)*R/n - 1
        assert np.array_equal(g[r], arr.reshape(2, -1)[r // 3])
    assert not lib.init_grids_from_array(L, 7, 2, arr.ctypes.data_as(C.POINTER(C.c_int)))
    assert lib.gasp_last_error() == 1
    p = lib.init_fates(9)
    assert np.all(np.ctypeslib.as_array(p, shape=(9,)) == -1)


def test_gridstates_bin_format(engine_lib, tmp_path, oracle):
    lib = engine_lib.lib()
    L, R = 7, 3  # 147 spins: the packed stream crosses grid and byte boundaries
    rs = np.random.default_rng(3)
    g1 = np.where(rs.random(R * L * L) < 0.4, 1, -1).astype(np.int32)
    g2 = -g1
    fn = str(tmp_path / "gridstates.bin").encode()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    assert lib.write_ising_grids(L, R, ip(g1), 0, None, None, b"magnetisation", -0.9, 0.9, fn) == 0
    assert lib.write_ising_grids(L, R, ip(g2), 100, None, None, b"magnetisation", -0.9, 0.9, fn) == 0
    raw = open(fn, "rb").read()
    nbytes = (R * L * L + 7) // 8
    assert len(raw) == 2 * (12 + nbytes)
    for k, (g, sweep) in enumerate([(g1, 0), (g2, 100)]):
        rec = raw[k * (12 + nbytes):(k + 1) * (12 + nbytes)]
        assert np.frombuffer(rec[:12], dtype=np.int32).tolist() == [L, R, sweep]
        body = np.frombuffer(rec[12:], dtype=np.uint8)
        assert np.array_equal(body, np.packbits((g == 1).astype(np.uint8), bitorder="little"))
        ref = np.zeros(nbytes, dtype=np.uint8)
        oracle.lib().orc_pack_bits(g, g.size, ref)
        assert np.array_equal(body, ref)


def test_gridinput_bin_reader(engine_lib, tmp_path, monkeypatch):
    lib = engine_lib.lib()
    L, R = 6, 4  # 36 spins: 4 padding bits in the last byte (set, to check they are ignored)
    rs = np.random.default_rng(5)
    g = np.where(rs.random(L * L) < 0.5, 1, -1).astype(np.int32)
    bits = np.packbits((g == 1).astype(np.uint8), bitorder="little")
    bits[-1] |= 0xF0
    monkeypatch.chdir(tmp_path)
    with open("gridinput.bin", "wb") as f:
        f.write(np.int32(L).tobytes() + bits.tobytes())
    p = lib.init_grids_from_file(L, R)
    assert p
    got = np.ctypeslib.as_array(p, shape=(R, L * L))
    assert all(np.array_equal(got[r], g) for r in range(R))
    assert not lib.init_grids_from_file(L + 1, R)  # size mismatch -> error, no exit
    assert lib.gasp_last_error() == 4
    os.remove("gridinput.bin")
    assert not lib.init_grids_from_file(L, R)
    assert lib.gasp_last_error() == 4


def test_bootstrap_matches_reference(engine_lib, oracle):
    gold = np.load(GOLD)
    lib = engine_lib.lib()
    sizes = gold["boot_sizes"]
    fates = np.split(gold["boot_fates"], np.cumsum(sizes)[:-1])
    oracle.lib().srand(1)
    got = [lib.bootstrap_errbar(len(f), len(f), np.ascontiguousarray(f).ctypes.data_as(C.POINTER(C.c_int)))
           for f in fates]
    assert np.array_equal(np.array(got).view(np.uint64), gold["boot_vals"].view(np.uint64))


@pytest.mark.parametrize("seed", [1, 7, 12345, 2**31 + 5])
def test_bootstrap_libc_batch_is_sequential_libc(engine_lib, seed):
    """gasp_bootstrap_errbar_libc (its own copy of glibc's rand() sequence,
    jumped ahead to each input, inputs on threads) equals successive
    bootstrap_errbar calls on the real libc rand() after srand(seed), bit for
    bit: inputs of 0/1/-1 fates (one all undecided: 0/0), 1 to 7 threads, so
    that threads start at offsets of several hundred thousand calls."""
    G = engine_lib
    lib = G.lib()
    libc = C.CDLL("libc.so.6")
    libc.srand.argtypes = [C.c_uint]
    rs = np.random.default_rng(seed % 1000)
    ninputs, n, m = 9, 333, 211
    f = rs.integers(-1, 2, size=(ninputs, n)).astype(np.int32)
    f[4] = -1
    libc.srand(seed)
    ref = np.array([lib.bootstrap_errbar(n, m, np.ascontiguousarray(f[i]).ctypes.data_as(C.POINTER(C.c_int)))
                    for i in range(ninputs)])
    for nthreads in (1, 2, 3, 7):
        got = G.bootstrap_libc(f, m, seed=seed, nthreads=nthreads)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(got[~np.isnan(ref)], ref[~np.isnan(ref)]), nthreads


def test_bootstrap_libc_edges(engine_lib):
    """No inputs; one fate per input (n = 1: every draw is index 0); more
    threads than inputs, or than the thread cap; an input with no decided
    fate gives NaN, as the reference's 0/0."""
    G = engine_lib
    assert G.bootstrap_libc(np.zeros((0, 5), dtype=np.int32), 3).shape == (0,)
    one = G.bootstrap_libc(np.array([[1], [0], [-1]], dtype=np.int32), 7, nthreads=16)
    assert one[0] == 0.0 and one[1] == 0.0 and np.isnan(one[2])
    f = np.array([[0, 1, 1, 0, 1]] * 3, dtype=np.int32)
    a = G.bootstrap_libc(f, 50, seed=3, nthreads=1)
    b = G.bootstrap_libc(f, 50, seed=3, nthreads=64)
    assert np.array_equal(a, b) and np.all(a > 0)
    # a thread count far above the cap of 256 threads (300 inputs)
    f = np.random.default_rng(5).integers(-1, 2, size=(300, 40), dtype=np.int32)
    a = G.bootstrap_libc(f, 20, seed=9, nthreads=1)
    b = G.bootstrap_libc(f, 20, seed=9, nthreads=100000)
    assert np.array_equal(a, b, equal_nan=True)


def test_threshold_table_layout():
    """Kernel table order j = 5*s01 + (up neighbours) is the reference index
    5(s+1) + nsum + 4 = 10 s01 + 2c (s = 2 s01 - 1, nsum = 2c - 4)."""
    for s01 in (0, 1):
        for c in range(5):
            s, nsum = 2 * s01 - 1, 2 * c - 4
            assert 5 * (s + 1) + nsum + 4 == 10 * s01 + 2 * c


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU behaviour")
def test_cli_without_gpu(engine_lib, tmp_path):
    cli = engine_lib.CLI_PATH
    assert os.path.exists(cli)
    r = subprocess.run([cli], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0 and "Usage : GPU_2DIsing nsweeps nreplicas" in r.stdout
    # without gridinput.bin an itask-1 run stops before any engine runs; the
    # old gridstates.bin is deleted at start all the same (src/ising.cu:84)
    (tmp_path / "gridstates.bin").write_bytes(b"old")
    r = subprocess.run([cli, "100", "64", "10", "10", "32", "0", "2", "0.54", "0.07", "1"],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0
    assert not (tmp_path / "gridstates.bin").exists()


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU behaviour")
def test_python_module_without_gpu(engine_lib):
    code = r"""
import sys, copy, numpy as np
sys.path.insert(0, sys.argv[1])
import gasp
assert gasp.gpu_nsms == 0 and gasp.grids == [] and gasp.magnetisation == [] and gasp.largest_cluster == []
s = gasp.GridSnapObject(3, 4, np.ones((5, 5), dtype=np.int8))
assert (s.isweep, s.igrid, s.magnetisation, s.lclus_size, s.committor) == (3, 4, 0.0, -1.0, -1.0)
s.magnetisation = 0.5
c = copy.deepcopy(s)
assert c.grid is not s.grid and c.magnetisation == 0.5 and np.array_equal(c.grid, s.grid)
for bad in (np.ones((5, 5), dtype=np.int32), np.ones(5, dtype=np.int8)):
    try:
        gasp.GridSnapObject(0, 0, bad); raise SystemExit("accepted a bad array")
    except TypeError:
        pass
try:
    gasp.run_nucleation_swarm(64, 128, 100, 0.54, 0.07, cv="energy"); raise SystemExit("bad cv")
except ValueError:
    pass
try:
    gasp.run_nucleation_swarm(64, 128, 100, 0.54, 0.07, max_keep_grids=-1); raise SystemExit("bad max")
except ValueError:
    pass
try:
    gasp.run_committor_calc(8, 64, 100, 0.54, 0.07, nsms=-1); raise SystemExit("bad nsms")
except ValueError:
    pass
try:
    gasp.run_committor_calc(8, 64, 100, 0.54, 0.07, grid_input="NumPy", grid_array=[], nsms=4)
    raise SystemExit("empty list")
except ValueError:
    pass
try:
    gasp.run_committor_calc(8, 64, 100, 0.54, 0.07, grid_input="NumPy",
                            grid_array=[np.ones((8, 8), dtype=np.int32)], nsms=4)
    raise SystemExit("int32 grid accepted")
except TypeError:
    pass
print("OK")
"""
    r = subprocess.run([sys.executable, "-c", code, engine_lib.PKG_DIR], capture_output=True, text=True)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout + r.stderr


def test_default_launch_shapes(engine_lib):
    """Launch shapes chosen without the tuner (256 CUs): method 0 runs a whole
    wave per replica from N >= 4096, and 2, 4 or 8 waves per replica (16 only from the tuner) when
    N >= 16384 and the replicas alone would not fill the device; method 2 keeps
    split words with 16 slots per replica at L=64 with 16384 replicas, splits
    further (16 or 8 replicas per workgroup) when halves would leave CUs
    without a workgroup, and never goes below 8 slots (method 1: 8).  An explicit 128-slot
    setting for method 0 implies two waves."""
    G = engine_lib
    G.set_tuning(0, 0)
    try:
        assert G.get_tuning(0, 64, 16384) == (64, 1)
        assert G.get_tuning(0, 64, 256) == (64, 1)
        assert G.get_tuning(0, 256, 256) == (512, 8)
        assert G.get_tuning(0, 256, 1024) == (256, 4)
        assert G.get_tuning(0, 256, 2048) == (128, 2)
        assert G.get_tuning(0, 256, 16384) == (64, 1)
        assert G.get_tuning(2, 64, 16384) == (16, 2)  # split words
        assert G.get_tuning(2, 64, 524288) == (8, 2)  # at least 8 slots with replicas to spare
        assert G.get_tuning(2, 160, 64) == (16, 8)  # one group over 8 workgroups (whole words would not fit)
        assert G.get_tuning(2, 96, 2048) == (16, 8)  # 32 groups: eighth words, 256 workgroups
        assert G.get_tuning(2, 64, 4096) == (16, 4)  # 64 groups: quarter words
        assert G.get_tuning(2, 64, 8192) == (16, 2)  # 128 groups: halves already fill the CUs
        # many replicas of a large lattice: the split grows until a workgroup's
        # lattice fits the LDS (64 KB assumed without a GPU: eighth words at L = 250)
        assert G.get_tuning(2, 250, 16384) == (16, 8)
        assert G.get_tuning(1, 64, 524288) == (8, 1)  # method 1: at least 8
        G.set_tuning(128, 0)
        assert G.get_tuning(0, 16, 96) == (128, 2)
        assert G.get_tuning(1, 16, 96)[0] == 64  # other methods stop at 64 slots
    finally:
        G.set_tuning(0, 0)
