"""The CPU fallback of the drop-in (no GPU visible): the product's mc_driver_cpu
and its helpers (gpu_2dising_amd/csrc/gasp_cpu.cpp) against runs of the
reference's own mc_driver_cpu (src/mc_cpu.c:106-308) compiled from its sources
(tests/golden/golden_cpu_driver.npz, written by make_golden_cpu_driver.py),
and the command line program and the `gasp` module falling back to it the way
src/ising.cu:127-155 and src/ising_ext.cu:1179-1187 do.

Compared exactly: the result array, the fates, every output-callback call
(sweep, grids, magnetisation / cluster sizes), the final grids and the
progress text."""
import ctypes as C
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "golden_cpu_driver.npz")


def _gpu_visible():
    try:
        hip = C.CDLL("libamdhip64.so")
        n = C.c_int(0)
        return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _bits(packed, n):
    return np.where(np.unpackbits(packed, bitorder="little")[:n] == 1, 1, -1).astype(np.int32)


def _case(gold, name):
    L, R, tot, mag, gout, itask, clus, ispin, nin, seed = (int(x) for x in gold[f"{name}_meta"])
    dn, up, beta, h = (float(x) for x in gold[f"{name}_thr"])
    return dict(L=L, R=R, tot=tot, mag=mag, gout=gout, itask=itask, clus=clus, ispin=ispin, nin=nin, seed=seed,
                dn=dn, up=up, beta=beta, h=h)


def _run_product(G, c, start):
    """The product's mc_driver_cpu through the C ABI, stdout captured."""
    lib = G.lib()
    lib.init_genrand.argtypes = [C.c_ulong]
    lib.preComputeProbs_cpu.argtypes = [C.c_double, C.c_double]
    lib.mc_driver_cpu.argtypes = [G.McGrids, C.c_double, C.c_double, C.POINTER(C.c_int), G.McSampler,
                                  G.McFunction, G.GridOutputFunc]
    L, R = c["L"], c["R"]
    grids = np.ascontiguousarray(start.copy())
    fates = np.full(R, -1, dtype=np.int32)
    nres = c["tot"] // c["mag"] if c["itask"] == 0 else c["nin"]
    result = np.zeros(max(nres, 1), dtype=np.float32)
    calls = []

    def cb(L_, n_, data, isweep, mg, lc, cv, d, u, fn):
        calls.append((isweep, np.ctypeslib.as_array(data, shape=(n_ * L_ * L_,)).copy(),
                      np.ctypeslib.as_array(mg, shape=(n_,)).copy(),
                      np.ctypeslib.as_array(lc, shape=(n_,)).copy() if lc else None))
        return 0

    cbf = G.GridOutputFunc(cb)
    lib.init_genrand(c["seed"])
    lib.preComputeProbs_cpu(c["beta"], c["h"])
    gr = G.McGrids(L, R, grids.ctypes.data_as(C.POINTER(C.c_int)))
    sm = G.McSampler(c["tot"], c["mag"], c["gout"])
    fc = G.McFunction(c["itask"], c["dn"], c["up"], c["nin"], c["ispin"],
                      b"largest_cluster" if c["clus"] else b"magnetisation",
                      result.ctypes.data_as(C.POINTER(C.c_float)), b"unused.bin")
    libc = C.CDLL(None)
    libc.fflush.argtypes = [C.c_void_p]
    libc.fflush(None)
    sys.stdout.flush()
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tf:
        os.dup2(tf.fileno(), 1)
        try:
            lib.mc_driver_cpu(gr, c["beta"], c["h"], fates.ctypes.data_as(C.POINTER(C.c_int)) if c["itask"] else None,
                              sm, fc, cbf)
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tf.seek(0)
        text = tf.read().decode()
    G.check()
    keep = [ln for ln in text.split("\n") if not ln.startswith("# Time taken") and not ln.startswith("pB estimate")]
    return result[:nres], fates, grids, calls, "\n".join(keep)


@pytest.mark.parametrize("name", ["nuc", "cli", "clus", "comm", "commclus", "uneven", "stop"])
def test_mc_driver_cpu_matches_reference(engine_lib, gold, name):
    c = _case(gold, name)
    L, R = c["L"], c["R"]
    start = _bits(gold[f"{name}_start"], R * L * L)
    result, fates, grids, calls, text = _run_product(engine_lib, c, start)
    assert np.array_equal(result, gold[f"{name}_result"])
    if c["itask"] == 1:
        assert np.array_equal(fates, gold[f"{name}_fates"])
    assert np.array_equal(grids, _bits(gold[f"{name}_final"], R * L * L))
    sweeps = gold[f"{name}_dump_sweeps"]
    assert [k[0] for k in calls] == list(sweeps)
    for i, (isweep, g, m, lc) in enumerate(calls):
        assert np.array_equal(g, _bits(gold[f"{name}_dumps"][i], R * L * L))
        assert np.array_equal(m, gold[f"{name}_mags"][i])
        if c["clus"]:
            assert np.array_equal(lc, gold[f"{name}_lclus"][i])
        else:
            assert lc is None
    assert text == str(gold[f"{name}_stdout"])


def test_cpu_stream_is_mt19937(engine_lib):
    """init_genrand / genrand_int32 / genrand_real3 of the product: the reference
    generator's first outputs and its 100000th (src/mt19937ar.c, compiled)."""
    ref = np.load(os.path.join(ROOT, "tests", "golden", "golden_ref.npz"))
    lib = engine_lib.lib()
    lib.init_genrand.argtypes = [C.c_ulong]
    lib.genrand_int32.restype = C.c_ulong
    lib.genrand_real3.restype = C.c_double
    for seed in (5489, 2894203475, 1, 4357):
        lib.init_genrand(seed)
        v = np.array([lib.genrand_int32() for _ in range(2000)], dtype=np.uint32)
        assert np.array_equal(v, ref[f"mt_{seed}"])
        for _ in range(100000 - 2000 - 1):
            lib.genrand_int32()
        assert lib.genrand_int32() == int(ref[f"mt_{seed}_100000th"][0])
    lib.init_genrand(5489)
    u = lib.genrand_int32()
    lib.init_genrand(5489)
    assert lib.genrand_real3() == (u + 0.5) / 2.0 ** 32


def test_cpu_sweep_matches_reference(engine_lib):
    """mc_sweep_cpu of the product against the reference's on the golden lattices
    (init_genrand(base + r) per replica, src/mc_cpu.c:41-89)."""
    ref = np.load(os.path.join(ROOT, "tests", "golden", "golden_ref.npz"))
    lib = engine_lib.lib()
    lib.init_genrand.argtypes = [C.c_ulong]
    lib.preComputeProbs_cpu.argtypes = [C.c_double, C.c_double]
    lib.mc_sweep_cpu.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.c_double, C.c_double, C.c_int]
    for L in (4, 5, 8, 16, 17, 32, 48, 64):
        _, nrep, nsweeps, base = (int(x) for x in ref[f"sweep_L{L}_meta"])
        g = _bits(ref[f"sweep_L{L}_start"], nrep * L * L).reshape(nrep, L * L)
        for r in range(nrep):
            lib.init_genrand(base + r)
            lib.preComputeProbs_cpu(0.54, 0.07)
            row = np.ascontiguousarray(g[r])
            lib.mc_sweep_cpu(L, row.ctypes.data_as(C.POINTER(C.c_int)), 0, 0.54, 0.07, nsweeps)
            g[r] = row
        assert np.array_equal(g.reshape(-1), _bits(ref[f"sweep_L{L}_final"], nrep * L * L)), L


def test_cpu_cluster_and_magnetisation(engine_lib):
    ref = np.load(os.path.join(ROOT, "tests", "golden", "golden_ref.npz"))
    lib = engine_lib.lib()
    lib.compute_largest_cluster_cpu.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int,
                                                C.POINTER(C.c_float)]
    lib.compute_magnetisation_cpu.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_float)]
    off = 0
    for k, L in enumerate(ref["clus_Ls"]):
        L = int(L)
        nb = (L * L + 7) // 8
        g = np.ascontiguousarray(_bits(ref["clus_bits"][off:off + nb], L * L))
        off += nb
        o = np.zeros(1, dtype=np.float32)
        lib.compute_largest_cluster_cpu(L, g.ctypes.data_as(C.POINTER(C.c_int)), 0, 1,
                                        o.ctypes.data_as(C.POINTER(C.c_float)))
        assert o[0] == ref["clus_up"][k]
        lib.compute_largest_cluster_cpu(L, g.ctypes.data_as(C.POINTER(C.c_int)), 0, -1,
                                        o.ctypes.data_as(C.POINTER(C.c_float)))
        assert o[0] == ref["clus_dn"][k]
    off = 0
    for k, L in enumerate(ref["mag_Ls"]):
        L = int(L)
        nb = (L * L + 7) // 8
        g = np.ascontiguousarray(_bits(ref["mag_bits"][off:off + nb], L * L))
        off += nb
        o = np.zeros(1, dtype=np.float32)
        lib.compute_magnetisation_cpu(L, g.ctypes.data_as(C.POINTER(C.c_int)), 0,
                                      o.ctypes.data_as(C.POINTER(C.c_float)))
        assert o[0] == ref["mag_vals"][k]


def _gridstates_bytes(gold, name):
    c = _case(gold, name)
    L, R = c["L"], c["R"]
    out = b""
    for i, isweep in enumerate(gold[f"{name}_dump_sweeps"]):
        out += struct.pack("<iii", L, R, int(isweep)) + gold[f"{name}_dumps"][i].tobytes()
    return out


@pytest.mark.skipif(_gpu_visible(), reason="forced fallback: needs a host without a visible GPU")
def test_cli_falls_back_to_cpu(engine_lib, gold, tmp_path):
    """`gasp` without a GPU (src/ising.cu:127-155): the reference's CPU engine with
    one MT19937 stream seeded with the seed; gridstates.bin and the report lines
    equal the compiled reference's mc_driver_cpu with write_ising_grids."""
    c = _case(gold, "cli")
    env = dict(os.environ, GASP_SEED=str(c["seed"]), GASP_L=str(c["L"]), HIP_VISIBLE_DEVICES="")
    args = [engine_lib.CLI_PATH, str(c["tot"]), str(c["R"]), str(c["mag"]), str(c["gout"]), "32", "0", "2",
            repr(c["beta"]), repr(c["h"]), "0"]
    (tmp_path / "gridstates.bin").write_bytes(b"stale")
    r = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("# isweep    nucleated fraction\n")
    assert "Falling back to CPU\n" in r.stdout
    lines = r.stdout.split("Falling back to CPU\n", 1)[1]
    keep = [ln for ln in lines.split("\n") if not ln.startswith("# Time taken")]
    assert "\n".join(keep) == str(gold["cli_stdout"])
    assert "# Time taken on CPU = " in r.stdout
    assert (tmp_path / "gridstates.bin").read_bytes() == _gridstates_bytes(gold, "cli")


@pytest.mark.skipif(_gpu_visible(), reason="forced fallback: needs a host without a visible GPU")
def test_python_module_falls_back_to_cpu(engine_lib, gold, tmp_path):
    """`import gasp` without a GPU prints "Falling back to CPU" and both run
    functions use the reference's CPU engine (src/ising_ext.cu:637-664,
    1007-1030): the nucleated fractions, the snapshot magnetisations and the pB
    values equal the compiled reference's mc_driver_cpu on the same seed."""
    c = _case(gold, "cli")
    k = _case(gold, "comm")
    starts = _bits(gold["comm_start"], k["R"] * k["L"] * k["L"]).reshape(k["R"], k["L"], k["L"])
    per = k["R"] // k["nin"]
    inputs = [starts[i * per].astype(np.int8) for i in range(k["nin"])]
    np.save(tmp_path / "inputs.npy", np.stack(inputs))
    code = f"""
import sys, numpy as np
sys.path.insert(0, {engine_lib.PKG_DIR!r})
import gasp
assert gasp.gpu_nsms == 0
res = gasp.run_nucleation_swarm({c['L']}, {c['R']}, {c['tot']}, {c['beta']!r}, {c['h']!r}, up_threshold={c['up']!r},
                                dn_threshold={c['dn']!r}, mag_output_int={c['mag']}, grid_output_int={c['gout']},
                                seed={c['seed']}, write_hdf5=False)
np.save("swarm.npy", res)
np.save("swarm_mags.npy", np.array(gasp.magnetisation, dtype=np.float32))
ins = list(np.load("inputs.npy"))
out = gasp.run_committor_calc({k['L']}, {k['R']}, {k['tot']}, {k['beta']!r}, {k['h']!r}, up_threshold={k['up']!r},
                              dn_threshold={k['dn']!r}, mag_output_int={k['mag']}, grid_output_int={k['gout']},
                              grid_input="NumPy", grid_array=ins, nsms={k['nin']}, seed={k['seed']},
                              write_hdf5=False)
np.save("comm.npy", np.array([p for p, e in out], dtype=np.float32))
print("OK")
"""
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=tmp_path, env=env)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout + r.stderr
    assert r.stdout.startswith("Falling back to CPU\n") or "\nFalling back to CPU\n" in r.stdout
    assert np.array_equal(np.load(tmp_path / "swarm.npy"), gold["cli_result"])
    mags = np.load(tmp_path / "swarm_mags.npy")
    assert np.array_equal(mags.reshape(-1), gold["cli_mags"].reshape(-1))
    assert np.array_equal(np.load(tmp_path / "comm.npy"), gold["comm_result"])
