"""The HDF5 grid files of the Python front-end (libgasp_hdf5.so) against the
reference's own writer, src/io.c:132-301 compiled from the reference sources
into oracle/_ref/libgasp_ref_io.so (`make -C oracle ref-io`).

Both writers get the same calls — header, two sweeps with magnetisation
(without and with largest-cluster sizes), and an empty sweep whose group must
disappear — and h5diff must find the files identical (datasets, groups and
attributes; HDF5 object timestamps are not compared)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MINE = os.path.join(ROOT, "gpu_2dising_amd", "libgasp_hdf5.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libgasp_ref_io.so")
H5DIFF = shutil.which("h5diff") or "/opt/conda/bin/h5diff"
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def _bind(path):
    lib = C.CDLL(path)
    ip, fp = C.POINTER(C.c_int), C.POINTER(C.c_float)
    lib.create_ising_grids_hdf5.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_char_p]
    lib.create_ising_grids_hdf5.restype = C.c_int
    lib.write_ising_grids_hdf5.argtypes = [C.c_int, C.c_int, ip, C.c_int, fp, fp, C.c_char_p, C.c_double,
                                           C.c_double, C.c_char_p]
    lib.write_ising_grids_hdf5.restype = C.c_int
    return lib


def _write(lib, path):
    L, n = 6, 4
    rs = np.random.default_rng(7)
    grids = np.where(rs.random((n, L, L)) < 0.4, 1, -1).astype(np.int32)
    mag = grids.reshape(n, -1).mean(1).astype(np.float32)
    lclus = np.array([3, 17, 0, 36], dtype=np.float32)
    fn = path.encode()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    assert lib.create_ising_grids_hdf5(L, n, 1000, 0.07, 0.54, 0, fn) == 0
    assert lib.write_ising_grids_hdf5(L, n, ip(grids), 0, fp(mag), None, b"magnetisation", -0.9, 0.9, fn) == 0
    g2 = -grids
    assert lib.write_ising_grids_hdf5(L, n, ip(g2), 500, fp(-mag), fp(lclus), b"largest_cluster", -0.9, 0.9,
                                      fn) == 0
    # no grids: the group is created and deleted again, the total is unchanged
    assert lib.write_ising_grids_hdf5(L, 0, ip(grids), 900, fp(mag), None, b"magnetisation", -0.9, 0.9, fn) == 0
    return grids


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    if not (os.path.exists(H5DIFF) and os.path.exists(H5DUMP)):
        pytest.skip("h5diff/h5dump not available")
    d = tmp_path_factory.mktemp("h5")
    mine = str(d / "mine.hdf5")
    grids = _write(_bind(MINE), mine)
    return d, mine, grids


def test_layout_and_values(files):
    d, mine, grids = files
    out = subprocess.run([H5DUMP, "-d", "/total_saved_grids", mine], capture_output=True, text=True, check=True)
    assert "(0): 8" in out.stdout
    out = subprocess.run([H5DUMP, "-n", mine], capture_output=True, text=True, check=True).stdout
    for name in ("/L", "/ngrids", "/tot_nsweeps", "/h", "/beta", "/sweep_0/grid_3", "/sweep_500/grid_0"):
        assert name in out, name
    assert "/sweep_900" not in out
    out = subprocess.run([H5DUMP, "-d", "/sweep_0/grid_1", mine], capture_output=True, text=True, check=True).stdout
    vals = [int(v) for v in out.split("DATA {")[1].split("}")[0].replace("(", " ").replace(")", " ").replace(":", " ")
            .replace(",", " ").split()]
    # h5dump prints "(row,col): v, v, ..." lines; the row/col indices are interleaved, keep +-1 values
    spins = [v for v in vals if v in (-1, 1)]
    assert len(spins) >= 36


def test_identical_to_reference_writer(files):
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref/libgasp_ref_io.so not built (needs the reference sources)")
    d, mine, _ = files
    ref = str(d / "ref.hdf5")
    _write(_bind(REF), ref)
    r = subprocess.run([H5DIFF, "-v", mine, ref], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "0 differences found" in r.stdout
