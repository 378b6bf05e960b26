"""gridstates.bin -> gridinput.bin (gpu_2dising_amd.gridfiles) against the
reference's own reader and writer (src/io.c compiled into
oracle/_ref/libgasp_ref_io.so) and the engine library's host functions.

The reference viewer extracts a snapshot only when L*L is a multiple of 8
(tools/vis_gridstates.py:44-51); L = 17 checks the bit-stream layout where
grids are not byte-aligned."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_IO = os.path.join(ROOT, "oracle", "_ref", "libgasp_ref_io.so")


def _bind_io(lib):
    ip, fp = C.POINTER(C.c_int), C.POINTER(C.c_float)
    lib.write_ising_grids.argtypes = [C.c_int, C.c_int, ip, C.c_int, fp, fp, C.c_char_p, C.c_double, C.c_double,
                                      C.c_char_p]
    lib.write_ising_grids.restype = C.c_int
    lib.read_input_grid.argtypes = [C.c_int, C.c_int, ip]
    return lib


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _grids(L, n, seed):
    rs = np.random.default_rng(seed)
    return np.where(rs.random((n, L, L)) < 0.45, 1, -1).astype(np.int32)


@pytest.mark.parametrize("L", [8, 17])
def test_roundtrip_through_reference_io(engine_lib, tmp_path, monkeypatch, L):
    from gpu_2dising_amd import gridfiles as GF

    if not os.path.exists(REF_IO):
        pytest.skip("oracle/_ref/libgasp_ref_io.so not built (needs the reference sources)")
    monkeypatch.chdir(tmp_path)
    ref = _bind_io(C.CDLL(REF_IO))
    mine = _bind_io(engine_lib.lib())
    n = 5
    recs = [(0, _grids(L, n, 1)), (100, _grids(L, n, 2)), (200, _grids(L, n, 3))]
    for lib, name in ((ref, "ref.bin"), (mine, "mine.bin")):
        for isweep, g in recs:
            assert lib.write_ising_grids(L, n, _ip(g), isweep, None, None, b"magnetisation", -0.9, 0.9,
                                         name.encode()) == 0
    assert open("ref.bin", "rb").read() == open("mine.bin", "rb").read()
    got = GF.read_gridstates("ref.bin")
    assert [s for s, _ in got] == [s for s, _ in recs]
    for (_, g), (_, exp) in zip(got, recs):
        assert g.dtype == np.int8 and np.array_equal(g, exp)
    # one replica of one snapshot -> gridinput.bin, read by both readers
    grid = GF.extract("ref.bin", 1, 3, "gridinput.bin")
    assert np.array_equal(grid, recs[1][1][3])
    want = recs[1][1][3]
    out = np.zeros((4, L, L), dtype=np.int32)
    mine.read_input_grid(L, 4, _ip(out))
    assert np.array_equal(out, np.broadcast_to(want, (4, L, L)))
    # The reference reader also stores the padding bits of the last byte
    # (src/io.c:51-58): for L*L not a multiple of 8 they land on the first
    # sites of the next grid and past the end of the array.  Give it slack and
    # compare the grid it leaves intact (all of them when L*L % 8 == 0).
    slack = np.zeros(4 * L * L + 8, dtype=np.int32)
    ref.read_input_grid(L, 4, _ip(slack))
    got = slack[:4 * L * L].reshape(4, L, L)
    intact = 4 if (L * L) % 8 == 0 else 1
    assert np.array_equal(got[:intact], np.broadcast_to(want, (intact, L, L)))


def test_extract_errors(tmp_path):
    from gpu_2dising_amd import gridfiles as GF

    p = tmp_path / "g.bin"
    p.write_bytes(np.array([4, 2, 0], dtype=np.int32).tobytes() + bytes(4))
    assert len(GF.read_gridstates(str(p))) == 1
    with pytest.raises(IndexError):
        GF.extract(str(p), 1, 0, str(tmp_path / "x.bin"))
    with pytest.raises(IndexError):
        GF.extract(str(p), 0, 2, str(tmp_path / "x.bin"))
    p.write_bytes(np.array([4, 2, 0], dtype=np.int32).tobytes() + bytes(3))
    with pytest.raises(ValueError):
        GF.read_gridstates(str(p))
