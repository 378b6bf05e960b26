#!/usr/bin/env python3
"""Writes tests/golden/golden_cpu_driver.npz: runs of the reference's own CPU
driver mc_driver_cpu (src/mc_cpu.c:106-308, one global MT19937 stream seeded
with init_genrand), compiled from its sources by `make -C oracle ref`
(oracle/_ref/libgasp_ref.so).

Each case records, as the reference produced them:
  <case>_result      the result array (nucleated fraction or pB per input)
  <case>_fates       grid_fate after the run (itask 1)
  <case>_final       the grids after the run, bit-packed (LSB first)
  <case>_dump_sweeps isweep of every output-callback call
  <case>_dumps       the grids handed to each callback call, bit-packed
  <case>_mags        the magnetisation array handed to each call
  <case>_lclus       the cluster-size array handed to each call (cluster CV)
  <case>_stdout      what the driver printed (minus the "Time taken" line)
  <case>_meta        L, R, tot, mag, grid, itask, cluster, initial_spin, ninputs, seed
  <case>_thr         dn, up, beta, h
  <case>_start       the start grids, bit-packed

Run from the repository root: python tests/golden/make_golden_cpu_driver.py
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_cpu_driver.npz")
SEED = 2894203475  # src/ising.cu:55 (commented development seed)

# name, L, R, tot, mag, grid, itask, cluster, initial_spin, ninputs, dn, up, beta, h, start
CASES = [
    ("nuc", 8, 6, 600, 50, 100, 0, 0, -1, 1, -0.9, 0.5, 0.5, 0.25, "down"),
    # the command line program's shape (GASP_L=8): checks every 50, grids every 100, +-0.90
    ("cli", 8, 4, 300, 50, 100, 0, 0, -1, 1, -0.9, 0.9, 0.45, 0.3, "down"),
    ("clus", 8, 5, 300, 30, 60, 0, 1, -1, 1, -0.9, 30.0, 0.5, 0.1, "down"),
    ("comm", 8, 8, 3000, 10, 50, 1, 0, -1, 2, -0.6, 0.6, 0.54, 0.07, "inputs"),
    ("commclus", 6, 6, 2000, 20, 40, 1, 1, -1, 3, 4.0, 20.0, 0.54, 0.07, "inputs"),
    # checks fall between launches (spc = grid = 30 < mag = 40) and tot % spc != 0
    ("uneven", 5, 3, 230, 40, 30, 0, 0, -1, 1, -0.9, 0.9, 0.44, 0.2, "down"),
    # everything is above the threshold at the first check: the run stops there
    ("stop", 4, 3, 100, 10, 10, 0, 0, -1, 1, -0.9, -1.5, 0.54, 0.07, "down"),
]

CB = C.CFUNCTYPE(C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_float),
                 C.POINTER(C.c_float), C.c_char_p, C.c_double, C.c_double, C.c_char_p)


class McGrids(C.Structure):
    _fields_ = [("L", C.c_int), ("ngrids", C.c_int), ("ising_grids", C.POINTER(C.c_int))]


class McSampler(C.Structure):
    _fields_ = [("tot_nsweeps", C.c_int), ("mag_output_int", C.c_int), ("grid_output_int", C.c_int)]


class McFunction(C.Structure):
    _fields_ = [("itask", C.c_int), ("dn_thr", C.c_double), ("up_thr", C.c_double),
                ("ninputs", C.c_int), ("initial_spin", C.c_int), ("cv", C.c_char_p),
                ("result", C.POINTER(C.c_float)), ("filename", C.c_char_p)]


def packbits(g):
    return np.packbits((np.asarray(g).reshape(-1) == 1).astype(np.uint8), bitorder="little")


def start_grids(kind, L, R, ninputs, rs):
    if kind == "down":
        return -np.ones((R, L * L), dtype=np.int32)
    ins = np.where(rs.random((ninputs, L * L)) < 0.45, 1, -1).astype(np.int32)
    return np.repeat(ins, R // ninputs, axis=0)


def run_case(R_lib, libc, case, rs):
    name, L, R, tot, mag, gout, itask, clus, ispin, nin, dn, up, beta, h, kind = case
    g0 = start_grids(kind, L, R, nin, rs)
    grids = np.ascontiguousarray(g0.copy())
    fates = np.full(R, -1, dtype=np.int32)
    nres = tot // mag if itask == 0 else nin
    result = np.zeros(max(nres, 1), dtype=np.float32)
    dumps, sweeps, mags, lcl = [], [], [], []

    def cb(L_, n_, data, isweep, mg, lc, cv, d, u, fn):
        a = np.ctypeslib.as_array(data, shape=(n_ * L_ * L_,)).copy()
        dumps.append(packbits(a))
        sweeps.append(isweep)
        mags.append(np.ctypeslib.as_array(mg, shape=(n_,)).copy())
        if lc:
            lcl.append(np.ctypeslib.as_array(lc, shape=(n_,)).copy())
        return 0

    cbf = CB(cb)
    R_lib.init_genrand(SEED)
    R_lib.preComputeProbs_cpu(beta, h)
    gr = McGrids(L, R, grids.ctypes.data_as(C.POINTER(C.c_int)))
    sm = McSampler(tot, mag, gout)
    fc = McFunction(itask, dn, up, nin, ispin, b"largest_cluster" if clus else b"magnetisation",
                    result.ctypes.data_as(C.POINTER(C.c_float)), b"unused.bin")
    # capture what the driver prints on the C stdout
    libc.fflush(None)
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tf:
        os.dup2(tf.fileno(), 1)
        try:
            R_lib.mc_driver_cpu(gr, C.c_double(beta), C.c_double(h),
                                fates.ctypes.data_as(C.POINTER(C.c_int)) if itask == 1 else None,
                                sm, fc, cbf)
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tf.seek(0)
        text = tf.read().decode()
    # the timing line varies; the itask-1 "pB estimate" line prints a pointer (a reference defect)
    keep = [ln for ln in text.split("\n") if not ln.startswith("# Time taken") and not ln.startswith("pB estimate")]
    out = {
        f"{name}_result": result[:nres],
        f"{name}_fates": fates,
        f"{name}_final": packbits(grids),
        f"{name}_start": packbits(g0),
        f"{name}_dump_sweeps": np.array(sweeps, dtype=np.int32),
        f"{name}_dumps": np.stack(dumps) if dumps else np.zeros((0, 1), np.uint8),
        f"{name}_mags": np.stack(mags) if mags else np.zeros((0, R), np.float32),
        f"{name}_lclus": np.stack(lcl) if lcl else np.zeros((0, R), np.float32),
        f"{name}_stdout": np.array("\n".join(keep)),
        f"{name}_meta": np.array([L, R, tot, mag, gout, itask, clus, ispin, nin, SEED], dtype=np.int64),
        f"{name}_thr": np.array([dn, up, beta, h], dtype=np.float64),
    }
    return out


def main():
    O.build(with_ref=True)
    R_lib = O.ref()
    if R_lib is None:
        raise SystemExit("oracle/_ref/libgasp_ref.so missing (needs the reference tree)")
    R_lib.mc_driver_cpu.argtypes = [McGrids, C.c_double, C.c_double, C.POINTER(C.c_int), McSampler,
                                    McFunction, CB]
    libc = C.CDLL(None)
    libc.fflush.argtypes = [C.c_void_p]
    rs = np.random.default_rng(777)
    out = {"cases": np.array([c[0] for c in CASES])}
    for case in CASES:
        out.update(run_case(R_lib, libc, case, rs))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
