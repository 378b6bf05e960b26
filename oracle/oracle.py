"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference
build (oracle/_ref/libgasp_ref.so, when present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgasp_ref.so")
REFERENCE_ROOT = "/root/reference"

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class MTState(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int)]


def build(with_ref: bool | None = None) -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if with_ref is None:
        with_ref = os.path.isdir(os.path.join(REFERENCE_ROOT, "src"))
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref", "ref-io"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build(with_ref=False)
    L = C.CDLL(LIB_PATH)
    L.orc_mt_seed.argtypes = [C.POINTER(MTState), C.c_uint32]
    L.orc_mt_u32.argtypes = [C.POINTER(MTState)]
    L.orc_mt_u32.restype = C.c_uint32
    L.orc_mt_real3.argtypes = [C.POINTER(MTState)]
    L.orc_mt_real3.restype = C.c_double
    L.orc_mt_fill.argtypes = [C.POINTER(MTState), _u32p, C.c_longlong]
    L.orc_pacc_cpu.argtypes = [C.c_double, C.c_double, _f64p]
    L.orc_pacc_gpu.argtypes = [C.c_double, C.c_double, _f32p]
    L.orc_threshold.argtypes = [C.c_double]
    L.orc_threshold.restype = C.c_uint64
    L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    L.orc_sweep_mt.argtypes = [C.c_int, _i32p, C.POINTER(MTState), _f64p, C.c_longlong]
    L.orc_sweep_philox.argtypes = [C.c_int, _i32p, C.c_uint64, C.c_uint32, C.c_uint64,
                                   C.c_longlong, _u64p]
    L.orc_magnetisation_cpu.argtypes = [C.c_int, _i32p]
    L.orc_magnetisation_cpu.restype = C.c_float
    L.orc_magnetisation_gpu.argtypes = [C.c_int, _i32p]
    L.orc_magnetisation_gpu.restype = C.c_float
    L.orc_largest_cluster.argtypes = [C.c_int, _i32p, C.c_int]
    L.orc_largest_cluster.restype = C.c_float
    L.orc_bootstrap_errbar.argtypes = [C.c_int, C.c_int, _i32p]
    L.orc_bootstrap_errbar.restype = C.c_double
    L.orc_pack_bits.argtypes = [_i32p, C.c_longlong, _u8p]
    L.orc_unpack_bits.argtypes = [_u8p, C.c_longlong, _i32p]
    L.orc_sweep_batch.argtypes = [C.c_int, C.c_int, _i32p, C.c_int, C.c_void_p, C.c_uint64,
                                  C.c_uint64, C.c_longlong, C.c_double, C.c_double, C.c_int]
    L.orc_sweep_batch.restype = C.c_double
    L.orc_sweep_sliced.argtypes = [C.c_int, _i32p, C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint64,
                                   C.c_longlong, _u64p, C.c_void_p]
    L.orc_sweep_sliced_batch.argtypes = [C.c_int, C.c_int, _i32p, C.c_int, C.c_uint64, C.c_uint64,
                                         C.c_longlong, C.c_double, C.c_double, C.c_void_p, C.c_int]
    L.orc_sweep_sliced_batch.restype = C.c_double
    L.orc_driver.argtypes = [C.c_int, C.c_int, _i32p, C.c_int, C.c_uint64, C.c_double, C.c_double,
                             C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                             C.c_double, C.c_double, C.c_int, C.c_void_p, _f32p,
                             C.c_void_p, C.c_void_p, C.c_int,
                             C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_int]
    L.orc_driver.restype = C.c_int
    L.srand = C.CDLL(None).srand
    L.srand.argtypes = [C.c_uint]
    _lib = L
    return L


def ref():
    """The reference's own C sources compiled by oracle/Makefile, or None."""
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_PATH):
        return None
    R = C.CDLL(REF_PATH)
    R.init_genrand.argtypes = [C.c_ulong]
    R.genrand_int32.restype = C.c_ulong
    R.genrand_real3.restype = C.c_double
    R.preComputeProbs_cpu.argtypes = [C.c_double, C.c_double]
    R.mc_sweep_cpu.argtypes = [C.c_int, _i32p, C.c_int, C.c_double, C.c_double, C.c_int]
    R.compute_magnetisation_cpu.argtypes = [C.c_int, _i32p, C.c_int, _f32p]
    R.compute_largest_cluster_cpu.argtypes = [C.c_int, _i32p, C.c_int, C.c_int, _f32p]
    R.bootstrap_errbar.argtypes = [C.c_int, C.c_int, _i32p]
    R.bootstrap_errbar.restype = C.c_double
    _ref = R
    return R


# ------------------------------------------------------------------ helpers
def mt_stream(seed: int, n: int) -> np.ndarray:
    s = MTState()
    lib().orc_mt_seed(C.byref(s), seed & 0xFFFFFFFF)
    out = np.empty(n, dtype=np.uint32)
    lib().orc_mt_fill(C.byref(s), out, n)
    return out


def pacc_cpu(beta: float, h: float) -> np.ndarray:
    out = np.zeros(20, dtype=np.float64)
    lib().orc_pacc_cpu(beta, h, out)
    return out


def pacc_gpu(beta: float, h: float) -> np.ndarray:
    out = np.zeros(20, dtype=np.float32)
    lib().orc_pacc_gpu(beta, h, out)
    return out


def thresholds(p: np.ndarray) -> np.ndarray:
    return np.array([lib().orc_threshold(float(x)) for x in p], dtype=np.uint64)


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def sweep_mt_per_replica(L: int, grids: np.ndarray, seed: int, nsweeps: int, beta: float,
                         h: float, nthreads: int = 8) -> np.ndarray:
    """Per-replica MT19937 (seed + r) sweeps of int32 grids [R, L, L], in place."""
    R = grids.shape[0]
    states = (MTState * R)()
    for r in range(R):
        lib().orc_mt_seed(C.byref(states[r]), (seed + r) & 0xFFFFFFFF)
    flat = np.ascontiguousarray(grids.reshape(R, -1), dtype=np.int32)
    lib().orc_sweep_batch(L, R, flat, 0, C.cast(states, C.c_void_p), 0, 0,
                          nsweeps * L * L, beta, h, nthreads)
    grids[...] = flat.reshape(grids.shape)
    return grids


def sweep_sliced(L: int, grids: np.ndarray, wb: int, seed: int, attempt0: int, nattempts: int,
                 beta: float, h: float, frozen=None, nthreads: int = 8) -> np.ndarray:
    """The bit-sliced multispin stream (engine gpu_method 3, wb = 64 or 32
    replicas per word group): every group advanced by nattempts attempts from
    attempt index attempt0, in place; frozen (R booleans or None) replicas do
    not change."""
    R = grids.shape[0]
    flat = np.ascontiguousarray(grids.reshape(R, -1), dtype=np.int32)
    fz = None
    if frozen is not None:
        fz = np.ascontiguousarray(np.asarray(frozen, dtype=np.uint8).reshape(R))
    lib().orc_sweep_sliced_batch(L, R, flat, wb, seed, attempt0, nattempts, beta, h,
                                 None if fz is None else fz.ctypes.data_as(C.c_void_p), nthreads)
    grids[...] = flat.reshape(grids.shape)
    return grids


def sweep_philox(L: int, grids: np.ndarray, seed: int, attempt0: int, nattempts: int,
                 beta: float, h: float, nthreads: int = 8) -> np.ndarray:
    R = grids.shape[0]
    flat = np.ascontiguousarray(grids.reshape(R, -1), dtype=np.int32)
    lib().orc_sweep_batch(L, R, flat, 1, None, seed, attempt0, nattempts, beta, h, nthreads)
    grids[...] = flat.reshape(grids.shape)
    return grids


def driver(L, R, grids, *, rng_kind, seed, beta, h, tot_nsweeps, mag_output_int,
           grid_output_int, itask=0, cv_kind=0, initial_spin=-1, dn_thr=-0.9, up_thr=0.9,
           ninputs=1, fates=None, max_dumps=0, nthreads=8):
    """Run the restated mc_driver_gpu schedule; returns a dict of outputs."""
    flat = np.ascontiguousarray(grids.reshape(R, L * L), dtype=np.int32)
    nres = tot_nsweeps // mag_output_int if itask == 0 else ninputs
    result = np.zeros(max(nres, 1), dtype=np.float32)
    if fates is None:
        fates = np.full(R, -1, dtype=np.int32)
    max_checks = tot_nsweeps // min(mag_output_int, grid_output_int) + 2
    cv = np.zeros((max_checks, R), dtype=np.float32)
    csw = np.zeros(max_checks, dtype=np.int32)
    dumps = np.zeros((max(max_dumps, 1), R, L * L), dtype=np.int32)
    dsw = np.zeros(max(max_dumps, 1), dtype=np.int32)
    nd = C.c_int(0)
    nchk = lib().orc_driver(L, R, flat, rng_kind, seed, beta, h, tot_nsweeps, mag_output_int,
                            grid_output_int, itask, cv_kind, initial_spin, dn_thr, up_thr,
                            ninputs, fates.ctypes.data, result, cv.ctypes.data, csw.ctypes.data,
                            max_checks, dumps.ctypes.data, dsw.ctypes.data, max_dumps,
                            C.byref(nd), nthreads)
    if nchk < 0:
        raise ValueError("orc_driver rejected its arguments")
    nd = min(nd.value, max_dumps)
    return {
        "grids": flat.reshape(R, L, L),
        "result": result[:nres],
        "fates": fates,
        "cv": cv[:min(nchk, max_checks)],
        "check_sweeps": csw[:min(nchk, max_checks)],
        "dumps": dumps[:nd],
        "dump_sweeps": dsw[:nd],
    }


def gpu_thresholds(beta: float, h: float) -> np.ndarray:
    """Integer thresholds (20 entries, reference index 5(s+1)+nsum+4) of the
    single-precision table src/mc_gpu.cu:21-37."""
    return thresholds(pacc_gpu(beta, h))


def sweep_philox_replicas(L: int, grids: np.ndarray, replicas, seed: int, attempt0: int, nattempts: int,
                          beta: float, h: float, nthreads: int = 8) -> np.ndarray:
    """Production-stream sweeps of selected replicas: grids[k] (int32 [L, L])
    is advanced as replica replicas[k] of an engine keyed by `seed`, from
    attempt index attempt0, in place (threads over the replicas)."""
    from concurrent.futures import ThreadPoolExecutor

    thr = np.ascontiguousarray(gpu_thresholds(beta, h), dtype=np.uint64)
    flat = np.ascontiguousarray(grids.reshape(len(replicas), L * L), dtype=np.int32)
    rows = [np.ascontiguousarray(flat[k]) for k in range(len(replicas))]

    def one(k):
        lib().orc_sweep_philox(L, rows[k], seed, int(replicas[k]), attempt0, nattempts, thr)

    with ThreadPoolExecutor(max_workers=max(1, nthreads)) as ex:
        list(ex.map(one, range(len(replicas))))
    out = np.stack(rows).reshape(grids.shape)
    grids[...] = out
    return grids


def magnetisation_gpu(grids: np.ndarray) -> np.ndarray:
    """(float)sum / (float)N per grid (src/mc_gpu.cu:445-449), grids [R, L, L]."""
    R = grids.shape[0]
    s = grids.reshape(R, -1).sum(1)
    return (s.astype(np.float32) / np.float32(grids[0].size)).astype(np.float32)


def shoot_replicas(L: int, grids: np.ndarray, replicas, seed: int, beta: float, h: float, chk: int,
                   dn: float, up: float, max_sweeps: int, nthreads: int = 8):
    """Committor shooting of selected replicas with decided replicas frozen:
    check the start state, then sweep chk sweeps at a time (attempt indices
    from 0, one counter for the whole engine) and check again, until the fate
    is decided (1: m > up, 0: m < dn, sticky) or max_sweeps.  Returns
    (fates, grids advanced in place, sweeps swept per replica)."""
    n = len(replicas)
    fates = np.full(n, -1, dtype=np.int32)
    swept = np.zeros(n, dtype=np.int64)
    N = L * L

    def decide(idx):
        m = magnetisation_gpu(grids[idx]).astype(np.float64)
        fates[idx] = np.where(m > up, 1, np.where(m < dn, 0, -1))

    decide(np.arange(n))
    done = 0
    while done < max_sweeps:
        act = np.flatnonzero(fates == -1)
        if act.size == 0:
            break
        g = grids[act].copy()
        sweep_philox_replicas(L, g, [replicas[i] for i in act], seed, done * N, chk * N, beta, h, nthreads)
        grids[act] = g
        swept[act] += chk
        done += chk
        decide(act)
    return fates, grids, swept
