"""CPU baseline with the reference's own code: src/mc_cpu.c's mc_sweep_cpu,
compiled from the reference sources into oracle/_ref/libgasp_ref.so.

TEST / BASELINE INFRASTRUCTURE ONLY (run by bench.py's cpu_baseline leg in a
child process that never touches the GPU).  The reference keeps one global
MT19937 state (src/mt19937ar.c:56-57), so its sweep cannot run on threads;
here every core gets its own process (fork), its own replicas and its own
init_genrand seed, all processes start together, and the attempt rate is the
total attempts over the slowest process's time.

  python3 -m oracle.refbench --cores 16 --seconds 12 --L 64
prints one JSON object.  --driver instead times the reference's literal
single-stream schedule, src/mc_cpu.c:106-308 mc_driver_cpu (one MT19937 stream
over all replicas, checks every 100 sweeps), on the plain run's shape (L = 32, 256
replicas, beta = 0.54, h = 0.07, all -1) for a bounded number of sweeps.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def _worker(i, L, nrep, nsweeps, beta, h, start, q):
    from oracle import oracle as O

    R = O.ref()
    g = -np.ones(nrep * L * L, dtype=np.int32)
    R.init_genrand(2894203475 + i)
    R.preComputeProbs_cpu(beta, h)
    start.wait()
    t0 = time.perf_counter()
    for r in range(nrep):
        R.mc_sweep_cpu(L, g, r, beta, h, nsweeps)
    q.put((nrep * nsweeps * L * L, time.perf_counter() - t0))


def run(cores: int, seconds: float, L: int = 64, beta: float = 0.54, h: float = 0.07, nrep: int = 4) -> dict:
    from oracle import oracle as O

    R = O.ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libgasp_ref.so is not built")
    # calibrate on one core: seconds per sweep of one replica
    g = -np.ones(L * L, dtype=np.int32)
    R.init_genrand(1)
    R.preComputeProbs_cpu(beta, h)
    t0 = time.perf_counter()
    R.mc_sweep_cpu(L, g, 0, beta, h, 20)
    per_sweep = (time.perf_counter() - t0) / 20
    nsweeps = int(max(1, min(10 ** 6, seconds / max(per_sweep * nrep, 1e-9))))
    ctx = mp.get_context("fork")
    start, q = ctx.Barrier(cores), ctx.Queue()
    procs = [ctx.Process(target=_worker, args=(i, L, nrep, nsweeps, beta, h, start, q)) for i in range(cores)]
    for p in procs:
        p.start()
    res = [q.get() for _ in procs]
    for p in procs:
        p.join()
    attempts = sum(a for a, _ in res)
    dt = max(t for _, t in res)
    return {"value": attempts / dt, "unit": "attempts/s", "cores": cores, "kind": "reference",
            "sample": f"src/mc_cpu.c mc_sweep_cpu (compiled reference), {cores} processes x {nrep} replicas x "
                      f"{nsweeps} sweeps, L={L}, beta={beta}, h={h}, all -1 start, {dt:.1f} s"}


def _driver_types(R):
    """ctypes signature of src/mc_cpu.c mc_driver_cpu on the compiled reference."""
    import ctypes as C

    class McGrids(C.Structure):  # include/mc_cpu.h:5-9
        _fields_ = [("L", C.c_int), ("ngrids", C.c_int), ("ising_grids", C.POINTER(C.c_int))]

    class McSampler(C.Structure):  # include/mc_cpu.h:11-15
        _fields_ = [("tot_nsweeps", C.c_int), ("mag_output_int", C.c_int), ("grid_output_int", C.c_int)]

    class McFunction(C.Structure):  # include/mc_cpu.h:17-26
        _fields_ = [("itask", C.c_int), ("dn_thr", C.c_double), ("up_thr", C.c_double), ("ninputs", C.c_int),
                    ("initial_spin", C.c_int), ("cv", C.c_char_p), ("result", C.POINTER(C.c_float)),
                    ("filename", C.c_char_p)]

    outf_t = C.CFUNCTYPE(C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_float),
                         C.POINTER(C.c_float), C.c_char_p, C.c_double, C.c_double, C.c_char_p)
    R.mc_driver_cpu.argtypes = [McGrids, C.c_double, C.c_double, C.POINTER(C.c_int), McSampler, McFunction, outf_t]
    R.mc_driver_cpu.restype = None
    return McGrids, McSampler, McFunction, outf_t


def _quiet_stdout():
    # the driver prints a line per check: keep it off this process's stdout
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)


def _driver_child(nsweeps, q):
    import ctypes as C
    from oracle import oracle as O

    _quiet_stdout()
    R = O.ref()
    McGrids, McSampler, McFunction, outf_t = _driver_types(R)
    L, n, beta, h, mag = 32, 256, 0.54, 0.07, 100
    grids = -np.ones(n * L * L, dtype=np.int32)
    fates = np.full(n, -1, dtype=np.int32)
    result = np.zeros(nsweeps // mag + 1, dtype=np.float32)
    noop = outf_t(lambda *args: 0)
    R.init_genrand(2894203475)
    R.preComputeProbs_cpu(beta, h)
    t0 = time.perf_counter()
    R.mc_driver_cpu(McGrids(L, n, grids.ctypes.data_as(C.POINTER(C.c_int))), beta, h,
                    fates.ctypes.data_as(C.POINTER(C.c_int)), McSampler(nsweeps, mag, 10 * nsweeps),
                    McFunction(0, -0.9, 0.9, 1, -1, b"magnetisation", result.ctypes.data_as(C.POINTER(C.c_float)),
                               b"gridstates.bin"), noop)
    dt = time.perf_counter() - t0
    # itask 0 stops at the first check where every replica has nucleated
    # (src/mc_cpu.c:218-220): that check's sweep count is what ran
    full = np.nonzero(result[: nsweeps // mag + 1] >= 1.0)[0]
    done = int(full[0]) * mag if full.size else nsweeps
    q.put((done * n * L * L, dt))


def _committor_child(i, grid, replicas, max_sweeps, seed, start, q):
    """src/mc_cpu.c:106-308 with itask 1 on one input grid: `replicas` copies
    (src/grid.c:41-71), one MT19937 stream seeded seed + i over all of them,
    checks every 100 sweeps against -0.90 / +0.90, every replica swept until
    every fate is decided or max_sweeps."""
    import ctypes as C
    from oracle import oracle as O

    _quiet_stdout()
    R = O.ref()
    McGrids, McSampler, McFunction, outf_t = _driver_types(R)
    L = grid.shape[-1]
    beta, h, mag = 0.54, 0.07, 100
    grids = np.ascontiguousarray(np.broadcast_to(grid.reshape(1, -1), (replicas, L * L)), dtype=np.int32)
    fates = np.full(replicas, -1, dtype=np.int32)
    result = np.zeros(1, dtype=np.float32)
    noop = outf_t(lambda *args: 0)
    R.init_genrand(seed + i)
    R.preComputeProbs_cpu(beta, h)
    start.wait()
    t0 = time.perf_counter()
    R.mc_driver_cpu(McGrids(L, replicas, grids.ctypes.data_as(C.POINTER(C.c_int))), beta, h,
                    fates.ctypes.data_as(C.POINTER(C.c_int)), McSampler(max_sweeps, mag, 100 * max_sweeps),
                    McFunction(1, -0.9, 0.9, 1, -1, b"magnetisation", result.ctypes.data_as(C.POINTER(C.c_float)),
                               b"gridstates.bin"), noop)
    q.put((i, float(result[0]), int((fates == -1).sum()), time.perf_counter() - t0))


def run_committor(cores: int, replicas: int = 64, max_sweeps: int = 20000, seed: int = 31) -> dict:
    """The reference's CPU committor (mc_driver_cpu, itask 1) to completion on
    a bounded sample: inputs 0 .. cores-1 of the committor workload (512
    droplet grids, L = 64), one per process, `replicas` replicas each; every
    process starts together and the rate is inputs over the slowest process's
    time."""
    from gpu_2dising_amd.surrogate import disc_grids
    from oracle import oracle as O

    if O.ref() is None:
        raise RuntimeError("oracle/_ref/libgasp_ref.so is not built")
    inputs = disc_grids(512, 64, 12345)[:cores]
    ctx = mp.get_context("fork")
    start, q = ctx.Barrier(cores), ctx.Queue()
    procs = [ctx.Process(target=_committor_child, args=(i, inputs[i], replicas, max_sweeps, seed, start, q))
             for i in range(cores)]
    for p in procs:
        p.start()
    res = sorted(q.get() for _ in procs)
    for p in procs:
        p.join()
    dt = max(r[3] for r in res)
    return {"value": cores / dt, "unit": "pB estimates/s", "cores": cores, "kind": "reference",
            "replicas_per_input": replicas, "inputs": list(range(cores)), "pB": [r[1] for r in res],
            "unresolved": sum(r[2] for r in res), "seconds": dt,
            "sample": f"src/mc_cpu.c mc_driver_cpu itask 1 (compiled reference, one MT19937 stream per process), "
                      f"committor inputs 0..{cores - 1} (L=64 droplets), {replicas} replicas each, checks every 100 "
                      f"sweeps, thresholds -0.90/+0.90, run to completion, {cores} processes, {dt:.1f} s"}


def run_driver(nsweeps: int = 1000) -> dict:
    """src/mc_cpu.c mc_driver_cpu at L = 32, 256 replicas, in a forked child."""
    from oracle import oracle as O

    if O.ref() is None:
        raise RuntimeError("oracle/_ref/libgasp_ref.so is not built")
    ctx = mp.get_context("fork")
    q = ctx.Queue()
    p = ctx.Process(target=_driver_child, args=(nsweeps, q))
    p.start()
    attempts, dt = q.get()
    p.join()
    return {"value": attempts / dt, "unit": "attempts/s", "cores": 1, "kind": "reference",
            "sample": f"src/mc_cpu.c mc_driver_cpu (compiled reference, one MT19937 stream over all replicas), "
                      f"L=32, 256 replicas x {nsweeps} sweeps, checks every 100, beta=0.54, h=0.07, all -1, "
                      f"{dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--cores", type=int, default=1)
    ap.add_argument("--seconds", type=float, default=12.0)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--driver", action="store_true", help="time mc_driver_cpu at L = 32, 256 replicas")
    ap.add_argument("--sweeps", type=int, default=1000, help="--driver: sweeps")
    ap.add_argument("--committor", action="store_true",
                    help="time mc_driver_cpu itask 1 on committor inputs 0..cores-1")
    ap.add_argument("--replicas", type=int, default=64, help="--committor: replicas per input")
    a = ap.parse_args()
    if a.committor:
        out = run_committor(a.cores, a.replicas)
    else:
        out = run_driver(a.sweeps) if a.driver else run(a.cores, a.seconds, a.L)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
