"""CPU baseline with the reference's own code: src/mc_cpu.c's mc_sweep_cpu,
compiled from the reference sources into oracle/_ref/libgasp_ref.so.

TEST / BASELINE INFRASTRUCTURE ONLY (run by bench.py's cpu_baseline leg in a
child process that never touches the GPU).  The reference keeps one global
MT19937 state (src/mt19937ar.c:56-57), so its sweep cannot run on threads;
here every core gets its own process (fork), its own replicas and its own
init_genrand seed, all processes start together, and the attempt rate is the
total attempts over the slowest process's time.

  python3 -m oracle.refbench --cores 16 --seconds 12 --L 64
prints one JSON object.
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--cores", type=int, default=1)
    ap.add_argument("--seconds", type=float, default=12.0)
    ap.add_argument("--L", type=int, default=64)
    a = ap.parse_args()
    print(json.dumps(run(a.cores, a.seconds, a.L)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
