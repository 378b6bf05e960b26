#!/usr/bin/env python3
"""Writes tests/golden/golden_ref.npz: known-answer vectors produced by the
reference's own C code, compiled from its sources by `make -C oracle ref`
(oracle/_ref/libgasp_ref.so: src/mc_cpu.c, src/mt19937ar.c, src/bootstrap.c).

Vectors (all small):
  mt_<seed>          first 2000 genrand_int32 outputs and the 100000th
  pacc_<k>           preComputeProbs_cpu tables (double, src/mc_cpu.c:14-38)
  sweep_L<L>_*       start grids and grids after mc_sweep_cpu with
                     init_genrand(base + r) per replica (src/mc_cpu.c:41-89)
  c1_*               BASELINE config 1 run: L=32, 256 replicas, 2000 sweeps,
                     per-replica seeds base+r, magnetisation before each
                     100-sweep chunk (compute_magnetisation_cpu) and the
                     final grids, bit-packed
  clus_*             compute_largest_cluster_cpu on random grids
  mag_*              compute_magnetisation_cpu on random grids
  boot_*             bootstrap_errbar after srand(1) on fixed fate vectors

Run from the repository root: python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_ref.npz")
BASE_SEED = 2894203475  # src/ising.cu:55 (commented development seed)
TABLES = [(0.54, 0.07), (0.3, 0.0), (0.6, 0.1), (0.5, 0.03), (1.2, 0.5)]


def packbits(g):
    return np.packbits((np.asarray(g).reshape(-1) == 1).astype(np.uint8), bitorder="little")


def main():
    O.build(with_ref=True)
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libgasp_ref.so missing (needs the reference tree)")
    libc = C.CDLL(None)
    out = {}

    for seed in (5489, BASE_SEED, 1, 4357):
        R.init_genrand(seed)
        v = np.array([R.genrand_int32() for _ in range(2000)], dtype=np.uint32)
        for _ in range(100000 - 2000 - 1):
            R.genrand_int32()
        out[f"mt_{seed}"] = v
        out[f"mt_{seed}_100000th"] = np.array([R.genrand_int32()], dtype=np.uint32)

    pacc = (C.c_double * 20).in_dll(R, "Pacc")
    for k, (beta, h) in enumerate(TABLES):
        R.preComputeProbs_cpu(beta, h)
        out[f"pacc_{k}"] = np.array(list(pacc), dtype=np.float64)
        out[f"pacc_{k}_beta_h"] = np.array([beta, h])

    rs = np.random.default_rng(20251202)
    beta, h = 0.54, 0.07
    for L, nrep, nsweeps, start in [(4, 6, 200, "random"), (5, 6, 200, "random"), (8, 6, 100, "random"),
                                    (16, 4, 60, "random"), (17, 4, 60, "down"), (32, 4, 40, "down"),
                                    (48, 3, 20, "random"), (64, 3, 20, "down")]:
        if start == "down":
            g0 = -np.ones((nrep, L, L), dtype=np.int32)
        else:
            g0 = np.where(rs.random((nrep, L, L)) < 0.5, 1, -1).astype(np.int32)
        g = g0.copy()
        for r in range(nrep):
            R.init_genrand(BASE_SEED + r)
            R.preComputeProbs_cpu(beta, h)
            flat = np.ascontiguousarray(g[r].reshape(-1))
            R.mc_sweep_cpu(L, flat, 0, beta, h, nsweeps)
            g[r] = flat.reshape(L, L)
        out[f"sweep_L{L}_start"] = packbits(g0)
        out[f"sweep_L{L}_final"] = packbits(g)
        out[f"sweep_L{L}_meta"] = np.array([L, nrep, nsweeps, BASE_SEED], dtype=np.int64)

    # config 1: L=32, 256 replicas, 2000 sweeps, checks every 100
    L, nrep, chunks, spc = 32, 256, 20, 100
    t0 = time.time()
    R.preComputeProbs_cpu(beta, h)
    mags = np.zeros((nrep, chunks), dtype=np.float32)
    final = np.zeros((nrep, L * L), dtype=np.int32)
    m1 = np.zeros(1, dtype=np.float32)
    for r in range(nrep):
        R.init_genrand(BASE_SEED + r)
        grid = -np.ones(L * L, dtype=np.int32)
        for c in range(chunks):
            R.compute_magnetisation_cpu(L, grid, 0, m1)
            mags[r, c] = m1[0]
            R.mc_sweep_cpu(L, grid, 0, beta, h, spc)
        final[r] = grid
    out["c1_mags"] = mags
    out["c1_final"] = packbits(final)
    out["c1_meta"] = np.array([L, nrep, chunks, spc, BASE_SEED], dtype=np.int64)
    print(f"config 1 reference run: {time.time() - t0:.1f} s")

    clus_in, clus_up, clus_dn = [], [], []
    for L in (4, 8, 16, 33):
        for p in (0.2, 0.5, 0.7):
            g = np.where(rs.random((L, L)) < p, 1, -1).astype(np.int32).reshape(-1)
            o = np.zeros(1, dtype=np.float32)
            R.compute_largest_cluster_cpu(L, g, 0, 1, o)
            up = o[0]
            R.compute_largest_cluster_cpu(L, g, 0, -1, o)
            clus_in.append(packbits(g))
            clus_up.append(up)
            clus_dn.append(o[0])
    out["clus_Ls"] = np.repeat(np.array([4, 8, 16, 33]), 3)
    out["clus_bits"] = np.concatenate(clus_in)
    out["clus_up"] = np.array(clus_up, dtype=np.float32)
    out["clus_dn"] = np.array(clus_dn, dtype=np.float32)

    mag_bits, mag_vals = [], []
    for L in (3, 8, 48):
        g = np.where(rs.random((L, L)) < 0.37, 1, -1).astype(np.int32).reshape(-1)
        o = np.zeros(1, dtype=np.float32)
        R.compute_magnetisation_cpu(L, g, 0, o)
        mag_bits.append(packbits(g))
        mag_vals.append(o[0])
    out["mag_Ls"] = np.array([3, 8, 48])
    out["mag_bits"] = np.concatenate(mag_bits)
    out["mag_vals"] = np.array(mag_vals, dtype=np.float32)

    libc.srand.argtypes = [C.c_uint]
    fates = [np.array(rs.choice([-1, 0, 1], size=n, p=[0.1, 0.6, 0.3]), dtype=np.int32) for n in (64, 200, 1024)]
    libc.srand(1)
    boot = [R.bootstrap_errbar(len(f), len(f), f) for f in fates]
    out["boot_fates"] = np.concatenate(fates)
    out["boot_sizes"] = np.array([len(f) for f in fates])
    out["boot_vals"] = np.array(boot, dtype=np.float64)

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
