"""The benchmark workloads at their full sizes on the GPU, checked against the
oracle on sampled replicas and against the reference's published physics.

- Committor shooting at its full shape: 512 droplet inputs (a -1 sea, a +1
  disc of radius U[4, 14], 5% of the spins flipped) x 1024 replicas at
  L = 64, beta = 0.54, h = 0.07, thresholds -0.90 / +0.90, checks every 100
  sweeps, decided replicas frozen, until every fate is decided or 20000
  sweeps (src/mc_gpu.cu:734-768, 812-826).  Freezing does not change a
  replica's fate, so every sampled replica's chain is replayed on the CPU
  from the same Philox stream: fate and final grid must be equal.
- The multispin throughput run at its full length: L = 64, 16384 replicas
  from all -1, 50000 sweeps, nucleated fraction every 100 sweeps; the rate k
  of 1 - exp(-k (t - t0)) must agree with the reference's 1.5032e-4 per sweep
  (Jupyter/GASP_rates.ipynb:178) within 6% (about 2 standard errors of the
  two estimates combined), and two replicas are replayed over the first
  2000 sweeps.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L, BETA, H = 64, 0.54, 0.07


def test_committor_full_shape(gpu, engine_lib, oracle):
    from gpu_2dising_amd.surrogate import disc_grids

    G = engine_lib
    n_in, per, chk, max_sweeps = 512, 1024, 100, 20000
    R = n_in * per
    seed = 777
    inputs = disc_grids(n_in, L, 12345)
    print(f"[committor] {n_in} inputs x {per} replicas, L={L}", flush=True)
    rng = G.RngState(R, seed)
    e = G.Engine(L, R, 2, rng, BETA, H)
    e.set_inputs(inputs)
    e.set_freeze(True)
    e.set_fates(np.full(R, -1, dtype=np.int32))
    t0 = time.perf_counter()
    e.sweep(0, G.CHECK_FATES, -0.9, 0.9)
    sweeps, c = 0, e.counts()
    while sweeps < max_sweeps and c[1] + c[2] < R:
        e.sweep(chk, G.CHECK_FATES, -0.9, 0.9)
        sweeps += chk
        c = e.counts()
    fates = e.fates()
    print(f"[committor] {sweeps} sweeps, {time.perf_counter() - t0:.2f} s, unresolved {R - c[1] - c[2]}", flush=True)
    assert c[1] + c[2] == R and c[1] == int((fates == 0).sum()) and c[2] == int((fates == 1).sum())
    pB = (fates.reshape(n_in, per) == 1).mean(1)
    # droplets of radius 4 shrink and of radius 14 grow: both ends are reached
    assert pB.min() < 0.1 and pB.max() > 0.9
    rs = np.random.default_rng(3)
    idx = np.unique(np.concatenate([[0, R - 1], rs.choice(R, 30, replace=False)]))
    start = inputs[idx // per].astype(np.int32)
    print(f"[committor] oracle replay of {idx.size} replicas", flush=True)
    ofates, ogrids, swept = oracle.shoot_replicas(L, start, idx, seed, BETA, H, chk, -0.9, 0.9, max_sweeps,
                                                  nthreads=16)
    assert np.array_equal(fates[idx], ofates)
    assert np.array_equal(e.gather(idx).astype(np.int32), ogrids)
    e.close()
    rng.free()


def test_nucleation_rate_full_run(gpu, engine_lib, oracle):
    from scipy.optimize import curve_fit

    G = engine_lib
    R, steps, spc = 16384, 500, 100
    seed = 20251202
    print(f"[rate] L={L}, {R} replicas, {steps * spc} sweeps", flush=True)
    rng = G.RngState(R, seed)
    e = G.Engine(L, R, 2, rng, BETA, H)
    g0 = -np.ones((R, L, L), dtype=np.int32)
    e.set_grids(g0)
    frac = []
    early = None
    for s in range(steps):
        e.sweep(spc, G.CHECK_COUNT, -0.9, 0.9)
        frac.append(e.counts()[0] / R)
        if s == 19:
            early = e.gather([0, 9999]).astype(np.int32)
        if s % 100 == 99:
            print(f"[rate] {s + 1} steps, nucleated {frac[-1]:.4f}", flush=True)
    e.close()
    rng.free()
    t = np.arange(1, steps + 1) * spc
    frac = np.array(frac)
    m = frac > 0
    (k, t0), cov = curve_fit(lambda t, k, t0: 1.0 - np.exp(-k * (t - t0)), t[m], frac[m], p0=(1.5e-4, 500.0),
                             maxfev=20000)
    print(f"[rate] k = {k:.5e} per sweep (reference 1.5032e-4), t0 = {t0:.0f}", flush=True)
    assert abs(k / 1.5032e-4 - 1.0) < 0.06
    assert frac[-1] > 0.99
    ref = -np.ones((2, L, L), dtype=np.int32)
    oracle.sweep_philox_replicas(L, ref, [0, 9999], seed, 0, 20 * spc * L * L, BETA, H)
    assert np.array_equal(early, ref)
