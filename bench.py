#!/usr/bin/env python3
"""Benchmark of the replica-parallel Metropolis engine on one MI355X.

Workload (multispin throughput): L = 64, periodic, J = 1, beta = 0.54,
h = 0.07, 16384 replicas (256 words of 64 replicas, gpu_method 2), all spins
-1 at start, Philox4x32-10 streams.  One step = one 100-sweep launch of the
sweep kernel with the fused magnetisation / nucleation check, the host reading
the nucleated count (the driver's mag_output_int cadence); every 10th step
also copies the grids (grid_output_int = 1000) to host memory as the body of a
bit-packed gridstates.bin record (8.4 MB, packed on the device).  The
default 500 steps are the full 50000-sweep run; the nucleated fraction curve
is fitted to 1 - exp(-k (t - t0)).

Prints one JSON line.  `value` is spin-flip attempts per second over the timed
steps (wall clock, HIP-synchronised on both sides).  `roofline` is for the
sweep kernel, timed with HIP events on the engine's stream: LDS bytes per the
algorithmic figure, with the VALU issue rate beside it and the limiting one
named.  `cpu_baseline` times the reference's own CPU sweep (src/mc_cpu.c,
compiled from its sources into oracle/_ref) on every host core this process
may use (the host's nproc, affinity set, cgroup quota and lscpu model are
recorded) over a bounded sample, with one core and the oracle's restatement
reported beside it.  The committor, surrogate and batched legs follow the
throughput leg; the committor leg carries the reference's CPU committor
measured to completion on one input per host core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

L, R, BETA, H = 64, 16384, 0.54, 0.07
SWEEPS_PER_STEP, DUMP_EVERY_STEPS = 100, 10
UP_THR, DN_THR = 0.90, -0.90
# Roofline of the LDS-resident multispin kernel: 5 reads of the 4-byte LDS
# bank word holding the spin bit per attempt; peak = 256 CU x 128 B/clk x 2.4 GHz.
LDS_BYTES_PER_ATTEMPT = 20
HBM_BYTES_PER_ATTEMPT = 5
HBM_PEAK_GBS = 8000.0
LDS_PEAK_GBS = 256 * 128 * 2.4
VALU_PEAK = 256 * 4 * 2.4e9 / 4
METRIC = "spin-flip attempts/sec (1 GPU, L=64) + committor pB estimates/sec at MAE≤0.05"


def committor_inputs(n: int, L: int = 64, seed: int = 12345) -> np.ndarray:
    """Committor shooting inputs: a -1 sea with a +1 disc of radius U[4, 14]
    at a uniformly random integer centre, then 5% of the sites flipped."""
    from gpu_2dising_amd.surrogate import disc_grids

    return disc_grids(n, L, seed)


def committor_leg(G, n_inputs: int = 512, per: int = 1024, max_sweeps: int = 20000,
                  chk: int = 100, time_cap: float = 240.0, method: int = 2) -> dict:
    """Committor shooting (itask 1): every input grid gets `per` replicas with
    independent Philox streams, run until all fates are decided against
    dn = -0.90 / up = +0.90 (the CLI thresholds) or max_sweeps; replicas with a
    decided fate stop sweeping.  pB per input and its bootstrap error
    (bootstrap_errbar's estimator on glibc's rand() sequence, as the
    reference, on host threads) are computed on the host.  method 3: the
    bit-sliced engine (the 64 replicas of a word share their sites)."""
    import ctypes as C

    inputs = committor_inputs(n_inputs, L)
    R_c = n_inputs * per
    rng = G.RngState(R_c, 777)
    e = G.Engine(L, R_c, method, rng, BETA, H)
    e.set_inputs(inputs)
    e.set_freeze(True)
    e.set_fates(np.full(R_c, -1, dtype=np.int32))
    e.sync()
    t0 = time.perf_counter()
    e.sweep(0, G.CHECK_FATES, DN_THR, UP_THR)
    sweeps, c = 0, e.counts()
    while sweeps < max_sweeps and c[1] + c[2] < R_c and time.perf_counter() - t0 < time_cap:
        e.sweep(chk, G.CHECK_FATES, DN_THR, UP_THR)
        sweeps += chk
        c = e.counts()
    fates = e.fates()
    t_gpu = time.perf_counter() - t0
    kern_ms = e.kernel_ms()
    e.close()
    rng.free()
    f = fates.reshape(n_inputs, per)
    nB = (f == 1).sum(1)
    nres = (f >= 0).sum(1)
    pB = np.where(nres > 0, nB / np.maximum(nres, 1), np.nan).astype(np.float32)
    # bootstrap_errbar's estimator on glibc's rand() sequence from srand(1)
    # (glibc's unseeded state), bit for bit as successive calls would draw it;
    # the library jumps its own copy of the sequence to each input, so the
    # inputs run on 16 host threads
    t1 = time.perf_counter()
    errs = G.bootstrap_libc(f, per, seed=1, nthreads=16)
    t_boot = time.perf_counter() - t1
    # spot check against the real libc rand(): the first two inputs, called in order after srand(1)
    lib = G.lib()
    libc = C.CDLL("libc.so.6")
    libc.srand.argtypes = [C.c_uint]
    libc.srand(1)
    seq = [lib.bootstrap_errbar(per, per, np.ascontiguousarray(f[i]).ctypes.data_as(C.POINTER(C.c_int)))
           for i in range(2)]
    boot_checked = bool(np.array_equal(np.array(seq), errs[:2], equal_nan=True))
    t1 = time.perf_counter()
    errs_gpu = G.bootstrap_gpu(f, per, seed=1)
    t_boot_gpu = time.perf_counter() - t1
    return {
        "metric": "committor pB estimates/sec (shooting)",
        "value": n_inputs / (t_gpu + t_boot),
        "unit": "pB estimates/s",
        "gpu_only_value": n_inputs / t_gpu,
        "inputs": n_inputs, "replicas_per_input": per, "sweeps": sweeps,
        "unresolved": int(R_c - c[1] - c[2]), "gpu_seconds": t_gpu, "kernel_seconds": kern_ms / 1e3,
        "bootstrap_seconds": t_boot,
        "bootstrap_threads": 16,
        "bootstrap_matches_sequential_libc": boot_checked,
        "value_gpu_bootstrap": n_inputs / (t_gpu + t_boot_gpu),
        "bootstrap_gpu_seconds": t_boot_gpu,
        "mean_bootstrap_error_gpu": float(np.nanmean(errs_gpu)),
        "pB_quantiles": [float(x) for x in np.nanquantile(pB, [0.0, 0.25, 0.5, 0.75, 1.0])],
        "mean_bootstrap_error": float(np.nanmean(errs)),
        "workload": "512 grids: -1 sea, +1 disc of radius U[4,14] at a random centre, 5% of spins "
                    "flipped; 1024 replicas each; thresholds -0.90/+0.90; <= 20000 sweeps",
        "_pB": pB,
        "_fates": f,
    }


def committor_sliced_leg(G, shooting: dict) -> dict:
    """The committor leg on the bit-sliced engine (method 3): the same 512
    inputs x 1024 replicas, with the replicas of each 64-replica word sharing
    their site sequence.  Reported against the per-replica-stream shooting:
    the MAE of the two pB estimates (two independent 1024-replica estimates
    differ by about sqrt(2 pB (1 - pB) / 1024) each), and the dispersion of
    the 16 word means of each input against the binomial variance pB (1 - pB)
    / 64 (1 when the replicas of a word are independent; inputs with pB in
    [0.05, 0.95])."""
    per = shooting["replicas_per_input"]
    out = committor_leg(G, per=per, method=3)
    pB, ref = out.pop("_pB"), shooting["_pB"]
    ok = np.isfinite(pB) & np.isfinite(ref)
    out["metric"] = "committor pB estimates/sec (shooting, bit-sliced engine)"
    out["mae_vs_per_replica_shooting"] = float(np.mean(np.abs(pB[ok] - ref[ok])))
    pm = 0.5 * (pB[ok] + ref[ok])
    out["expected_mae_two_independent_estimates"] = float(np.mean(np.sqrt(2.0 * pm * (1.0 - pm) / per)) * np.sqrt(2.0 / np.pi))
    out["speedup_vs_per_replica_shooting"] = out["value"] / shooting["value"]
    f = out.pop("_fates")
    words = (f == 1).reshape(f.shape[0], per // 64, 64).mean(2)  # [input][word]
    p = (f == 1).mean(1)
    sel = (p > 0.05) & (p < 0.95) & (f >= 0).all(1)
    disp = words[sel].var(1, ddof=1) / (p[sel] * (1.0 - p[sel]) / 64.0)
    out["word_dispersion"] = {"mean": float(disp.mean()), "inputs": int(sel.sum()),
                              "note": "variance of the 16 word means of an input over p(1-p)/64; 1 = independent replicas"}
    # the same statistic on the per-replica streams (independent by construction)
    fr = shooting["_fates"]
    wr = (fr == 1).reshape(fr.shape[0], per // 64, 64).mean(2)
    pr = (fr == 1).mean(1)
    sr = (pr > 0.05) & (pr < 0.95) & (fr >= 0).all(1)
    out["word_dispersion_per_replica_streams"] = float((wr[sr].var(1, ddof=1) / (pr[sr] * (1.0 - pr[sr]) / 64.0)).mean())
    # correlated replicas inflate an estimate's variance by the dispersion (the
    # design effect), so equal error needs that many times the replicas
    out["speedup_at_equal_error"] = out["speedup_vs_per_replica_shooting"] / max(out["word_dispersion"]["mean"], 1.0)
    return out


SURROGATE_PARAMS = os.path.join(ROOT, "gpu_2dising_amd", "surrogate_L64.npz")
# trained on 20000 nucleation-trajectory grids only (the training set as the surrogate workload states it)
SURROGATE_PARAMS_TRAJ = os.path.join(ROOT, "gpu_2dising_amd", "surrogate_L64_traj.npz")


def _training_cost(path: str, blob) -> dict:
    """GPU seconds of the training data's labels and of the training (stored with the parameters,
    or in the summary written beside them)."""
    out = {k: float(blob[k]) for k in ("val_mae", "n_train", "label_seconds", "train_seconds", "n_trajectory",
                                        "n_discs") if k in blob.files}
    js = path[:-4] + ".json"
    if "label_seconds" not in out and os.path.exists(js):
        with open(js) as f:
            meta = json.load(f)
        for k in ("label_seconds", "train_seconds", "n_trajectory", "n_discs"):
            if k in meta:
                out[k] = float(meta[k])
    return out


def surrogate_leg(G, shooting: dict, n_inputs: int = 512, fallback_reps: int = 128) -> dict:
    """Committor estimates for the shooting leg's 512 inputs from the trained network, with shooting
    (fallback_reps replicas) for predictions in [0.2, 0.8]; beside it plain shooting of every input at
    16 .. 512 replicas.  Accuracy is measured against the shooting leg's 1024-replica pB (MAE over
    all inputs and over those with pB in [0.2, 0.8]); the gain is over plain shooting at the smallest
    replica count that reaches the surrogate path's MAE, and the training cost (labels + training,
    GPU seconds) is reported per estimate of this run and as the break-even number of estimates."""
    from gpu_2dising_amd import surrogate as SG

    inputs = committor_inputs(n_inputs, L).astype(np.int8)
    ref = shooting["_pB"]
    ok = ~np.isnan(ref)
    band = ok & (ref >= 0.2) & (ref <= 0.8)

    def accuracy(est):
        err = np.abs(est - ref)
        return {"mae_vs_shooting": float(err[ok].mean()),
                "mae_band_0.2_0.8": float(err[band].mean()) if band.any() else None,
                "max_error": float(err[ok].max())}

    def network_only(tm):
        # the network's own estimates (no fallback): accuracy, rate over the inference time, and the
        # plain shooting that matches its MAE
        acc = accuracy(tm["pred"])
        match = next((k for k in plain if plain[k]["mae_vs_shooting"] <= acc["mae_vs_shooting"]), None)
        return {**acc, "value": n_inputs / tm["predict_seconds"], "plain_shooting_same_mae": match,
                "gain_vs_plain_same_mae": (plain[match]["seconds"] / tm["predict_seconds"]) if match else None}

    plain = {}
    for reps in (16, 32, 64, 128, 256, 512):
        t0 = time.perf_counter()
        pb, _, _, _ = SG.label_by_shooting(inputs, reps=reps, seed=5050 + reps)
        secs = time.perf_counter() - t0
        plain[str(reps)] = {"seconds": secs, "value": n_inputs / secs, **accuracy(pb)}
    out = {"metric": "committor pB estimates/sec (surrogate, shooting fallback for pB in [0.2, 0.8])",
           "unit": "pB estimates/s", "inputs": n_inputs, "band_inputs": int(band.sum()),
           "reference": "pB of the shooting leg (1024 replicas per input, the same 512 inputs)",
           "plain_shooting": plain}
    for name, path in (("mixed", SURROGATE_PARAMS), ("trajectory", SURROGATE_PARAMS_TRAJ)):
        if not os.path.exists(path):
            continue
        blob = np.load(path)
        t0 = time.perf_counter()
        est, unsure, tm = SG.estimate(inputs, blob["params"], reps=fallback_reps, seed=4040)
        secs = time.perf_counter() - t0
        acc = accuracy(est)
        # plain shooting at the smallest replica count whose MAE is at most the surrogate path's
        match = next((k for k in plain if plain[k]["mae_vs_shooting"] <= acc["mae_vs_shooting"]), None)
        cost = _training_cost(path, blob)
        train_s = cost.get("label_seconds", 0.0) + cost.get("train_seconds", 0.0)
        res = {"value": n_inputs / secs, "seconds": secs, **acc, "fallback_fraction": float(unsure.mean()),
               "fallback_replicas": fallback_reps, "predict_seconds": tm["predict_seconds"],
               "fallback_seconds": tm["fallback_seconds"],
               "network_only": network_only(tm),
               "calibration": SG.calibration(tm["pred"][ok], ref[ok], bins=5),
               "training": cost,
               "plain_shooting_same_mae": match,
               "gain_vs_plain_same_mae": plain[match]["seconds"] / secs if match else None,
               "gain_vs_plain_fallback_replicas": plain[str(fallback_reps)]["seconds"] / secs,
               "training_gpu_seconds_per_estimate_this_run": train_s / n_inputs}
        if match and plain[match]["seconds"] > secs:
            res["break_even_estimates"] = train_s / ((plain[match]["seconds"] - secs) / n_inputs)
        out[name] = res
    if "mixed" in out:
        out["value"] = out["mixed"]["value"]
    return out


def batched_leg(G, replays: int = 10, chk: int = 20, seed: int = 2026) -> dict:
    """Uneven batch in one captured graph: L in {32, 48, 64, 96} (multispin,
    LDS) with {8192, 4096, 4096, 2048} replicas and L = 256 (HBM path) with
    256; beta ~ U[0.50, 0.60], h ~ U[0.03, 0.10] per group; itask 0 (count
    nucleated) and itask 1 (fates from droplet inputs) alternate.  Launch
    shapes and sweeps per launch come from the autotuner, probed at each
    group's own (beta, h), then refined by a coordinate search on the captured
    graph (`Batch._refine`).  One replay = `chk` sweeps of every group plus its
    threshold check."""
    from gpu_2dising_amd.batched import Batch

    t0 = time.perf_counter()
    b = Batch(chk=chk, seed=seed, refine=True)
    setup_s = time.perf_counter() - t0
    attempts = b.attempts_per_replay
    graph = b.graph
    graph.launch(1)  # warm-up
    ms0 = graph.ms()
    t0 = time.perf_counter()
    graph.launch(replays)
    wall = time.perf_counter() - t0
    g_ms = (graph.ms() - ms0) / replays
    # the same work one group after another on the engines' own streams
    t0 = time.perf_counter()
    for _ in range(replays):
        for g in b.groups:
            g.engine.sweep(chk, G.CHECK_COUNT if g.itask == 0 else G.CHECK_FATES, b.dn, b.up)
            g.engine.sync()
    seq_ms = (time.perf_counter() - t0) * 1e3 / replays
    groups = [{"L": g.L, "replicas": g.R, "method": g.method, "slots": g.slots, "waves": g.waves,
               "sweeps_per_launch": g.launch_sweeps, "probe_rate": g.probe_rate, "beta": g.beta, "h": g.h,
               "itask": g.itask, "conflict_filter": g.engine.conflict_filter(),
               "lattice_staging": g.engine.lattice_staging()} for g in b.groups]
    order = [b.groups[i].L for i in b.order]
    tuned_ms, refine_log = b.tuned_ms, b.refine_log
    b.close()
    return {"metric": "batched spin-flip attempts/sec (one captured graph, uneven L)",
            "value": attempts / (g_ms * 1e-3), "unit": "attempts/s",
            "wall_value": attempts * replays / wall, "graph_ms_per_replay": g_ms,
            "sequential_ms_per_replay": seq_ms, "sweeps_per_replay": chk, "replays": replays,
            "attempts_per_replay": attempts, "capture_order_L": order, "groups": groups,
            "tuned_ms_per_replay": tuned_ms, "refinement": refine_log, "setup_seconds": setup_s}


def fit_rate(sweeps: np.ndarray, frac: np.ndarray):
    """Least-squares fit of 1 - exp(-k (t - t0)) (the reference notebooks' model)."""
    try:
        from scipy.optimize import curve_fit

        m = frac > 0
        if m.sum() < 3:
            return None, None
        f = lambda t, k, t0: 1.0 - np.exp(-k * (t - t0))  # noqa: E731
        (k, t0), _ = curve_fit(f, sweeps[m], frac[m], p0=(1.5e-4, 500.0), maxfev=20000)
        return float(k), float(t0)
    except Exception:
        return None, None


def _port_baseline(ncores: int, max_seconds: float) -> dict:
    """The oracle's per-replica restatement of mc_cpu.c on ncores threads."""
    from oracle import oracle as O

    nrep = 8 * ncores
    g = -np.ones((nrep, L, L), dtype=np.int32)
    t0 = time.perf_counter()
    O.sweep_mt_per_replica(L, g, 2894203475, 20, BETA, H, nthreads=ncores)
    dt = (time.perf_counter() - t0) / 20
    nsweeps = int(max(1, min(200000, max_seconds / max(dt, 1e-6))))
    g = -np.ones((nrep, L, L), dtype=np.int32)
    t0 = time.perf_counter()
    O.sweep_mt_per_replica(L, g, 2894203475, nsweeps, BETA, H, nthreads=ncores)
    dt = time.perf_counter() - t0
    return {"value": nrep * nsweeps * L * L / dt, "unit": "attempts/s", "cores": ncores, "kind": "port",
            "sample": f"L={L}, {nrep} replicas x {nsweeps} sweeps from all -1, beta={BETA}, h={H}, "
                      f"per-replica MT19937 (src/mc_cpu.c arithmetic restated in oracle/), {dt:.1f} s"}


def host_description() -> dict:
    """The host cores this process may use: nproc, the affinity set, the
    cgroup CPU quota and lscpu's model line."""
    import subprocess

    out = {"nproc": os.cpu_count()}
    try:
        out["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        out["affinity"] = os.cpu_count()
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        out["cgroup_cpus"] = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        out["cgroup_cpus"] = None
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        info = dict(line.split(":", 1) for line in txt.splitlines() if ":" in line)
        out["model"] = " ".join(info.get("Model name", "").split())
        out["cpu_family"] = info.get("CPU family", "").strip()
        out["cpu_model"] = info.get("Model", "").strip()
        out["threads_per_core"] = info.get("Thread(s) per core", "").strip()
    except Exception:
        pass
    return out


def host_cores(host: dict) -> int:
    """Every core of the affinity set, bounded by the cgroup quota."""
    n = host.get("affinity") or 1
    if host.get("cgroup_cpus"):
        n = min(n, max(1, int(host["cgroup_cpus"])))
    return max(1, int(n))


def cpu_baseline(max_seconds: float = 12.0) -> dict:
    """The reference's own mc_cpu.c (compiled from its sources into oracle/_ref) on every host core
    this process may use, one process per core (its MT19937 state is global), run in a child process
    that does not use the GPU; the oracle's threaded restatement is reported beside it and stands
    in when oracle/_ref was not built."""
    import subprocess

    host = host_description()
    ncores = host_cores(host)
    port = _port_baseline(ncores, max_seconds / 2)
    try:
        r = subprocess.run([sys.executable, "-m", "oracle.refbench", "--cores", str(ncores), "--seconds",
                            str(max_seconds), "--L", str(L)], cwd=ROOT, capture_output=True, text=True,
                           timeout=10 * max_seconds + 60)
        ref = json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
    except Exception:
        ref = None
    if ref is None:
        port["host"] = host
        return port
    ref["port"] = port
    ref["host"] = host
    try:
        r1 = subprocess.run([sys.executable, "-m", "oracle.refbench", "--cores", "1", "--seconds", "4", "--L",
                             str(L)], cwd=ROOT, capture_output=True, text=True, timeout=120)
        if r1.returncode == 0:
            ref["one_core"] = json.loads(r1.stdout.strip().splitlines()[-1])
        # the reference's literal single-stream driver (one MT19937 over all replicas) at L = 32
        r2 = subprocess.run([sys.executable, "-m", "oracle.refbench", "--driver", "--sweeps", "1000"], cwd=ROOT,
                            capture_output=True, text=True, timeout=180)
        if r2.returncode == 0:
            ref["driver_1_core"] = json.loads(r2.stdout.strip().splitlines()[-1])
    except Exception:
        pass
    return ref


def cpu_committor(ncores: int, replicas: int = 64) -> dict | None:
    """The reference's CPU committor, src/mc_cpu.c mc_driver_cpu with itask 1 (compiled into
    oracle/_ref), run to completion on committor inputs 0 .. ncores-1 with `replicas` replicas each,
    one process per core, in a child process that does not use the GPU."""
    import subprocess

    try:
        r = subprocess.run([sys.executable, "-m", "oracle.refbench", "--committor", "--cores", str(ncores),
                            "--replicas", str(replicas)], cwd=ROOT, capture_output=True, text=True, timeout=600)
        return json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
    except Exception:
        return None


def throughput_run(G, method: int, steps: int, warmup: int, dump: bool):
    """The multispin throughput workload with `method`'s kernel: `warmup` untimed
    steps, a restart from all -1, then `steps` timed steps of one 100-sweep
    launch with the fused check and the host reading the count, a bit-packed
    gridstates.bin record to host memory every 10th step.  Returns the wall
    time, the sweep kernel's time (HIP events), the nucleated fractions and
    the engine's conflict filter (1 row/column masks, 2 mark plane)."""
    rng = G.RngState(R, 20251202)
    eng = G.Engine(L, R, method, rng, BETA, H)
    g0 = -np.ones((R, L, L), dtype=np.int32)
    buf = np.empty((R * L * L + 7) // 8, dtype=np.uint8)  # gridstates.bin record body
    eng.set_grids(g0)
    for _ in range(warmup):
        eng.sweep(SWEEPS_PER_STEP, G.CHECK_COUNT, DN_THR, UP_THR)
        eng.counts()
    eng.set_grids(g0)
    eng.sync()
    ms0 = eng.kernel_ms()
    frac = []
    t0 = time.perf_counter()
    for step in range(steps):
        eng.sweep(SWEEPS_PER_STEP, G.CHECK_COUNT, DN_THR, UP_THR)
        c = eng.counts()  # synchronises
        frac.append(c[0] / R)
        if dump and (step + 1) % DUMP_EVERY_STEPS == 0:
            eng.gridstates(buf)
    eng.sync()
    wall = time.perf_counter() - t0
    kern_ms = eng.kernel_ms() - ms0
    filt = eng.conflict_filter()
    eng.close()
    rng.free()
    return wall, kern_ms, frac, filt


def sliced_leg(G, steps: int, warmup: int) -> dict:
    """The same workload with the bit-sliced multispin kernel (engine gpu_method
    3: the 64 replicas of a word share one site per attempt; neighbour sums by
    bitwise adders, acceptance by bit-sliced comparison of per-replica draws),
    at the tuner's launch shape."""
    S3, split3 = G.get_tuning(3, L, R)
    wall, kern_ms, frac, _ = throughput_run(G, 3, steps, warmup, True)
    attempts = float(steps) * SWEEPS_PER_STEP * L * L * R
    per_launch_ms = kern_ms / max(steps, 1)
    wb = 64 // split3
    sweeps = np.arange(1, steps + 1) * SWEEPS_PER_STEP
    k, t0fit = fit_rate(sweeps.astype(float), np.array(frac))
    # LDS: five 4- or 8-byte word reads per attempt serve the wb replicas of the word
    bpa = 5.0 * (wb // 8) / wb
    achieved = SWEEPS_PER_STEP * L * L * R * bpa / (per_launch_ms / 1e3) / 1e9
    return {"metric": "spin-flip attempts/sec, bit-sliced multispin (shared site per word)",
            "value": attempts / wall, "unit": "attempts/s", "ms_per_step": wall * 1e3 / max(steps, 1),
            "steps": steps, "kernel": f"sweep_sliced<{wb}, {S3}, true>", "waves_per_workgroup": S3,
            "replicas_per_word": wb, "kernel_ms_per_launch": per_launch_ms,
            "kernel_attempts_per_s": attempts / (kern_ms / 1e3),
            "lds": {"bytes_per_attempt": bpa, "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / LDS_PEAK_GBS},
            "nucleation_fit": {"k_per_sweep": k, "t0": t0fit, "final_fraction": frac[-1] if frac else None,
                               "reference_k_per_sweep": 1.5032e-4}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--method", type=int, default=2)
    ap.add_argument("--slots", type=int, default=0, help="slots per replica (0 = tuner)")
    ap.add_argument("--word-split", type=int, default=0, choices=(0, 1, 2, 4, 8),
                    help="replicas per workgroup: 1 = 64 (whole words), 2 = 32, 4 = 16, 8 = 8 (split words), "
                         "0 = tuner")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-dump", action="store_true", help="skip the grid dumps")
    ap.add_argument("--no-committor", action="store_true", help="skip the committor shooting leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (graph) leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate committor leg")
    ap.add_argument("--no-sliced", action="store_true", help="skip the bit-sliced multispin leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): replicas are independent "
              "chains and this build measures one GPU's share.", file=sys.stderr)
        return 2

    import gpu_2dising_amd as G

    G.init_device(-1)
    G.set_tuning(args.slots, args.word_split)
    S, split = G.get_tuning(args.method, L, R)
    wall, kern_ms, frac, filt = throughput_run(G, args.method, args.steps, args.warmup, not args.no_dump)

    attempts = float(args.steps) * SWEEPS_PER_STEP * L * L * R
    value = attempts / wall
    kern_rate = attempts / (kern_ms / 1e3)
    per_launch_ms = kern_ms / max(args.steps, 1)
    # methods 1/2: LDS bytes (5 reads of the 4-byte bank word holding the spin bit); method 0: HBM
    # bytes (5 spin bytes read per attempt)
    hbm = args.method == 0
    # method 3 (bit-sliced): five 8-byte word reads per attempt serve the word's 64 replicas
    bpa = HBM_BYTES_PER_ATTEMPT if hbm else (5 * 8 / 64 if args.method == 3 else LDS_BYTES_PER_ATTEMPT)
    peak = HBM_PEAK_GBS if hbm else LDS_PEAK_GBS
    achieved = SWEEPS_PER_STEP * L * L * R * bpa / (per_launch_ms / 1e3) / 1e9
    multi = {1: "LAT_MULTISPIN", 2: "LAT_MULTIHALF", 4: "LAT_MULTI16", 8: "LAT_MULTI8"}
    kname = {0: "LAT_HBM", 1: "LAT_BITS", 2: multi.get(split, "LAT_MULTISPIN"), 3: "sliced"}[args.method]
    sweeps = np.arange(1, args.steps + 1) * SWEEPS_PER_STEP
    k, t0fit = fit_rate(sweeps.astype(float), np.array(frac))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (all spins -1 at start)",
        "config": {
            "workload": "multispin throughput: L=64, 16384 replicas (64 per 64-bit word), "
                        "beta=0.54, h=0.07, all -1, 100 sweeps per step with fused magnetisation "
                        "and nucleation check, bit-packed gridstates.bin record (8.4 MB) to host memory "
                        "every 1000 sweeps",
            "L": L, "replicas": R, "sweeps_per_step": SWEEPS_PER_STEP,
            "total_sweeps": args.steps * SWEEPS_PER_STEP, "gpu_method": args.method,
            "rng": "philox4x32-10 per replica", "slots_per_replica": S,
            "replicas_per_workgroup": 64 // max(split, 1) if args.method == 2 else None,
            "conflict_filter": ({1: "row/column masks", 2: "mark plane"}[filt] if args.method == 2 else None),
            "grid_dumps": not args.no_dump,
        },
        "roofline": {
            "bound": "hbm" if hbm else "lds",
            "achieved": achieved,
            "peak": peak,
            "unit": "GB/s",
            "frac": achieved / peak,
            "traffic": None,
            "bytes_per_attempt": bpa,
            "kernel": (f"sweep_sliced<{64 // max(split, 1)}, {S}, true>" if args.method == 3
                       else f"sweep_kernel<{kname}, RNG_PHILOX, S{', MOFF' if filt == 2 else ''}>"),
            "kernel_ms_per_launch": per_launch_ms,
            "kernel_attempts_per_s": kern_rate,
        },
        "nucleation_fit": {"k_per_sweep": k, "t0": t0fit, "final_fraction": frac[-1] if frac else None,
                           "reference_k_per_sweep": 1.5032e-4},
    }
    traffic_file = os.path.join(ROOT, "profiles", "traffic_sweep_multispin_L64.json")
    if args.method == 2 and os.path.exists(traffic_file):
        with open(traffic_file) as f:
            prof = json.load(f)
        # the counters hold for the kernel source they were measured on
        import hashlib

        with open(os.path.join(ROOT, "gpu_2dising_amd", "csrc", "gasp_kernels.h"), "rb") as f:
            src_sha = hashlib.sha256(f.read()).hexdigest()
        current = prof.get("kernel_source_sha256") == src_sha
        out["roofline"]["traffic"] = prof.get("hbm_bytes_per_launch") if current else None
        vpa = prof.get("valu_instructions_per_attempt") if current else None
        out["roofline"]["counters"] = {"file": "profiles/traffic_sweep_multispin_L64.json",
                                       "kernel_source_sha256": prof.get("kernel_source_sha256"),
                                       "matches_current_source": current}
        if vpa:
            # VALU issue: wave64 instructions per attempt (rocprofv3 SQ_INSTS_VALU over the same
            # launch) x the live attempt rate, against 256 CU x 4 SIMD x 2.4 GHz / 4 cycles
            out["roofline"]["valu"] = {"instructions_per_attempt": vpa, "achieved": kern_rate * vpa,
                                       "peak": VALU_PEAK, "unit": "wave64 instr/s",
                                       "frac": kern_rate * vpa / VALU_PEAK}
            # the LDS pipe's busy share (SQ_LDS_IDX_ACTIVE over the kernel's cycles, bank
            # conflicts included) from the same counter file
            busy = prof.get("lds_active_share")
            if busy:
                out["roofline"]["lds_pipe"] = {"active_share": busy,
                                               "bank_conflict_share": prof.get("lds_bank_conflict_share")}
            # the binding resource is the highest of the fractions
            fr = {"valu": kern_rate * vpa / VALU_PEAK, "lds bandwidth": achieved / peak, "lds pipe": busy or 0.0}
            out["roofline"]["limiter"] = max(fr, key=fr.get)
    G.set_tuning(0, 0)
    if not args.no_sliced and args.method != 3:
        out["sliced"] = sliced_leg(G, args.steps, args.warmup)
    if not args.no_committor:
        out["committor"] = committor_leg(G)
        if not args.no_sliced:
            out["committor_sliced"] = committor_sliced_leg(G, out["committor"])
        if not args.no_surrogate and (os.path.exists(SURROGATE_PARAMS) or os.path.exists(SURROGATE_PARAMS_TRAJ)):
            out["surrogate"] = surrogate_leg(G, out["committor"])
    if not args.no_batched:
        out["batched"] = batched_leg(G)
    if not args.no_cpu:
        cb = cpu_baseline()
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
        if "committor" in out:
            # The reference's CPU committor (src/mc_cpu.c:190-267: every replica of the input swept
            # until all its fates are decided), measured to completion on one input per host core at
            # 64 replicas per input.  An estimate at the shooting leg's 1024 replicas sweeps 16 times
            # the replicas for at least as many sweeps, so its rate is at most 64/1024 of that.
            cm = out["committor"]
            cc = cpu_committor(host_cores(cb.get("host", {})))
            if cc is not None:
                ref_pB = cm["_pB"][cc["inputs"]]
                cc["mae_vs_gpu_shooting_1024"] = float(np.nanmean(np.abs(np.array(cc["pB"]) - ref_pB)))
                cc["value_1024_upper_bound"] = cc["value"] * cc["replicas_per_input"] / cm["replicas_per_input"]
                cm["cpu"] = cc
                cm["gpu_over_cpu_1024"] = cm["value"] / cc["value_1024_upper_bound"]
    if "committor" in out:
        out["committor"].pop("_pB")
        out["committor"].pop("_fates")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
