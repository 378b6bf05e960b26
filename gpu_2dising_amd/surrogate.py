"""Learned committor surrogate: data from the engine, training on the GPU,
estimation with shooting as the fallback.

Pipeline (defaults follow the benchmark specification at L = 64):
  1. sample_trajectories: a nucleation swarm (all spins -1, beta = 0.54,
     h = 0.07, Philox streams); grids are sampled every `every` sweeps.
  2. select: 14000 grids, stratified by magnetisation over the transition
     region (grids outside the thresholds have pB 0 or 1 by definition), plus
     6000 committor-input-style grids (disc_grids: a +1 disc in a -1 sea with
     5% of the spins flipped) so that the network also covers shooting inputs.
  3. label_by_shooting: each grid gets `reps` replicas (256) run until every
     fate is decided against dn/up thresholds; the label is nB / (nA + nB).
  4. train: the CNN of gasp_cnn_* (BF16 MFMA kernels, Adam, batch 256) on a
     train split; MAE and calibration on the held-out split.
  5. estimate: surrogate pB for every grid; grids whose prediction falls in
     [lo, hi] are re-estimated by shooting.

Everything runs through the engine's C ABI (gpu_2dising_amd.Engine /
Surrogate); there is no CPU path.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import CHECK_FATES, Engine, RngState, Surrogate, init_device, set_tuning

BETA, H = 0.54, 0.07
DN, UP = -0.90, 0.90


def disc_grids(n: int, L: int = 64, seed: int = 12345, r_lo: float = 4.0, r_hi: float = 14.0,
               flip: float = 0.05) -> np.ndarray:
    """Committor shooting inputs: a -1 sea with a +1 disc of radius U[r_lo, r_hi]
    (sites within periodic distance R of a uniformly random integer centre),
    then a fraction `flip` of the sites (uniform, without replacement) flipped."""
    rs = np.random.default_rng(seed)
    r = np.arange(L)
    out = np.empty((n, L, L), dtype=np.int32)
    for k in range(n):
        rad = rs.uniform(r_lo, r_hi)
        cy, cx = rs.integers(0, L, 2)
        dy = np.minimum(np.abs(r - cy), L - np.abs(r - cy))[:, None]
        dx = np.minimum(np.abs(r - cx), L - np.abs(r - cx))[None, :]
        g = np.where(dy * dy + dx * dx <= rad * rad, 1, -1).astype(np.int32)
        fl = rs.choice(L * L, size=int(round(flip * L * L)), replace=False)
        g.reshape(-1)[fl] *= -1
        out[k] = g
    return out


def sample_trajectories(L=64, R=16384, sweeps=50000, every=500, seed=101, keep_per_snapshot=1024,
                        window=(-0.88, 0.88), rng_seed=0, log=None):
    """Returns int8 grids, their magnetisation and sweep.  At every sample all
    replicas inside `window` are kept, plus `keep_per_snapshot` others drawn
    at random (metastable or nucleated)."""
    rs = np.random.default_rng(rng_seed)
    st = RngState(R, seed)
    grids, mags, times = [], [], []
    with Engine(L, R, 2, st, BETA, H) as e:
        e.set_grids(-np.ones((R, L, L), dtype=np.int32))
        for k in range(sweeps // every):
            e.sweep(every)
            m = e.magnetisation()
            inside = np.nonzero((m > window[0]) & (m < window[1]))[0]
            others = np.setdiff1d(np.arange(R), inside)
            extra = rs.choice(others, size=min(keep_per_snapshot, others.size), replace=False)
            idx = np.concatenate([inside, extra])
            g = e.gather(idx)
            grids.append(g)
            mags.append(m[idx])
            times.append(np.full(idx.size, (k + 1) * every))
            if log and (k + 1) % 10 == 0:
                log(f"sampled to sweep {(k + 1) * every}: {sum(len(x) for x in grids)} grids")
    st.free()
    return np.concatenate(grids), np.concatenate(mags), np.concatenate(times)


def select(grids, mags, n=20000, frac_transition=0.7, bins=18, seed=1):
    """Stratified choice: frac_transition * n grids with DN < m < UP, spread
    evenly over `bins` magnetisation bins (a bin short of grids gives its
    share to the others), the rest from outside the thresholds."""
    rs = np.random.default_rng(seed)
    inside = np.nonzero((mags > DN) & (mags < UP))[0]
    outside = np.nonzero(~((mags > DN) & (mags < UP)))[0]
    want = min(inside.size, int(frac_transition * n))
    edges = np.linspace(DN, UP, bins + 1)
    members = [inside[(mags[inside] >= a) & (mags[inside] < b)] for a, b in zip(edges[:-1], edges[1:])]
    take = np.zeros(bins, dtype=int)
    while take.sum() < want:
        open_ = [k for k in range(bins) if take[k] < members[k].size]
        if not open_:
            break
        share = max(1, (want - take.sum()) // len(open_))
        for k in open_:
            take[k] = min(members[k].size, take[k] + share)
            if take.sum() >= want:
                break
    chosen = [rs.choice(m, t, replace=False) for m, t in zip(members, take) if t > 0]
    no = min(outside.size, n - int(take.sum()))
    idx = np.concatenate(chosen + [rs.choice(outside, no, replace=False)])
    rs.shuffle(idx)
    return idx


def label_by_shooting(grids, reps=256, seed=202, max_sweeps=20000, chk=100, chunk=4096, L=None, log=None):
    """pB of each grid by shooting: `reps` replicas per grid with independent
    Philox streams, until every fate is decided (or max_sweeps).  Returns
    (pB, nA, nB, seconds)."""
    n = grids.shape[0]
    L = L or grids.shape[1]
    nA = np.zeros(n, dtype=np.int64)
    nB = np.zeros(n, dtype=np.int64)
    t0 = time.perf_counter()
    for i0 in range(0, n, chunk):
        g = np.ascontiguousarray(grids[i0:i0 + chunk], dtype=np.int32)
        k = g.shape[0]
        R = k * reps
        st = RngState(R, seed + i0)
        with Engine(L, R, 2, st, BETA, H) as e:
            e.set_inputs(g)
            e.set_freeze(True)
            e.set_fates(np.full(R, -1, dtype=np.int32))
            e.sweep(0, CHECK_FATES, DN, UP)
            c, sw = e.counts(), 0
            while c[1] + c[2] < R and sw < max_sweeps:
                e.sweep(chk, CHECK_FATES, DN, UP)
                sw += chk
                c = e.counts()
            f = e.fates().reshape(k, reps)
        st.free()
        nA[i0:i0 + k] = (f == 0).sum(1)
        nB[i0:i0 + k] = (f == 1).sum(1)
        if log:
            log(f"labelled {i0 + k}/{n} grids ({sw} sweeps in this chunk), {time.perf_counter() - t0:.1f} s")
    pB = nB / np.maximum(nA + nB, 1)
    return pB.astype(np.float32), nA, nB, time.perf_counter() - t0


def init_params(seed: int, nparam: int = 30241) -> np.ndarray:
    """He-uniform conv / fc weights, zero biases, in the gasp_cnn parameter order."""
    rs = np.random.default_rng(seed)
    shapes = [("w", 9 * 32, 9), ("b", 32, 0), ("w", 9 * 32 * 32, 288), ("b", 32, 0),
              ("w", 9 * 32 * 32, 288), ("b", 32, 0), ("w", 9 * 32 * 32, 288), ("b", 32, 0),
              ("w", 32 * 64, 32), ("b", 64, 0), ("w", 64, 64), ("b", 1, 0)]
    out = []
    for kind, size, fan in shapes:
        out.append(rs.uniform(-np.sqrt(6.0 / fan), np.sqrt(6.0 / fan), size) if kind == "w" else np.zeros(size))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == nparam
    return p


def train(grids, labels, *, epochs=40, batch=256, lr=2e-3, seed=3, val_frac=0.1, log=print):
    """Trains on a random (1 - val_frac) split; returns params and metrics."""
    n = grids.shape[0]
    L = grids.shape[1]
    rs = np.random.default_rng(seed)
    perm = rs.permutation(n)
    nval = int(round(val_frac * n))
    val, tr = perm[:nval], perm[nval:]
    net = Surrogate(L, max(batch, 256))
    net.set_params(init_params(seed))
    net.set_data(grids, labels)
    t0 = time.perf_counter()
    hist = []
    for ep in range(epochs):
        lr_ep = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))  # cosine decay
        loss = net.train_epoch(rs.permutation(tr), batch, float(lr_ep))
        hist.append(loss)
        if log and (ep % 10 == 0 or ep == epochs - 1):
            pv = net.predict(grids[val])
            log(f"epoch {ep}: train loss {loss:.4f}, val MAE {np.mean(np.abs(pv - labels[val])):.4f}")
    t_train = time.perf_counter() - t0
    params = net.params()
    pv = net.predict(grids[val])
    net.close()
    return params, {"train_seconds": t_train, "epochs": epochs, "loss_history": hist,
                    "val_idx": val, "val_pred": pv, "val_mae": float(np.mean(np.abs(pv - labels[val])))}


def calibration(pred, label, bins=10):
    """Mean label per bin of predicted pB (a reliability curve)."""
    edges = np.linspace(0, 1, bins + 1)
    out = []
    for a, b in zip(edges[:-1], edges[1:]):
        m = (pred >= a) & (pred < b if b < 1 else pred <= b)
        if m.any():
            out.append({"bin": [float(a), float(b)], "n": int(m.sum()), "mean_pred": float(pred[m].mean()),
                        "mean_label": float(label[m].mean())})
    return out


def estimate(grids, params, *, lo=0.2, hi=0.8, reps=128, seed=303, max_sweeps=20000):
    """Surrogate pB for all grids; shooting (`reps` replicas each) for
    predictions in [lo, hi].  Returns estimates, the fallback mask, and a dict
    of timings and the network's raw predictions ("pred").  128 replicas bound
    the shooting standard error by sqrt(0.25 / 128) = 0.044."""
    L = grids.shape[1]
    t0 = time.perf_counter()
    net = Surrogate(L, 4096)
    net.set_params(params)
    pred = net.predict(grids)
    net.close()
    t_pred = time.perf_counter() - t0
    unsure = (pred >= lo) & (pred <= hi)
    est = pred.copy()
    t_shoot = 0.0
    if unsure.any():
        pb, _, _, t_shoot = label_by_shooting(grids[unsure], reps=reps, seed=seed, max_sweeps=max_sweeps)
        est[unsure] = pb
    return est, unsure, {"predict_seconds": t_pred, "fallback_seconds": t_shoot,
                         "seconds": time.perf_counter() - t0, "pred": pred}


FWD_FLOP = 2.29e8  # per L=64 grid: conv1 2.36 MFLOP + 3 x 75.5 MFLOP (32 -> 32, 3x3)


def bench(params, n: int = 4096, train_batch: int = 256, seed: int = 5) -> dict:
    """Inference (one batch of n) and training-epoch throughput of the CNN at L = 64."""
    grids = disc_grids(n, 64, seed=seed).astype(np.int8)
    out = {}
    with Surrogate(64, n) as net:
        net.set_params(params)
        net.predict(grids)
        t0 = time.perf_counter()
        for _ in range(5):
            net.predict(grids)
        dt = (time.perf_counter() - t0) / 5
        out.update(predict_grids_per_s=n / dt, predict_ms=dt * 1e3, forward_tflops=n * FWD_FLOP / dt / 1e12)
    with Surrogate(64, train_batch) as net:
        net.set_params(params)
        net.set_data(grids, np.random.default_rng(seed).random(n).astype(np.float32))
        order = np.arange(n, dtype=np.int32)
        net.train_epoch(order, train_batch, 1e-4)
        t0 = time.perf_counter()
        for _ in range(3):
            net.train_epoch(order, train_batch, 1e-4)
        dt = (time.perf_counter() - t0) / 3
        out.update(train_grids_per_s=n / dt, train_tflops=n * 3 * FWD_FLOP / dt / 1e12)
    return out


def main():
    import argparse

    ap = argparse.ArgumentParser(description="build the committor surrogate (data, labels, training)")
    ap.add_argument("--out", default="gpurun_out/surrogate")
    ap.add_argument("--n", type=int, default=14000, help="grids from nucleation trajectories")
    ap.add_argument("--discs", type=int, default=6000, help="disc + noise grids (committor-input style)")
    ap.add_argument("--reps", type=int, default=256)
    ap.add_argument("--sweeps", type=int, default=50000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--name", default="surrogate_L64", help="file name of the parameters and summary")
    ap.add_argument("--bench", action="store_true",
                    help="only time inference and training with the packaged parameters")
    args = ap.parse_args()
    if args.bench:
        init_device(-1)
        here = os.path.dirname(os.path.abspath(__file__))
        print(json.dumps(bench(np.load(os.path.join(here, "surrogate_L64.npz"))["params"])), flush=True)
        return
    os.makedirs(args.out, exist_ok=True)
    init_device(-1)
    set_tuning(0, 0)
    t = time.perf_counter()
    log = lambda msg: print(msg, flush=True)  # noqa: E731
    g, m, tt = sample_trajectories(sweeps=args.sweeps, log=log)
    print(f"sampled {g.shape[0]} grids in {time.perf_counter() - t:.1f} s", flush=True)
    idx = select(g, m, args.n)
    g, m = g[idx], m[idx]
    if args.discs:
        d = disc_grids(args.discs, 64, seed=777001).astype(np.int8)
        g = np.concatenate([g, d])
        m = np.concatenate([m, d.reshape(len(d), -1).mean(1, dtype=np.float64).astype(np.float32)])
    kind = np.concatenate([np.zeros(len(idx), dtype=np.int8), np.ones(args.discs, dtype=np.int8)])
    pB, nA, nB, ts = label_by_shooting(g, reps=args.reps, log=log)
    print(f"labelled {g.shape[0]} grids by shooting ({args.reps} replicas) in {ts:.1f} s", flush=True)
    np.savez_compressed(os.path.join(args.out, "dataset.npz"), grids=np.packbits(g.reshape(len(g), -1) > 0, axis=1),
                        mag=m, pB=pB, nA=nA, nB=nB, kind=kind)
    params, met = train(g, pB, epochs=args.epochs)
    np.save(os.path.join(args.out, "params.npy"), params)
    val = met["val_idx"]
    pv, yv, kv = met["val_pred"], pB[val], kind[val]
    unsure = (pv >= 0.2) & (pv <= 0.8)
    summary = {"n": int(g.shape[0]), "n_trajectory": int(len(idx)), "n_discs": int(args.discs), "reps": args.reps,
               "label_seconds": ts, "train_seconds": met["train_seconds"], "epochs": args.epochs,
               "val_n": int(len(val)), "val_mae": met["val_mae"],
               "val_mae_trajectory": float(np.mean(np.abs(pv - yv)[kv == 0])),
               "val_mae_discs": float(np.mean(np.abs(pv - yv)[kv == 1])),
               "val_fallback_fraction": float(unsure.mean()),
               "val_mae_outside_fallback": float(np.mean(np.abs(pv - yv)[~unsure])),
               "calibration": calibration(pv, yv),
               "loss_history": [float(x) for x in met["loss_history"]]}
    np.save(os.path.join(args.out, "val_idx.npy"), val)
    np.savez(os.path.join(args.out, args.name + ".npz"), params=params, val_mae=summary["val_mae"],
             n_train=len(g) - len(val), epochs=args.epochs, reps=args.reps, label_seconds=ts,
             train_seconds=met["train_seconds"], n_trajectory=len(idx), n_discs=args.discs)
    with open(os.path.join(args.out, args.name + ".json"), "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps({k: v for k, v in summary.items() if k != "loss_history"}), flush=True)


if __name__ == "__main__":
    main()
