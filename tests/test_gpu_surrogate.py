"""The trained committor surrogate (gpu_2dising_amd/surrogate_L64.npz) against
shooting, through the C ABI.

- On fresh committor inputs (disc_grids with a seed the training never used),
  surrogate estimates with shooting fallback for predictions in [0.2, 0.8]
  are within MAE 0.05 of 1024-replica shooting, and take less GPU time than
  that shooting and than plain shooting of every input at the smallest
  replica count that is as accurate (the gain the surrogate exists for).
- The network's forward pass with the trained parameters agrees with the
  float64 oracle evaluated with bfloat16 rounding at the GPU's rounding
  points (tolerance as in test_gpu_cnn).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PARAMS = os.path.join(ROOT, "gpu_2dising_amd", "surrogate_L64.npz")


@pytest.fixture(scope="module")
def params():
    return np.load(PARAMS)["params"]


def test_surrogate_estimates_beat_shooting(gpu, engine_lib, params):
    from gpu_2dising_amd import surrogate as SG

    print("[surrogate] 96 inputs vs 1024-replica shooting", flush=True)
    grids = SG.disc_grids(96, 64, seed=99001).astype(np.int8)
    ref, nA, nB, t_ref = SG.label_by_shooting(grids, reps=1024, seed=5)
    est, unsure, tm = SG.estimate(grids, params, reps=128, seed=6)
    mae = float(np.mean(np.abs(est - ref)))
    print(f"[surrogate] MAE {mae:.4f}, fallback {unsure.mean():.2f}, "
          f"{tm['seconds']:.2f} s vs shooting {t_ref:.2f} s", flush=True)
    assert mae <= 0.05
    assert tm["seconds"] < t_ref
    # grids the network answers alone are those with confident predictions
    assert np.all((est[~unsure] < 0.2) | (est[~unsure] > 0.8))
    # what the network gains over the method without it: plain shooting of every input needs more
    # GPU time at the smallest replica count that is as accurate (or at 256 replicas, if none of
    # 32 .. 256 is)
    for reps in (32, 64, 128, 256):
        pb, _, _, t_plain = SG.label_by_shooting(grids, reps=reps, seed=7 + reps)
        mae_plain = float(np.mean(np.abs(pb - ref)))
        print(f"[surrogate] plain shooting {reps} replicas: MAE {mae_plain:.4f}, {t_plain:.2f} s", flush=True)
        if mae_plain <= mae:
            break
    assert tm["seconds"] < t_plain, (tm["seconds"], reps, t_plain)


def test_trained_forward_matches_oracle(gpu, engine_lib, params):
    from oracle import cnn_oracle as CO

    G = engine_lib
    rs = np.random.default_rng(3)
    from gpu_2dising_amd import surrogate as SG

    grids = np.concatenate([SG.disc_grids(4, 64, seed=5),
                            np.where(rs.random((4, 64, 64)) < 0.3, 1, -1)]).astype(np.int8)
    with G.Surrogate(64, 256) as net:
        net.set_params(params)
        got = net.predict(grids)
    ref = CO.forward(params.astype(np.float64), grids.astype(np.float64), bf16=True)["pB"]
    assert np.max(np.abs(got - ref)) < 2e-3
