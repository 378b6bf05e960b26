"""The committor-surrogate CNN's BF16 MFMA kernels against the CPU oracle
(oracle/cnn_oracle.py, float64 with bfloat16 rounding at the same points):
forward pB, the full parameter gradient of the loss, one Adam step, and a
training run that lowers the loss.  The oracle's own gradient is pinned by
finite differences in tests/test_cnn_oracle.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(L, n, seed):
    from oracle import cnn_oracle as K

    rs = np.random.default_rng(seed)
    p = K.init_params(seed).astype(np.float64)
    p += rs.normal(0.0, 0.05, p.shape)  # non-zero biases, asymmetric weights
    p = p.astype(np.float32)
    frac = rs.uniform(0.1, 0.9, n)[:, None, None]
    spins = np.where(rs.random((n, L, L)) < frac, 1, -1).astype(np.int8)
    y = rs.uniform(0, 1, n).astype(np.float32)
    return K, p, spins, y


@pytest.mark.parametrize("L,n", [(16, 5), (64, 3)])
def test_forward_matches_oracle(gpu, engine_lib, L, n):
    K, p, spins, y = _case(L, n, 1)
    print(f"[cnn fwd] L={L} n={n}", flush=True)
    with engine_lib.Surrogate(L, 8) as net:
        net.set_params(p)
        got = net.predict(spins)
    ref = K.forward(p.astype(np.float64), spins.astype(np.float64), bf16=True)["pB"]
    assert np.max(np.abs(got - ref)) < 2e-3, np.max(np.abs(got - ref))


@pytest.mark.parametrize("L,n", [(16, 6), (32, 4), (64, 3), (128, 2)])
def test_gradient_matches_oracle(gpu, engine_lib, L, n):
    K, p, spins, y = _case(L, n, 2)
    print(f"[cnn grad] L={L} n={n}", flush=True)
    with engine_lib.Surrogate(L, 8) as net:
        net.set_params(p)
        loss, g = net.gradient(spins, y)
    rl, rg, _ = K.loss_and_grad(p.astype(np.float64), spins.astype(np.float64), y.astype(np.float64), bf16=True)
    assert abs(loss - rl) < 1e-3 * max(1.0, abs(rl))
    Gg, Gr = K.unflatten(g), K.unflatten(rg)
    for name, _ in K.SHAPES:
        a, b = Gg[name].ravel(), Gr[name].ravel()
        err = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
        assert err < 2e-2, (name, err)


def test_adam_step_and_training(gpu, engine_lib):
    K, p, spins, y = _case(16, 8, 3)
    print("[cnn adam]", flush=True)
    with engine_lib.Surrogate(16, 8) as net:
        net.set_params(p)
        _, g = net.gradient(spins, y)
        net.set_params(p)
        net.train_step(spins, y, 1e-3)
        p1 = net.params()
    ref, _, _ = K.adam_step(p.astype(np.float64), g.astype(np.float64), np.zeros(K.NPARAM), np.zeros(K.NPARAM), 1)
    assert np.max(np.abs(p1 - ref)) < 1e-6
    # learn pB = (m + 1) / 2 from 64 random grids: the loss must fall
    rs = np.random.default_rng(5)
    frac = rs.uniform(0, 1, 64)[:, None, None]
    s = np.where(rs.random((64, 16, 16)) < frac, 1, -1).astype(np.int8)
    lab = ((s.reshape(64, -1).mean(1) + 1) / 2).astype(np.float32)
    with engine_lib.Surrogate(16, 32) as net:
        net.set_params(K.init_params(7))
        net.set_data(s, lab)
        first = net.train_epoch(np.arange(64), 32, 3e-3)
        for ep in range(60):
            last = net.train_epoch(rs.permutation(64), 32, 3e-3)
        pred = net.predict(s)
    assert last < first - 0.02, (first, last)
    assert np.mean(np.abs(pred - lab)) < 0.08
