"""Pins the CNN oracle (oracle/cnn_oracle.py): its analytic gradient against
central finite differences in float64, and its Adam step against the
published update rule (Kingma & Ba 2015, Algorithm 1)."""
import numpy as np


def test_gradient_finite_differences():
    from oracle import cnn_oracle as K

    rs = np.random.default_rng(0)
    p = K.init_params(1).astype(np.float64) + rs.normal(0, 0.05, K.NPARAM)
    spins = np.where(rs.random((3, 8, 8)) < 0.4, 1, -1)
    y = np.array([0.1, 0.7, 0.95])
    _, g, _ = K.loss_and_grad(p, spins, y)
    idx = np.concatenate([rs.choice(K.NPARAM, 80, replace=False), [0, 287, 288, 320, 9535, 9536, 28063, 30240]])
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (K.loss_and_grad(pp, spins, y)[0] - K.loss_and_grad(pm, spins, y)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * (abs(fd) + abs(g[i])) + 1e-10, (i, fd, g[i])


def test_param_layout_and_bf16():
    from oracle import cnn_oracle as K

    assert K.NPARAM == 30241
    d = K.unflatten(np.arange(K.NPARAM, dtype=np.float64))
    assert d["b1"][0] == 288 and d["w2"][0, 0, 0] == 320 and d["fw1"][0, 0] == 28064 and d["fb2"][0] == 30240
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1415926, 65504.0, 1e-3])
    r = K.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125  # ties to even, then round up
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


def test_adam_rule():
    from oracle import cnn_oracle as K

    p, g = np.array([1.0, -2.0]), np.array([0.5, -0.25])
    m, v = np.zeros(2), np.zeros(2)
    p1, m, v = K.adam_step(p, g, m, v, 1, lr=0.1)
    # first step of Adam moves each parameter by lr * sign(g)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-7)
