"""CPU oracle of the committor-surrogate CNN (numpy, float64 arithmetic).

TEST INFRASTRUCTURE ONLY: the checker for the HIP training and inference
kernels (gpu_2dising_amd/csrc/gasp_cnn.hip); the product never imports it.
The network has no counterpart in the reference (it is the learned addition
of the build); it is specified as:

  input  x0 = spins (+-1), 1 x 64 x 64 per grid
  conv1  1 -> 32, 3x3 circular, bias, ReLU      a1
  conv2  32 -> 32, 3x3 circular, bias, ReLU     a2
  conv3, conv4 likewise                         a3, a4
  GAP    g[c] = mean over positions of a4[., c]
  fc1    32 -> 64, bias, ReLU                   h
  fc2    64 -> 1, bias                          z;  pB = sigmoid(z)
  loss   binary cross-entropy against the shooting label y in [0, 1],
         mean over the batch: softplus(z) - y z

Convolution is cross-correlation with circular wrap:
  out[y, x, co] = b[co] + sum_{t, ci} w[t, ci, co] in[(y + dy_t) % L, (x + dx_t) % L, ci],
  t = 3 (dy + 1) + (dx + 1).

Flat parameter vector (float32, 30241 entries), in this order:
  w1[9][32] b1[32] w2[9][32][32] b2[32] w3 b3 w4 b4 fw1[32][64] fb1[64] fw2[64] fb2[1]

`bf16=True` rounds (round-to-nearest-even) at the points where the GPU
kernels store bfloat16: conv weights, and each conv layer's post-ReLU
activations (forward) and its masked output gradient (backward).
"""
from __future__ import annotations

import numpy as np

C = 32
HID = 64
TAPS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]

SHAPES = [("w1", (9, C)), ("b1", (C,)),
          ("w2", (9, C, C)), ("b2", (C,)),
          ("w3", (9, C, C)), ("b3", (C,)),
          ("w4", (9, C, C)), ("b4", (C,)),
          ("fw1", (C, HID)), ("fb1", (HID,)),
          ("fw2", (HID,)), ("fb2", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
assert NPARAM == 30241


def unflatten(p: np.ndarray) -> dict:
    out, off = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[off:off + n], dtype=np.float64).reshape(shp)
        off += n
    return out


def flatten(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[n], dtype=np.float64).reshape(-1) for n, _ in SHAPES])


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest, ties to even); returns float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed: int) -> np.ndarray:
    """He-uniform conv/fc weights, zero biases (float32)."""
    rs = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("b") or name.startswith("fb"):
            d[name] = np.zeros(shp)
        else:
            fan_in = {"w1": 9, "w2": 9 * C, "w3": 9 * C, "w4": 9 * C, "fw1": C, "fw2": HID}[name]
            lim = np.sqrt(6.0 / fan_in)
            d[name] = rs.uniform(-lim, lim, size=shp)
    return flatten(d).astype(np.float32)


def _shift(a: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """s[b, y, x] = a[b, (y + dy) % L, (x + dx) % L]."""
    return np.roll(a, shift=(-dy, -dx), axis=(1, 2))


def conv(a: np.ndarray, w: np.ndarray, b: np.ndarray) -> np.ndarray:
    """a [n, L, L, ci], w [9, ci, co] -> [n, L, L, co]."""
    out = np.zeros(a.shape[:3] + (w.shape[2],))
    for t, (dy, dx) in enumerate(TAPS):
        out += _shift(a, dy, dx) @ w[t]
    return out + b


def forward(p: np.ndarray, spins: np.ndarray, bf16: bool = False) -> dict:
    P = unflatten(p)
    rnd = bf16_round if bf16 else (lambda v: v)
    x0 = np.asarray(spins, dtype=np.float64)[..., None]
    acts = [x0]
    a = x0
    for k, (wn, bn) in enumerate([("w1", "b1"), ("w2", "b2"), ("w3", "b3"), ("w4", "b4")]):
        w = P[wn] if k > 0 else P[wn][:, None, :]
        a = rnd(np.maximum(conv(a, rnd(w), P[bn]), 0.0))
        acts.append(a)
    g = acts[-1].mean(axis=(1, 2))
    hpre = g @ P["fw1"] + P["fb1"]
    h = np.maximum(hpre, 0.0)
    z = h @ P["fw2"] + P["fb2"][0]
    return {"acts": acts, "g": g, "hpre": hpre, "h": h, "z": z, "pB": 1.0 / (1.0 + np.exp(-z))}


def loss_and_grad(p: np.ndarray, spins: np.ndarray, y: np.ndarray, bf16: bool = False):
    P = unflatten(p)
    rnd = bf16_round if bf16 else (lambda v: v)
    f = forward(p, spins, bf16)
    n = spins.shape[0]
    z, y = f["z"], np.asarray(y, dtype=np.float64)
    loss = float(np.mean(np.logaddexp(0.0, z) - y * z))
    G = {}
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    G["fw2"] = f["h"].T @ dz
    G["fb2"] = np.array([dz.sum()])
    dh = np.outer(dz, P["fw2"]) * (f["hpre"] > 0)
    G["fw1"] = f["g"].T @ dh
    G["fb1"] = dh.sum(0)
    dg = dh @ P["fw1"].T
    L = spins.shape[1]
    da = np.broadcast_to(dg[:, None, None, :] / (L * L), f["acts"][4].shape)
    for k in (4, 3, 2, 1):
        a_out, a_in = f["acts"][k], f["acts"][k - 1]
        dzk = rnd(da * (a_out > 0))
        wn, bn = f"w{k}", f"b{k}"
        G[bn] = dzk.sum(axis=(0, 1, 2))
        if k == 1:
            G[wn] = np.stack([np.einsum("nyx,nyxo->o", _shift(a_in[..., 0], dy, dx), dzk) for dy, dx in TAPS])
            break
        w = rnd(P[wn])
        G[wn] = np.stack([np.einsum("nyxi,nyxo->io", _shift(a_in, dy, dx), dzk) for dy, dx in TAPS])
        # d a_in[q, ci] = sum_t sum_co dzk[q - off_t, co] w[t, ci, co]
        da = np.zeros(a_in.shape)
        for t, (dy, dx) in enumerate(TAPS):
            da += _shift(dzk, -dy, -dx) @ w[t].T
    return loss, flatten(G), f


def adam_step(p, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """Adam (Kingma & Ba 2015) with bias correction; float64 arithmetic."""
    p = np.asarray(p, dtype=np.float64)
    g = np.asarray(grad, dtype=np.float64) + wd * p
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
