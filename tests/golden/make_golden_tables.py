"""Writes tests/golden/acceptance_tables.json: the reference's two acceptance
tables as hex, at the committed workloads' (beta, h).

  single precision, src/mc_gpu.cu:25-31 (host code):
      P[5(s+1)+nsum+4] = expf(-(float)beta * (2.0f * (float)s * ((float)nsum + (float)h)))
  double precision, src/mc_cpu.c:28-35:
      P[...] = exp(-beta * 2.0 * s * (nsum + h))

evaluated here with numpy float32 / float64 arithmetic in the same order and
the C library's expf / exp called through ctypes (the host code's own calls),
plus the exact integer thresholds T = ceil(P 2^32 - 0.5) (2^32 for P >= 1)
that "(u + 0.5) 2^-32 < P" reduces to.  (beta, h): the reference workloads'
0.54 / 0.07, the batched configuration's five groups (beta ~ U[0.50, 0.60],
h ~ U[0.03, 0.10] from numpy's default_rng(2026), as gpu_2dising_amd/batched.py
draws them), and a few signs and edge values of h.

    python tests/golden/make_golden_tables.py
"""
import ctypes
import ctypes.util
import json
import math
import os
from fractions import Fraction

import numpy as np

libm = ctypes.CDLL(ctypes.util.find_library("m"))
libm.expf.argtypes = [ctypes.c_float]
libm.expf.restype = ctypes.c_float
libm.exp.argtypes = [ctypes.c_double]
libm.exp.restype = ctypes.c_double


def threshold(p: float) -> int:
    """Smallest-u threshold: (u + 1/2) 2^-32 < p  <=>  u < T, for u in [0, 2^32)."""
    if p >= 1.0:
        return 1 << 32
    x = Fraction(p) * (1 << 32) - Fraction(1, 2)
    return max(0, math.ceil(x))


def tables(beta: float, h: float):
    f32, f64 = [0.0] * 20, [0.0] * 20
    for s in (-1, 1):
        for nsum in (-4, -2, 0, 2, 4):
            idx = 5 * (s + 1) + nsum + 4
            de = np.float32(2.0) * np.float32(s) * (np.float32(nsum) + np.float32(h))
            f32[idx] = float(libm.expf(float(-np.float32(beta) * np.float32(de))))
            f64[idx] = libm.exp(-beta * 2.0 * s * (nsum + h))
    return f32, f64


def main() -> None:
    rs = np.random.default_rng(2026)
    points = [(0.54, 0.07)]
    for _ in range(5):
        beta, h = float(rs.uniform(0.50, 0.60)), float(rs.uniform(0.03, 0.10))
        points.append((beta, h))
    points += [(0.3, 0.07), (0.45, -0.3), (0.45, 0.0), (0.45, 2.5), (0.16, 0.07), (0.35, 0.2)]
    out = []
    for beta, h in points:
        f32, f64 = tables(beta, h)
        out.append({"beta": beta, "h": h,
                    "f32_hex": [np.float32(x).view(np.uint32).item().to_bytes(4, "big").hex() for x in f32],
                    "f64_hex": [float(x).hex() for x in f64],
                    "thr_f32": [threshold(float(np.float32(x))) for x in f32],
                    "thr_f64": [threshold(x) for x in f64]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "acceptance_tables.json")
    with open(path, "w") as f:
        json.dump({"index": "5(s+1) + nsum + 4; odd entries are unused (0)", "tables": out}, f, indent=1)
    print(f"wrote {path}: {len(out)} (beta, h) points")


if __name__ == "__main__":
    main()
