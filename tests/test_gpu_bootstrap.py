"""GPU bootstrap error bars (gasp_bootstrap_errbar_gpu) against the
reference's estimator: the libc-rand() bootstrap_errbar (bit-identical to
src/bootstrap.c, itself checked against the compiled reference in
test_oracle.py) and the binomial standard error sqrt(p (1 - p) / n).  The two
bootstraps use different random sequences, so they are compared as
estimates: with m = 4096 resamples each is within about 1% (one standard
deviation) of the true bootstrap standard deviation."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fates(n, nA, nB, seed):
    f = np.full(n, -1, dtype=np.int32)
    f[:nA] = 0
    f[nA:nA + nB] = 1
    np.random.default_rng(seed).shuffle(f)
    return f


def _libc_boot(G, f, m):
    return G.lib().bootstrap_errbar(f.size, m, np.ascontiguousarray(f).ctypes.data_as(C.POINTER(C.c_int)))


@pytest.mark.parametrize("n,nA,nB", [(1024, 700, 300), (1024, 524, 300), (333, 33, 300), (1024, 1000, 24)])
def test_matches_reference_estimator(gpu, engine_lib, n, nA, nB):
    G = engine_lib
    print(f"[bootstrap] n={n} nA={nA} nB={nB}", flush=True)
    f = _fates(n, nA, nB, n + nA)
    m = 4096
    got = G.bootstrap_gpu(f[None, :], m, seed=11)[0]
    ref = _libc_boot(G, f, m)
    # resampling n draws from all n fates (unresolved included): the resolved count varies
    p = nB / (nA + nB)
    binom = np.sqrt(p * (1 - p) / ((nA + nB) * n / n))
    assert abs(got - ref) < 0.06 * ref, (got, ref)
    assert abs(got - binom) < 0.12 * binom, (got, binom)


def test_batch_seed_and_edge_cases(gpu, engine_lib):
    G = engine_lib
    print("[bootstrap] batch", flush=True)
    fs = np.stack([_fates(512, 256, 256, 1), _fates(512, 500, 12, 2), np.full(512, -1, dtype=np.int32),
                   _fates(512, 0, 512, 3)])
    a = G.bootstrap_gpu(fs, 1024, seed=5)
    b = G.bootstrap_gpu(fs, 1024, seed=5)
    c = G.bootstrap_gpu(fs, 1024, seed=6)
    assert np.array_equal(a, b, equal_nan=True)
    assert a[0] != c[0]
    assert abs(a[0] - 0.5 / np.sqrt(512)) < 0.1 * 0.5 / np.sqrt(512)
    assert np.isnan(a[2])            # no resolved fate: 0/0, as the reference
    assert a[3] == 0.0               # all B
    with pytest.raises(G.GaspError):
        G.bootstrap_gpu(fs, 0)
