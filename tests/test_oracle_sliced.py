"""Pins of the oracle's bit-sliced multispin stream (oracle.sweep_sliced, the
checker of engine gpu_method 3), none of which retypes its arithmetic:

  - the site sequence, read back from the oracle itself at beta = 0 (every
    attempt then flips its site in every replica), equals floor(w N / 2^32) of
    an independent vectorised Philox4x32-10 (checked against the Random123
    known answers) and is uniform over the sites (chi-square, several L);
  - the acceptance draws: from a state where every replica is in one class,
    single attempts flip each replica with the class's probability, and the
    flip counts of the 64 replicas of a word follow Binomial(64, p) (the
    replicas' draws are independent), with pair frequencies p^2;
  - the stationary distribution: every replica of a 64-replica word on a 3 x 3
    lattice visits the 512 states with their exact Boltzmann weights, with
    standard errors from batch means (which measure the autocorrelation and
    the correlation between replicas that share sites; the integrated
    autocorrelation time is reported);
  - frozen replicas do not change, and freezing one replica leaves the others'
    trajectories alone.
"""
import numpy as np
import pytest
from scipy import stats

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
DOMAIN = 0x534C4300


def philox_np(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 (Salmon et al. 2011) on uint64 arrays holding 32-bit words."""
    c = [np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF for x in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0), np.uint64(k1)
    m = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & m, p1 & m, ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & m, p0 & m]
        k0 = (k0 + np.uint64(W0)) & m
        k1 = (k1 + np.uint64(W1)) & m
    return c


def test_numpy_philox_known_answers():
    for ctr, key, out in [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
                          ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
                          ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
                           [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]:
        got = philox_np(*[np.array([x]) for x in ctr], *key)
        assert [int(x[0]) for x in got] == out


@pytest.mark.parametrize("L,wb", [(4, 64), (7, 32), (16, 64), (33, 64)])
def test_site_sequence_from_oracle(oracle, L, wb):
    N, R, seed, n = L * L, wb, 123456789012345 + L, 4000
    g = -np.ones((R, L, L), dtype=np.int32)
    sites = np.empty(n, dtype=np.int64)
    for i in range(n):
        before = g[0].copy()
        oracle.sweep_sliced(L, g, wb, seed, 1000 + i, 1, 0.0, 0.0, nthreads=1)
        diff = np.flatnonzero(g[0].reshape(-1) != before.reshape(-1))
        assert diff.size == 1
        sites[i] = diff[0]
        # beta = 0: every replica of the word flips the same site
        assert np.array_equal(g[1:] , np.broadcast_to(g[0], g[1:].shape))
    a = 1000 + np.arange(n, dtype=np.uint64)
    w = philox_np(a & np.uint64(0xFFFFFFFF), a >> np.uint64(32), np.zeros(n), np.full(n, DOMAIN),
                  seed & 0xFFFFFFFF, seed >> 32)
    assert np.array_equal(sites, ((w[0] * np.uint64(N)) >> np.uint64(32)).astype(np.int64))
    # uniform over the sites
    counts = np.bincount(sites, minlength=N)
    assert stats.chisquare(counts).pvalue > 1e-4


def test_acceptance_draws_independent_across_replicas(oracle):
    """All spins down at h = 0.07: every replica's site is in class (down, no
    up neighbour), accepted with p = expf(-beta 2 (4 - h)) (single precision
    table); 4000 single attempts from that state."""
    L, wb, seed = 8, 64, 2024
    beta, h = 0.16, 0.07
    p = float(oracle.pacc_gpu(beta, h)[0])
    assert 0.2 < p < 0.4
    n = 4000
    flips = np.empty((n, wb), dtype=bool)
    for i in range(n):
        g = -np.ones((wb, L, L), dtype=np.int32)
        oracle.sweep_sliced(L, g, wb, seed, i, 1, beta, h, nthreads=1)
        flips[i] = (g == 1).reshape(wb, -1).any(1)
    # each replica: frequency p
    f = flips.mean(0)
    assert np.all(np.abs(f - p) < 5 * np.sqrt(p * (1 - p) / n))
    # flip counts per attempt ~ Binomial(64, p): the draws of a word are independent
    k = flips.sum(1)
    assert abs(k.var() / (wb * p * (1 - p)) - 1) < 0.1
    obs = np.bincount(k, minlength=wb + 1)
    exp = stats.binom.pmf(np.arange(wb + 1), wb, p) * n
    keep = exp > 5
    chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum() + (obs[~keep].sum() - exp[~keep].sum()) ** 2 / max(
        exp[~keep].sum(), 1e-9)
    assert stats.chi2.sf(chi2, keep.sum()) > 1e-4
    # pairs of replicas: joint flip frequency p^2 (k (k - 1) / (64 * 63) per attempt)
    pair = k * (k - 1) / (wb * (wb - 1))
    assert abs(pair.mean() - p * p) < 5 * pair.std() / np.sqrt(n)


def _boltzmann_3x3(beta, h):
    L = 3
    states = np.array([[1 if (k >> i) & 1 else -1 for i in range(9)] for k in range(512)]).reshape(512, L, L)
    E = np.array([-np.sum(s * np.roll(s, 1, 0)) - np.sum(s * np.roll(s, 1, 1)) - h * np.sum(s) for s in states])
    w = np.exp(-beta * (E - E.min()))
    return w / w.sum()


def _state_index(g):
    R = g.shape[0]
    return ((g.reshape(R, 9) == 1) * (1 << np.arange(9))).sum(1)


@pytest.mark.parametrize("wb", [64, 32])
def test_boltzmann_3x3_sliced(oracle, wb):
    L, beta, h, seed = 3, 0.35, 0.2, 97
    exact = _boltzmann_3x3(beta, h)
    R = wb
    g = -np.ones((R, L, L), dtype=np.int32)
    oracle.sweep_sliced(L, g, wb, seed, 0, 200 * 9, beta, h, nthreads=1)  # burn-in
    nb, per = 40, 500
    freq = np.zeros((nb, 512))
    e_series = []
    a = 200 * 9
    for b in range(nb):
        for _ in range(per):
            oracle.sweep_sliced(L, g, wb, seed, a, 9, beta, h, nthreads=1)
            a += 9
            freq[b] += np.bincount(_state_index(g), minlength=512)
            e_series.append(g[0].sum())
        freq[b] /= per * R
    mean = freq.mean(0)
    se = freq.std(0, ddof=1) / np.sqrt(nb)
    # states seen about 20 times per batch or more one by one, the rare ones pooled
    big = exact * per * R >= 20
    z = (mean[big] - exact[big]) / se[big]
    assert big.sum() > 50
    assert np.max(np.abs(z)) < 5.5, np.max(np.abs(z))
    assert 0.6 < np.mean(z * z) < 1.6, np.mean(z * z)
    rare = freq[:, ~big].sum(1)
    assert abs(rare.mean() - exact[~big].sum()) < 5 * rare.std(ddof=1) / np.sqrt(nb)
    # integrated autocorrelation time of one replica's magnetisation (in sweeps),
    # from batch means against the naive variance
    x = np.array(e_series, dtype=float)
    bm = x.reshape(nb, per).mean(1)
    tau = per * bm.var(ddof=1) / x.var() / 2
    print(f"tau_int(m) = {tau:.2f} sweeps")
    assert tau < per / 10


def test_freeze_keeps_replicas_and_others_unchanged(oracle):
    L, wb, seed, n = 16, 64, 5, 3 * 256
    rs = np.random.default_rng(1)
    g0 = np.where(rs.random((100, L, L)) < 0.5, 1, -1).astype(np.int32)
    free = oracle.sweep_sliced(L, g0.copy(), wb, seed, 0, n, 0.3, 0.07)
    frozen = np.zeros(100, dtype=bool)
    frozen[[0, 5, 70, 99]] = True
    got = oracle.sweep_sliced(L, g0.copy(), wb, seed, 0, n, 0.3, 0.07, frozen=frozen)
    assert np.array_equal(got[frozen], g0[frozen])
    assert np.array_equal(got[~frozen], free[~frozen])
