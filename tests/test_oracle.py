"""Pins the CPU oracle (oracle/gasp_oracle.c) against
  - golden vectors produced by the reference's own compiled C code
    (tests/golden/golden_ref.npz, made by tests/golden/make_golden.py),
  - independent implementations (numpy's MT19937 with init_genrand seeding,
    the C++ standard's MT19937 value, scipy's connected-component labelling),
  - what the mathematics fixes (closed forms checked by brute force, the
    Boltzmann distribution of a 3x3 lattice by exact enumeration, detailed
    balance of the acceptance table, Random123 Philox known answers).
"""
import ctypes as C
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_ref.npz")
BASE_SEED = 2894203475


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def unpack(bits, n):
    return np.where(np.unpackbits(bits, bitorder="little")[:n] == 1, 1, -1).astype(np.int32)


# ------------------------------------------------------------------ MT19937
@pytest.mark.parametrize("seed", [5489, BASE_SEED, 1, 4357])
def test_mt_matches_reference_build(oracle, gold, seed):
    s = oracle.mt_stream(seed, 100000)
    assert np.array_equal(s[:2000], gold[f"mt_{seed}"])
    assert s[-1] == gold[f"mt_{seed}_100000th"][0]


def test_mt_matches_independent_implementations(oracle):
    # C++ standard [rand.predef]: 10000th output of mt19937 from seed 5489
    assert oracle.mt_stream(5489, 10000)[-1] == 4123659995
    for seed in (0, 12345, 0xFFFFFFFF):
        bg = np.random.MT19937()
        bg._legacy_seeding(seed)
        assert np.array_equal(oracle.mt_stream(seed, 5000), bg.random_raw(5000).astype(np.uint32))


# ------------------------------------------------------- exact integer maps
def test_real3_floor_closed_form():
    """floor(L * (u + 0.5) 2^-32) == ((2u + 1) L) >> 33 for every L and u tried,
    including every u adjacent to a cell boundary k 2^32 / L."""
    rs = np.random.default_rng(0)
    for L in list(range(1, 130)) + [255, 256, 1000, 4096, 8191, 8192]:
        u = rs.integers(0, 2**32, 20000, dtype=np.uint64)
        k = np.arange(L + 1, dtype=np.uint64)
        edge = (k * (2**32)) // np.uint64(L)
        u = np.concatenate([u, edge, np.maximum(edge, 1) - 1, np.minimum(edge + 1, 2**32 - 1)])
        u = np.clip(u, 0, 2**32 - 1).astype(np.uint64)
        lit = np.floor(L * ((u.astype(np.float64) + 0.5) * (1.0 / 4294967296.0))).astype(np.int64)
        closed = (((2 * u + 1) * np.uint64(L)) >> np.uint64(33)).astype(np.int64)
        assert np.array_equal(lit, closed), L


def test_threshold_exact(oracle):
    """u < T(p)  <=>  (u + 0.5) 2^-32 < p, at the boundary of every threshold."""
    ps = list(oracle.pacc_cpu(0.54, 0.07)) + list(oracle.pacc_gpu(0.54, 0.07).astype(np.float64))
    ps += list(np.random.default_rng(1).random(2000)) + [1e-12, 2.0**-33, 2.0**-32, 0.5, 1 - 2.0**-40, 1.0, 7.0]
    for p in ps:
        T = oracle.lib().orc_threshold(float(p))
        assert 0 <= T <= 2**32
        for u in (T - 2, T - 1, T, T + 1):
            if 0 <= u < 2**32:
                lit = ((float(u) + 0.5) * (1.0 / 4294967296.0)) < p
                assert lit == (u < T), (p, u, T)


def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32 10 rounds
    assert [hex(x) for x in oracle.philox([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(x) for x in oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(x) for x in oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                          [0xa4093822, 0x299f31d0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


# --------------------------------------------------------- acceptance tables
def test_pacc_cpu_matches_reference(oracle, gold):
    for k in range(5):
        beta, h = gold[f"pacc_{k}_beta_h"]
        got = oracle.pacc_cpu(beta, h)
        assert np.array_equal(got.view(np.uint64), gold[f"pacc_{k}"].view(np.uint64))


def _delta_e_bruteforce(s, nup, h):
    """Energy change of flipping the centre of a 3x3 periodic patch computed from
    H = -sum_<ij> s_i s_j - h sum_i s_i over all bonds (independent of the
    5(s+1)+nsum+4 indexing)."""
    L = 3
    g = -np.ones((L, L), dtype=np.int64)
    g[1, 1] = s
    for k, (r, c) in enumerate([(0, 1), (2, 1), (1, 0), (1, 2)]):
        g[r, c] = 1 if k < nup else -1

    def H(x):
        e = -np.sum(x * np.roll(x, 1, 0)) - np.sum(x * np.roll(x, 1, 1))
        return e - h * np.sum(x)

    f = g.copy()
    f[1, 1] = -s
    return H(f) - H(g)


@pytest.mark.parametrize("beta,h", [(0.54, 0.07), (0.3, -0.2), (0.9, 0.5)])
def test_pacc_is_metropolis_of_hamiltonian(oracle, beta, h):
    pd, pf = oracle.pacc_cpu(beta, h), oracle.pacc_gpu(beta, h)
    for s in (-1, 1):
        for nup in range(5):
            nsum = 2 * nup - 4
            idx = 5 * (s + 1) + nsum + 4
            de = _delta_e_bruteforce(s, nup, h)
            assert pd[idx] == pytest.approx(np.exp(-beta * de), rel=1e-15)
            # single precision table of the GPU path (src/mc_gpu.cu:28-29): the
            # argument is formed in float arithmetic, then expf (within 1 ulp)
            f32 = np.float32
            arg = -f32(beta) * (f32(2.0) * f32(s) * (f32(nsum) + f32(h)))
            assert abs(float(pf[idx]) - np.exp(float(arg))) <= np.spacing(f32(np.exp(float(arg))))
            assert float(pf[idx]) == pytest.approx(np.exp(-beta * de), rel=1e-5)


TABLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "acceptance_tables.json")


def test_tables_match_hex_fixtures(oracle):
    """Both acceptance tables and their integer thresholds, bit for bit, against
    tests/golden/acceptance_tables.json (written by make_golden_tables.py with the
    C library's expf / exp on the reference expressions) at the workloads' (beta, h)
    and the batched configuration's five sampled ones."""
    import json

    with open(TABLES) as f:
        fx = json.load(f)["tables"]
    assert len(fx) >= 6
    for t in fx:
        pf = oracle.pacc_gpu(t["beta"], t["h"])
        pd = oracle.pacc_cpu(t["beta"], t["h"])
        assert [x.view(np.uint32).item().to_bytes(4, "big").hex() for x in pf] == t["f32_hex"], t["beta"]
        assert [float(x).hex() for x in pd] == t["f64_hex"], t["beta"]
        assert [int(x) for x in oracle.thresholds(pf.astype(np.float64))] == t["thr_f32"]
        assert [int(x) for x in oracle.thresholds(pd)] == t["thr_f64"]


@pytest.mark.parametrize("L", [3, 16, 20, 64])
def test_production_site_map_uniform(oracle, L):
    """The production Philox stream's sites as the oracle uses them, read back at
    beta = 0 (every attempt flips its site): equal to floor(w N / 2^32) of an
    independent Philox (words 0 / 2 of block (p_lo, p_hi, replica, "GASP") for
    attempts 2p / 2p + 1) and uniform over the sites (chi-square)."""
    from test_oracle_sliced import philox_np

    N, seed, replica, n = L * L, 987654321987, 3, 6000
    g = -np.ones((1, N), dtype=np.int32)
    thr = oracle.thresholds(oracle.pacc_gpu(0.0, 0.0))
    sites = np.empty(n, dtype=np.int64)
    for i in range(n):
        before = g[0].copy()
        oracle.lib().orc_sweep_philox(L, g[0], seed, replica, 500 + i, 1, thr)
        d = np.flatnonzero(g[0] != before)
        assert d.size == 1
        sites[i] = d[0]
    a = 500 + np.arange(n, dtype=np.uint64)
    p = a >> np.uint64(1)
    w = philox_np(p & np.uint64(0xFFFFFFFF), p >> np.uint64(32), np.full(n, replica), np.full(n, 0x47415350),
                  seed & 0xFFFFFFFF, seed >> 32)
    site_word = np.where(a & np.uint64(1), w[2], w[0])
    assert np.array_equal(sites, ((site_word * np.uint64(N)) >> np.uint64(32)).astype(np.int64))
    from scipy import stats

    assert stats.chisquare(np.bincount(sites, minlength=N)).pvalue > 1e-4


# ------------------------------------------------------------------ sweeps
@pytest.mark.parametrize("L", [4, 5, 8, 16, 17, 32, 48, 64])
def test_sweep_matches_reference_build(oracle, gold, L):
    _, nrep, nsweeps, seed = gold[f"sweep_L{L}_meta"]
    g0 = unpack(gold[f"sweep_L{L}_start"], nrep * L * L).reshape(nrep, L, L)
    ref = unpack(gold[f"sweep_L{L}_final"], nrep * L * L).reshape(nrep, L, L)
    got = oracle.sweep_mt_per_replica(int(L), g0.copy(), int(seed), int(nsweeps), 0.54, 0.07)
    assert np.array_equal(got, ref)


def test_plain_run_matches_reference_build(oracle, gold):
    """The plain reference run (L=32, 256 replicas, 2000 sweeps, seeds base+r):
    magnetisation at every check and the final grids, bit for bit."""
    L, nrep, chunks, spc, seed = (int(x) for x in gold["c1_meta"])
    g = -np.ones((nrep, L, L), dtype=np.int32)
    states = (oracle.MTState * nrep)()
    for r in range(nrep):
        oracle.lib().orc_mt_seed(C.byref(states[r]), (seed + r) & 0xFFFFFFFF)
    flat = np.ascontiguousarray(g.reshape(nrep, -1))
    mags = np.zeros((nrep, chunks), dtype=np.float32)
    for c in range(chunks):
        for r in range(nrep):
            mags[r, c] = oracle.lib().orc_magnetisation_cpu(L, flat[r])
        oracle.lib().orc_sweep_batch(L, nrep, flat, 0, C.cast(states, C.c_void_p), 0, 0, spc * L * L,
                                     0.54, 0.07, 8)
    assert np.array_equal(mags, gold["c1_mags"])
    assert np.array_equal(flat.reshape(-1), unpack(gold["c1_final"], nrep * L * L))


def test_boltzmann_3x3_exact_enumeration(oracle):
    """Philox-stream chains of 32 replicas on a 3x3 lattice visit the 512 states
    with their exact Boltzmann weights.  Standard errors come from 40 batch means,
    which measure the chains' autocorrelation (the integrated autocorrelation
    time of the magnetisation is reported) instead of assuming a factor."""
    from test_oracle_sliced import _boltzmann_3x3, _state_index

    L, beta, h, R, seed = 3, 0.35, 0.2, 32, 77
    exact = _boltzmann_3x3(beta, h)
    g = -np.ones((R, L, L), dtype=np.int32)
    oracle.sweep_philox(L, g, seed, 0, 200 * 9, beta, h, nthreads=1)  # burn-in
    nb, per = 40, 500
    freq = np.zeros((nb, 512))
    m_series = []
    a = 200 * 9
    for b in range(nb):
        for _ in range(per):
            oracle.sweep_philox(L, g, seed, a, 9, beta, h, nthreads=1)
            a += 9
            freq[b] += np.bincount(_state_index(g), minlength=512)
            m_series.append(g[0].sum())
        freq[b] /= per * R
    mean = freq.mean(0)
    se = freq.std(0, ddof=1) / np.sqrt(nb)
    big = exact * per * R >= 20
    z = (mean[big] - exact[big]) / se[big]
    assert big.sum() > 50
    assert np.max(np.abs(z)) < 5.5, np.max(np.abs(z))
    assert 0.6 < np.mean(z * z) < 1.6, np.mean(z * z)
    rare = freq[:, ~big].sum(1)
    assert abs(rare.mean() - exact[~big].sum()) < 5 * rare.std(ddof=1) / np.sqrt(nb)
    x = np.array(m_series, dtype=float)
    tau = per * x.reshape(nb, per).mean(1).var(ddof=1) / x.var() / 2
    print(f"tau_int(m) = {tau:.2f} sweeps")
    assert tau < per / 10


def _periodic_largest(grid, spin):
    from scipy import ndimage

    m = (grid == spin)
    lab, n = ndimage.label(m)
    if n == 0:
        return 0
    parent = list(range(n + 1))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    Lr, Lc = grid.shape
    for r in range(Lr):
        a, b = lab[r, 0], lab[r, Lc - 1]
        if a and b:
            parent[find(a)] = find(b)
    for c in range(Lc):
        a, b = lab[0, c], lab[Lr - 1, c]
        if a and b:
            parent[find(a)] = find(b)
    sizes = {}
    for v in lab[lab > 0]:
        rt = find(int(v))
        sizes[rt] = sizes.get(rt, 0) + 1
    return max(sizes.values())



def test_largest_cluster_matches_reference_and_scipy(oracle, gold):
    off = 0
    for i, L in enumerate(gold["clus_Ls"]):
        L = int(L)
        nb = (L * L + 7) // 8
        g = unpack(gold["clus_bits"][off:off + nb], L * L)
        off += nb
        up = oracle.lib().orc_largest_cluster(L, g, 1)
        dn = oracle.lib().orc_largest_cluster(L, g, -1)
        assert up == gold["clus_up"][i] and dn == gold["clus_dn"][i]
        assert up == _periodic_largest(g.reshape(L, L), 1)
        assert dn == _periodic_largest(g.reshape(L, L), -1)


def test_magnetisation_matches_reference(oracle, gold):
    off = 0
    for i, L in enumerate(gold["mag_Ls"]):
        L = int(L)
        nb = (L * L + 7) // 8
        g = unpack(gold["mag_bits"][off:off + nb], L * L)
        off += nb
        assert oracle.lib().orc_magnetisation_cpu(L, g) == gold["mag_vals"][i]
        # GPU form: float sum / float N, exact sum, one rounding
        assert oracle.lib().orc_magnetisation_gpu(L, g) == np.float32(g.sum()) / np.float32(L * L)


# ------------------------------------------------------------------ bootstrap
def test_bootstrap_matches_reference(oracle, gold):
    sizes = gold["boot_sizes"]
    fates = np.split(gold["boot_fates"], np.cumsum(sizes)[:-1])
    oracle.lib().srand(1)
    got = [oracle.lib().orc_bootstrap_errbar(len(f), len(f), np.ascontiguousarray(f)) for f in fates]
    assert np.array_equal(np.array(got).view(np.uint64), gold["boot_vals"].view(np.uint64))
    # statistical sanity: close to the binomial standard error
    for f, e in zip(fates, got):
        res = f[f >= 0]
        p = (res == 1).mean()
        assert e == pytest.approx(np.sqrt(p * (1 - p) / len(f)), rel=0.35)


# ---------------------------------------------------------- driver schedule
def test_driver_schedule_small_cases(oracle):
    """Hand-checkable properties of the restated mc_driver_gpu schedule."""
    L, R = 8, 16
    g = -np.ones((R, L, L), dtype=np.int32)
    out = oracle.driver(L, R, g, rng_kind=1, seed=5, beta=0.54, h=0.07, tot_nsweeps=100,
                        mag_output_int=10, grid_output_int=20, max_dumps=10)
    assert list(out["check_sweeps"]) == list(range(0, 100, 10))
    assert list(out["dump_sweeps"]) == [0, 20, 40, 60, 80]
    # dump 0 is the initial state; the first check sees all -1
    assert np.all(out["dumps"][0] == -1) and np.all(out["cv"][0] == -1.0)
    # itask 0 result = fraction above up_thr at each check
    frac = (out["cv"] > 0.9).mean(axis=1).astype(np.float32)
    assert np.array_equal(out["result"][:len(frac)], frac)
    # all replicas nucleated immediately when up_thr is below the start value
    g = -np.ones((R, L, L), dtype=np.int32)
    out = oracle.driver(L, R, g, rng_kind=1, seed=5, beta=0.54, h=0.07, tot_nsweeps=100,
                        mag_output_int=10, grid_output_int=10, up_thr=-2.0)
    assert np.all(out["result"] == 1.0) and len(out["check_sweeps"]) == 1
    # itask 1: thresholds that every replica crosses at the first check
    g = np.ones((R, L, L), dtype=np.int32)
    fates = np.full(R, -1, dtype=np.int32)
    out = oracle.driver(L, R, g, rng_kind=1, seed=5, beta=0.54, h=0.07, tot_nsweeps=100,
                        mag_output_int=10, grid_output_int=10, itask=1, up_thr=0.5, dn_thr=-0.5,
                        ninputs=2, fates=fates)
    assert np.all(out["fates"] == 1) and np.array_equal(out["result"], np.ones(2, dtype=np.float32))
