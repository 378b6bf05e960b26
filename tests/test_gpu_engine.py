"""Bit-exact parity of the HIP sweep kernels with the CPU oracle.

Every lattice layout (gpu_method 0/1/2, method 2 with whole or split 64-bit
site words) and every speculation width S is run
through the C ABI (gasp_engine_*) and compared spin for spin with
  - oracle.sweep_mt_per_replica: src/mc_cpu.c:41-89 restated, per-replica
    MT19937 streams seeded seed + r (the reference CPU stream), and
  - oracle.sweep_philox: the production Philox4x32-10 stream.
Random +-1 starts at beta = 0.3 accept about half the attempts, so the
speculative rounds hit conflicts often; uniform starts at the reference's
beta = 0.54, h = 0.07 exercise the metastable regime.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BETA, H = 0.54, 0.07
# (method, slots per replica S, second tuning value: waves for 0/1, word split for 2:
# 1, 2, 4, 8 = 64, 32, 16, 8 replicas per workgroup)
# Method 0 with 128 / 256 / 512 / 1024 slots spreads a replica over 2 / 4 / 8 / 16 waves (sweep_hbm_waves).
CASES = ([(2, s, 1) for s in (1, 2, 4, 8, 16)] + [(2, s, 2) for s in (1, 4, 16, 32)] +
         [(2, s, 4) for s in (1, 8, 32)] + [(2, s, 8) for s in (2, 16, 32)] +
         [(1, s, 1) for s in (1, 2, 8, 64)] + [(0, s, 1) for s in (1, 4, 64)] + [(0, 128, 2), (0, 256, 4), (0, 512, 8), (0, 1024, 16)])


def _start(kind, R, L, seed):
    if kind == "down":
        return -np.ones((R, L, L), dtype=np.int32)
    rs = np.random.default_rng(seed)
    return np.where(rs.random((R, L, L)) < 0.5, 1, -1).astype(np.int32)


def _run_engine(G, L, R, method, S, g0, seed, mt, beta, h, chunks, W=1):
    G.set_tuning(S, W)
    try:
        rng = G.RngState(R, seed, mt=mt)
        with G.Engine(L, R, method, rng, beta, h) as e:
            e.set_grids(g0)
            for n in chunks:
                e.sweep(n)
            out = e.get_grids()
            mag = e.magnetisation()
        rng.free()
    finally:
        G.set_tuning(0, 0)
    return out, mag


@pytest.mark.parametrize("method,S,W", CASES)
@pytest.mark.parametrize("L", [16, 17])
def test_mt_stream_bit_exact(gpu, engine_lib, oracle, method, S, W, L):
    G, O = engine_lib, oracle
    R, seed, beta = 96, 1234567, 0.3
    print(f"[mt] method={method} S={S} W={W} L={L} R={R}", flush=True)
    g0 = _start("random", R, L, seed)
    got, mag = _run_engine(G, L, R, method, S, g0, seed, True, beta, 0.07, [3, 2], W)
    ref = O.sweep_mt_per_replica(L, g0.copy(), seed, 5, beta, 0.07)
    assert np.array_equal(got, ref)
    ref_m = np.array([O.lib().orc_magnetisation_gpu(L, ref[r].reshape(-1)) for r in range(R)],
                     dtype=np.float32)
    assert np.array_equal(mag, ref_m)


@pytest.mark.parametrize("method,S,W", CASES)
@pytest.mark.parametrize("start", ["random", "down"])
def test_philox_stream_bit_exact(gpu, engine_lib, oracle, method, S, W, start):
    G, O = engine_lib, oracle
    L, R, seed = 16, 100, 987654321
    beta = 0.3 if start == "random" else BETA
    print(f"[philox] method={method} S={S} W={W} start={start}", flush=True)
    g0 = _start(start, R, L, seed)
    got, mag = _run_engine(G, L, R, method, S, g0, seed, False, beta, H, [4, 3], W)
    ref = O.sweep_philox(L, g0.copy(), seed, 0, 7 * L * L, beta, H)
    assert np.array_equal(got, ref)
    ref_m = (ref.reshape(R, -1).sum(1) / np.float32(L * L)).astype(np.float32)
    assert np.array_equal(mag, ref_m)


@pytest.mark.parametrize("method", [0, 1, 2])
def test_methods_agree_at_L64(gpu, engine_lib, oracle, method):
    """The reference's L=64 at the default tuning: same trajectory in every layout."""
    G, O = engine_lib, oracle
    L, R, seed = 64, 128, 2894203475
    print(f"[L64] method={method}", flush=True)
    g0 = _start("down", R, L, seed)
    got, _ = _run_engine(G, L, R, method, 0, g0, seed, False, BETA, H, [10])
    ref = O.sweep_philox(L, g0.copy(), seed, 0, 10 * L * L, BETA, H)
    assert np.array_equal(got, ref)


def test_split_words_at_L64(gpu, engine_lib, oracle):
    """Split 64-bit words (32 replicas per workgroup) at the widest
    speculation, L = 64, with a partial last group (R = 100)."""
    G, O = engine_lib, oracle
    L, R, seed = 64, 100, 2894203475
    print("[L64 split] S=32", flush=True)
    g0 = _start("down", R, L, seed)
    got, mag = _run_engine(G, L, R, 2, 32, g0, seed, False, BETA, H, [6, 4], W=2)
    ref = O.sweep_philox(L, g0.copy(), seed, 0, 10 * L * L, BETA, H)
    assert np.array_equal(got, ref)
    assert np.array_equal(mag, (ref.reshape(R, -1).sum(1) / np.float32(L * L)).astype(np.float32))


@pytest.mark.parametrize("filt", [1, 2])
@pytest.mark.parametrize("W,S", [(2, 2), (2, 4), (2, 8), (2, 16), (2, 32), (4, 8), (4, 16), (4, 32), (8, 8), (8, 16), (8, 32)])
def test_conflict_filters_bit_exact(gpu, engine_lib, oracle, filt, W, S):
    """Method 2 on split words (W = 2, 4, 8: half, quarter, eighth words)
    with either conflict filter (1 row/column masks, 2 the mark plane; on
    quarter and eighth words for Philox streams only): random starts at
    beta = 0.3, where half the attempts accept and conflicts are common, and
    the uniform start at the reference's (beta, h); L = 17 (N odd, so the
    second launch starts at an odd attempt and its first round with a
    disabled attempt); L = 64 with a partial last word group; L = 72, whose
    mark plane sits 8 to 32 KB above the lattice by the layout."""
    G, O = engine_lib, oracle
    G.set_conflict_filter(filt)
    marks = filt == 2

    def run(L, R, g0, seed, mt, beta, chunks):
        G.set_tuning(S, W)
        try:
            rng = G.RngState(R, seed, mt=mt)
            with G.Engine(L, R, 2, rng, beta, H) as e:
                assert e.shape() == (S, W)
                assert e.conflict_filter() == (2 if marks and (W == 2 or not mt) else 1)
                e.set_grids(g0)
                for n in chunks:
                    e.sweep(n)
                out, mag = e.get_grids(), e.magnetisation()
            rng.free()
        finally:
            G.set_tuning(0, 0)
        return out, mag

    try:
        for L in (16, 17):
            R, seed = 96, 1234567
            print(f"[filter mt] filter={filt} W={W} S={S} L={L}", flush=True)
            g0 = _start("random", R, L, seed)
            got, _ = run(L, R, g0, seed, True, 0.3, [3, 2])
            assert np.array_equal(got, O.sweep_mt_per_replica(L, g0.copy(), seed, 5, 0.3, H))
        for L, R, start, beta, chunks in ((16, 100, "random", 0.3, [4, 3]), (16, 100, "down", BETA, [4, 3]),
                                          (17, 100, "random", 0.3, [3, 2]), (64, 100, "down", BETA, [6, 4]),
                                          (72, 40, "random", 0.3, [1, 1])):
            seed = 987654321
            print(f"[filter philox] filter={filt} W={W} S={S} L={L} start={start}", flush=True)
            g0 = _start(start, R, L, seed)
            got, mag = run(L, R, g0, seed, False, beta, chunks)
            ref = O.sweep_philox(L, g0.copy(), seed, 0, sum(chunks) * L * L, beta, H)
            assert np.array_equal(got, ref)
            assert np.array_equal(mag, (ref.reshape(R, -1).sum(1) / np.float32(L * L)).astype(np.float32))
    finally:
        G.set_conflict_filter(0)


def test_conflict_filter_choice(gpu, engine_lib):
    """The mark plane needs 4 N <= 32768 (L <= 90) and is chosen by default
    only where its doubled LDS leaves enough workgroups resident: at L = 64
    for 16384 replicas (two workgroups per CU), not for 524288 replicas on
    8 slots (sixty-four workgroups per CU to place).  Whole words, other
    methods and one slot per replica have no choice."""
    G = engine_lib

    def filt(L, R, method=2, S=16, W=2):
        G.set_tuning(S, W)
        try:
            rng = G.RngState(R, 5)
            with G.Engine(L, R, method, rng, BETA, H) as e:
                return e.conflict_filter()
        finally:
            G.set_tuning(0, 0)
            rng.free()

    try:
        G.set_conflict_filter(2)
        assert (filt(90, 64), filt(91, 64)) == (2, 1)
        # eighth words: a byte per site, so the plane fits up to L = 181
        assert (filt(181, 64, S=16, W=8), filt(182, 64, S=16, W=8)) == (2, 1)
        assert filt(64, 64, S=4, W=4) == 1  # quarter and eighth words from 8 slots
        G.set_conflict_filter(0)
        assert filt(64, 16384) == 2
        assert filt(64, 524288, S=8) == 1
        assert filt(64, 64, S=16, W=1) == 1
        assert filt(64, 64, S=1) == 1
        assert filt(64, 64, method=1, S=8, W=1) == 1
        assert filt(64, 2048, S=16, W=4) == 1  # quarter words at L <= 64 below 32 slots: masks
        assert filt(64, 2048, S=32, W=4) == 2
        assert filt(96, 2048, S=16, W=8) == 2
    finally:
        G.set_conflict_filter(0)


def test_lattice_staging_choice(gpu, engine_lib):
    """Automatic staging of method 0 over several waves on a 160 KB-LDS
    device: as bits where that leaves the dependence tables more entries than
    bytes (from L = 179 at one workgroup per CU; at L = 128 once four
    workgroups share a CU; L = 512, where bytes do not fit), else as bytes;
    1, 2 and 3 force HBM, bytes and bits."""
    G = engine_lib

    def stg(L, R, S=512):
        G.set_tuning(S, S // 64)
        rng = G.RngState(R, 5)
        try:
            with G.Engine(L, R, 0, rng, BETA, H) as e:
                return e.lattice_staging()
        finally:
            G.set_tuning(0, 0)
            rng.free()

    assert (stg(178, 64), stg(179, 64)) == (2, 3)
    assert stg(64, 64) == 2
    assert (stg(128, 256, S=256), stg(128, 1024, S=256)) == (2, 3)
    assert stg(512, 64) == 3
    try:
        for forced, (L, R) in ((1, (256, 64)), (2, (256, 64)), (3, (64, 64))):
            G.set_lattice_staging(forced)
            assert stg(L, R) == forced
    finally:
        G.set_lattice_staging(0)
    # with at most 8192 entries per table, bits no longer leave more entries at L = 179
    try:
        G.set_dependence_tables(8192)
        assert stg(179, 64) == 2
        assert stg(512, 64) == 3
    finally:
        G.set_dependence_tables(0)


@pytest.mark.parametrize("mt", [False, True])
def test_dependence_table_limit_bit_exact(gpu, engine_lib, oracle, mt):
    """gasp_set_dependence_tables: 64 entries per table for 2025 sites (about
    32 sites share each entry, so most deferrals are spurious) on the
    bits-staged lattice over four waves; the trajectories stay those of the
    sequential chain."""
    G, O = engine_lib, oracle
    L, R, S, seed, beta = 45, 5, 256, 9091, 0.3
    G.set_lattice_staging(3)
    G.set_dependence_tables(64)
    try:
        print(f"[table limit] L={L} S={S} mt={mt}", flush=True)
        g0 = _start("random", R, L, seed)
        got, mag = _run_engine(G, L, R, 0, S, g0, seed, mt, beta, H, [2, 1], W=S // 64)
        if mt:
            ref = O.sweep_mt_per_replica(L, g0.copy(), seed, 3, beta, H)
        else:
            ref = O.sweep_philox(L, g0.copy(), seed, 0, 3 * L * L, beta, H)
        assert np.array_equal(got, ref)
    finally:
        G.set_lattice_staging(0)
        G.set_dependence_tables(0)


def test_set_attempt_rewinds_streams(gpu, engine_lib, oracle):
    """gasp_engine_set_attempt: after sweeps, rewinding to attempt 0 and
    reloading the start grids replays the same trajectory; moving to attempt
    n continues the oracle's stream from there; MT19937 states are refused."""
    G, O = engine_lib, oracle
    L, R, seed = 16, 64, 4321
    g0 = _start("random", R, L, seed)
    rng = G.RngState(R, seed)
    try:
        with G.Engine(L, R, 2, rng, 0.3, H) as e:
            e.set_grids(g0)
            e.sweep(3)
            first = e.get_grids()
            e.set_attempt(0)
            e.set_grids(g0)
            e.sweep(3)
            assert np.array_equal(e.get_grids(), first)
            e.set_attempt(7 * L * L)
            e.set_grids(g0)
            e.sweep(2)
            assert np.array_equal(e.get_grids(), O.sweep_philox(L, g0.copy(), seed, 7 * L * L, 2 * L * L, 0.3, H))
    finally:
        rng.free()
    mt = G.RngState(R, seed, mt=True)
    try:
        with G.Engine(L, R, 2, mt, 0.3, H) as e:
            with pytest.raises(G.GaspError):
                e.set_attempt(0)
    finally:
        mt.free()


@pytest.mark.parametrize("staging", [1, 2, 3])
@pytest.mark.parametrize("S", [128, 256, 512, 1024])
def test_hbm_waves_at_L256(gpu, engine_lib, oracle, S, staging):
    """A replica spread over 2, 4, 8 or 16 waves at the batched run's L = 256
    shape (R = 6, so the conflict table has fewer entries than sites and
    sites share entries), with the byte lattice in HBM (staging 1) or staged
    in LDS as bytes (2) or bits (3), from all -1 at the reference's (beta, h)
    and from a random start at beta = 0.3 (about half the attempts
    accepted)."""
    G, O = engine_lib, oracle
    L, R, seed = 256, 6, 4242
    G.set_lattice_staging(staging)
    try:
        rng = G.RngState(R, 1)
        G.set_tuning(S, S // 64)
        with G.Engine(L, R, 0, rng, BETA, H) as e:
            assert e.lattice_staging() == staging
        G.set_tuning(0, 0)
        rng.free()
        for start, beta in (("down", BETA), ("random", 0.3)):
            print(f"[L256 waves] S={S} staging={staging} start={start}", flush=True)
            g0 = _start(start, R, L, seed)
            got, mag = _run_engine(G, L, R, 0, S, g0, seed, False, beta, H, [1, 2], W=S // 64)
            ref = O.sweep_philox(L, g0.copy(), seed, 0, 3 * L * L, beta, H)
            assert np.array_equal(got, ref)
            assert np.array_equal(mag, (ref.reshape(R, -1).sum(1) / np.float32(L * L)).astype(np.float32))
    finally:
        G.set_lattice_staging(0)


@pytest.mark.parametrize("mt", [False, True])
@pytest.mark.parametrize("L,S", [(17, 128), (45, 512), (64, 256)])
def test_hbm_waves_bits_staged(gpu, engine_lib, oracle, L, S, mt):
    """The byte lattice staged in LDS as bits (staging 3) on lattices whose
    site count is not a multiple of 32 (the last dword is partial: 289 and
    2025 sites) and on one that is, with both RNG streams, from a random start
    at beta = 0.3."""
    G, O = engine_lib, oracle
    R, seed, beta = 7, 31337, 0.3
    G.set_lattice_staging(3)
    try:
        rng = G.RngState(R, 1)
        G.set_tuning(S, S // 64)
        with G.Engine(L, R, 0, rng, beta, H) as e:
            assert e.lattice_staging() == 3
        G.set_tuning(0, 0)
        rng.free()
        print(f"[bits staged] L={L} S={S} mt={mt}", flush=True)
        g0 = _start("random", R, L, seed)
        got, mag = _run_engine(G, L, R, 0, S, g0, seed, mt, beta, H, [2, 1], W=S // 64)
        if mt:
            ref = O.sweep_mt_per_replica(L, g0.copy(), seed, 3, beta, H)
        else:
            ref = O.sweep_philox(L, g0.copy(), seed, 0, 3 * L * L, beta, H)
        ref_m = np.array([O.lib().orc_magnetisation_gpu(L, ref[r].reshape(-1)) for r in range(R)],
                         dtype=np.float32)
        assert np.array_equal(got, ref)
        assert np.array_equal(mag, ref_m)
    finally:
        G.set_lattice_staging(0)


def test_hbm_waves_small_table(gpu, engine_lib, oracle):
    """600 replicas of L = 256 over four waves each: the dependence tables
    are sized for three workgroups per CU (4096 entries each on a 160 KB-LDS
    device), so sixteen sites share each entry and spurious dependences are
    common, which must only delay attempts, never change them."""
    G, O = engine_lib, oracle
    L, R, seed = 256, 600, 777
    for start, beta in (("down", BETA), ("random", 0.3)):
        print(f"[L256 small table] start={start}", flush=True)
        g0 = _start(start, R, L, seed)
        got, mag = _run_engine(G, L, R, 0, 256, g0, seed, False, beta, H, [1], W=4)
        ref = O.sweep_philox(L, g0.copy(), seed, 0, L * L, beta, H)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("L", [8, 16, 33, 64])
def test_largest_cluster(gpu, engine_lib, oracle, L):
    G, O = engine_lib, oracle
    R = 40
    print(f"[cluster] L={L}", flush=True)
    rs = np.random.default_rng(L)
    p = np.linspace(0.2, 0.8, R)[:, None, None]
    g0 = np.where(rs.random((R, L, L)) < p, 1, -1).astype(np.int32)
    rng = G.RngState(R, 1)
    with G.Engine(L, R, 1, rng, BETA, H) as e:
        e.set_grids(g0)
        got_up = e.largest_cluster(1)
        got_dn = e.largest_cluster(-1)
    rng.free()
    ref_up = np.array([O.lib().orc_largest_cluster(L, g0[r].reshape(-1), 1) for r in range(R)])
    ref_dn = np.array([O.lib().orc_largest_cluster(L, g0[r].reshape(-1), -1) for r in range(R)])
    assert np.array_equal(got_up, ref_up)
    assert np.array_equal(got_dn, ref_dn)


@pytest.mark.parametrize("method,S,W", [(1, 0, 0), (2, 0, 0), (2, 0, 2), (0, 256, 4)])
def test_inputs_and_freeze_leave_fates_unchanged(gpu, engine_lib, method, S, W):
    """gasp_engine_set_inputs replicates input grids exactly as uploading the
    replicated grids; freezing decided replicas changes no fate at any check,
    and no undecided replica's trajectory."""
    G = engine_lib
    L, nin, per = 16, 4, 24
    R = nin * per
    inputs = np.stack([np.where(np.add.outer(np.arange(L), np.arange(L)) % (k + 3) == 0, 1, -1)
                       for k in range(nin)]).astype(np.int32)
    print(f"[freeze] method={method} S={S} W={W}", flush=True)
    runs = []
    G.set_tuning(S, W)
    for freeze in (False, True):
        rng = G.RngState(R, 99)
        with G.Engine(L, R, method, rng, 0.54, 0.07) as e:
            if freeze:
                e.set_inputs(inputs)
            else:
                e.set_grids(np.repeat(inputs, per, axis=0))
            assert np.array_equal(e.get_grids(), np.repeat(inputs, per, axis=0))
            e.set_freeze(freeze)
            e.set_fates(np.full(R, -1, dtype=np.int32))
            e.sweep(0, G.CHECK_FATES, -0.8, 0.6)
            hist = []
            for _ in range(30):
                e.sweep(5, G.CHECK_FATES, -0.8, 0.6)
                f = e.fates()
                hist.append((e.counts().copy(), f, e.get_grids()))
        rng.free()
        runs.append(hist)
    G.set_tuning(0, 0)
    for (c0, f0, g0), (c1, f1, g1) in zip(*runs):
        assert np.array_equal(c0[1:], c1[1:]) and np.array_equal(f0, f1)
        und = f1 == -1
        assert np.array_equal(g0[und], g1[und])
    assert (runs[1][-1][1] != -1).sum() > 0  # some replicas did freeze


@pytest.mark.parametrize("method", [0, 1, 2])
def test_gather_grids_int8(gpu, engine_lib, method):
    """gasp_engine_gather_grids returns exactly the selected rows of the full
    int32 grids, as int8, for any order and repeats of replica indices."""
    G = engine_lib
    L, R = 24, 130
    print(f"[gather] method={method}", flush=True)
    g0 = _start("random", R, L, 5)
    rng = G.RngState(R, 5)
    with G.Engine(L, R, method, rng, 0.3, 0.0) as e:
        e.set_grids(g0)
        e.sweep(3)
        full = e.get_grids()
        idx = np.array([129, 0, 64, 63, 7, 7, 128, 65], dtype=np.int32)
        got = e.gather(idx)
        with pytest.raises(G.GaspError):
            e.gather([R])
    rng.free()
    assert got.dtype == np.int8
    assert np.array_equal(got, full[idx].astype(np.int8))


@pytest.mark.parametrize("method,L", [(0, 17), (1, 16), (2, 17), (2, 64)])
def test_gridstates_record_body(gpu, engine_lib, tmp_path, method, L):
    """gasp_engine_get_gridstates equals write_ising_grids' bit packing of the
    engine's int32 grids (src/io.c:67-130), also when L*L is not a multiple of 8."""
    import ctypes as C

    G = engine_lib
    R = 70
    print(f"[gridstates] method={method} L={L}", flush=True)
    rng = G.RngState(R, 3)
    with G.Engine(L, R, method, rng, 0.3, 0.0) as e:
        e.set_grids(_start("random", R, L, 9))
        e.sweep(2)
        full = e.get_grids()
        body = e.gridstates()
    rng.free()
    exp = np.packbits((full.reshape(-1) > 0).astype(np.uint8), bitorder="little")
    assert np.array_equal(body, exp)
    # the same bytes as the record write_ising_grids appends (after its 12-byte header)
    lib = G.lib()
    fn = str(tmp_path / "g.bin").encode()
    assert lib.write_ising_grids(L, R, full.ctypes.data_as(C.POINTER(C.c_int)), 0, None, None,
                                 b"magnetisation", -0.9, 0.9, fn) == 0
    assert open(fn, "rb").read()[12:] == body.tobytes()


@pytest.mark.parametrize("method,S,W", [(2, 1, 1), (2, 16, 1), (2, 32, 2), (1, 4, 1), (0, 4, 1), (0, 128, 2)])
@pytest.mark.parametrize("L", [2, 3, 4, 5])
def test_tiny_lattices(gpu, engine_lib, oracle, method, S, W, L):
    """L = 2..5: the up and down (left and right) neighbours coincide at L = 2,
    and every attempt of a round is a conflict-filter candidate.  Both random
    streams, random starts at beta = 0.3."""
    G, O = engine_lib, oracle
    R, seed, beta = 70, 31337, 0.3
    print(f"[tiny] method={method} S={S} W={W} L={L}", flush=True)
    g0 = _start("random", R, L, seed)
    got, _ = _run_engine(G, L, R, method, S, g0, seed, True, beta, H, [7, 5], W)
    assert np.array_equal(got, O.sweep_mt_per_replica(L, g0.copy(), seed, 12, beta, H))
    got, _ = _run_engine(G, L, R, method, S, g0, seed, False, beta, H, [7, 5], W)
    assert np.array_equal(got, O.sweep_philox(L, g0.copy(), seed, 0, 12 * L * L, beta, H))


@pytest.mark.parametrize("L,S,W", [(128, 16, 1), (142, 16, 1), (143, 32, 2)])
def test_method2_largest_lattices(gpu, engine_lib, oracle, L, S, W):
    """The largest lattices of the multispin layouts: whole 64-bit words up to
    L = 142 (161 KB of LDS per workgroup), split words beyond; a partial last
    group (R = 70)."""
    G, O = engine_lib, oracle
    R, seed = 70, 5150
    print(f"[large] L={L} S={S} W={W}", flush=True)
    g0 = _start("random", R, L, seed)
    got, mag = _run_engine(G, L, R, 2, S, g0, seed, False, 0.3, H, [1, 1], W)
    ref = O.sweep_philox(L, g0.copy(), seed, 0, 2 * L * L, 0.3, H)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("method,S,W,L", [(2, 16, 2, 96), (2, 32, 2, 100), (1, 16, 1, 96), (2, 8, 2, 200),
                                           (2, 16, 4, 96), (2, 16, 8, 200), (2, 32, 8, 100),
                                           (2, 16, 8, 300)])  # eighth words past halves' L = 201
@pytest.mark.parametrize("start", ["random", "down"])
def test_folded_filter_large_L(gpu, engine_lib, oracle, method, S, W, L, start):
    """L > 64 not a power of two: the conflict filter folds rows and columns
    mod 64 (an entry is the union over every row with that residue), which
    may only add candidates.  Random starts at beta = 0.3 (many conflicts)
    and all-down starts at beta = 0.54 (the filter skips most rounds)."""
    G, O = engine_lib, oracle
    R, seed = 70, 6061
    beta = 0.3 if start == "random" else BETA
    print(f"[fold] method={method} S={S} W={W} L={L} start={start}", flush=True)
    g0 = _start(start, R, L, seed)
    got, mag = _run_engine(G, L, R, method, S, g0, seed, False, beta, H, [1, 2], W)
    ref = O.sweep_philox(L, g0.copy(), seed, 0, 3 * L * L, beta, H)
    assert np.array_equal(got, ref)


def test_method2_whole_words_refused_past_L142(gpu, engine_lib):
    """At L = 143 one group's 64-bit words exceed the 160 KB of LDS per
    workgroup: the engine reports it instead of launching."""
    G = engine_lib
    print("[large] L=143 whole words", flush=True)
    G.set_tuning(16, 1)
    try:
        rng = G.RngState(64, 1)
        with pytest.raises(G.GaspError):
            G.Engine(143, 64, 2, rng, BETA, H)
        rng.free()
    finally:
        G.set_tuning(0, 0)


@pytest.mark.parametrize("W", [1, 2])
def test_adaptive_slots_keep_trajectories(gpu, engine_lib, W):
    """A freezing engine raises its slots per replica as replicas are decided
    (here from 4 slots at 2^18 replicas).  Its fates at every check and its
    final grids equal those of an engine held at one slot by gasp_set_tuning."""
    G = engine_lib
    L, nin, per = 16, 64, 4096
    R = nin * per
    inputs = np.stack([np.where(np.add.outer(np.arange(L), np.arange(L)) % (k % 7 + 3) == 0, 1, -1)
                       for k in range(nin)]).astype(np.int32)
    print(f"[adaptive] R={R} W={W}", flush=True)
    runs = []
    for slots in (1, 0):  # fixed at one slot, then automatic (adaptive)
        G.set_tuning(slots, W)
        try:
            rng = G.RngState(R, 4321)
            with G.Engine(L, R, 2, rng, 0.54, 0.07) as e:
                e.set_inputs(inputs)
                e.set_freeze(True)
                e.set_fates(np.full(R, -1, dtype=np.int32))
                e.sweep(0, G.CHECK_FATES, -0.8, 0.6)
                e.counts()
                hist = []
                for _ in range(40):
                    e.sweep(5, G.CHECK_FATES, -0.8, 0.6)
                    hist.append((e.counts().copy(), e.fates()))
                grids = e.get_grids()
            rng.free()
        finally:
            G.set_tuning(0, 0)
        runs.append((hist, grids))
    (h0, g0), (h1, g1) = runs
    for (c0, f0), (c1, f1) in zip(h0, h1):
        assert np.array_equal(c0, c1) and np.array_equal(f0, f1)
    assert np.array_equal(g0, g1)
    # most replicas decided, so the adaptive engine ran several slot counts
    assert (h1[-1][1] != -1).mean() > 0.6
