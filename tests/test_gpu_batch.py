"""Batched execution: one captured hipGraph advancing several engines
(different L, layouts, temperatures and tasks).  Every replay is checked
against the oracle's restatement of the sweeps on the same Philox streams
(src/mc_cpu.c:41-89 arithmetic with the single-precision table of
src/mc_gpu.cu:21-37), replica for replica, including fates and counts; the
graph also has to continue the chain correctly when direct sweeps of the
same engines come between its replays; the batched configuration's five
groups run at their full shapes; and the autotuner returns valid shapes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DN, UP = -0.5, 0.4


def _groups():
    # (L, R, method, beta, h, check)
    return [(16, 128, 2, 0.52, 0.08, 1), (32, 64, 1, 0.58, 0.05, 2), (24, 96, 2, 0.55, 0.03, 2),
            (64, 32, 0, 0.50, 0.10, 1)]


def _start(L, R, sd):
    rs = np.random.default_rng(sd)
    return np.where(rs.random((R, L, L)) < 0.3, 1, -1).astype(np.int32)


def _make(G, groups, seeds):
    engines, rngs = [], []
    for (L, R, m, b, h, chk), sd in zip(groups, seeds):
        rng = G.RngState(R, sd)
        e = G.Engine(L, R, m, rng, b, h)
        e.set_grids(_start(L, R, sd))
        if chk == 2:
            e.set_fates(np.full(R, -1, dtype=np.int32))
        engines.append(e)
        rngs.append(rng)
    return engines, rngs


class OracleGroup:
    """The oracle's state of one group: all replicas, attempt counter, fates."""

    def __init__(self, O, L, R, beta, h, seed, start, check):
        self.O, self.L, self.R, self.beta, self.h, self.seed = O, L, R, beta, h, seed
        self.g = start.copy()
        self.a = 0
        self.check = check
        self.fates = np.full(R, -1, dtype=np.int32)

    def sweep(self, nsweeps):
        n = nsweeps * self.L * self.L
        self.O.sweep_philox(self.L, self.g, self.seed, self.a, n, self.beta, self.h)
        self.a += n
        m = self.O.magnetisation_gpu(self.g).astype(np.float64)
        counts = np.zeros(3, dtype=np.int32)
        if self.check == 1:
            counts[0] = int((m > UP).sum())
        elif self.check == 2:
            und = self.fates == -1
            self.fates[und & (m > UP)] = 1
            self.fates[und & (m < DN) & ~(m > UP)] = 0
            counts[1], counts[2] = int((self.fates == 0).sum()), int((self.fates == 1).sum())
        return m.astype(np.float32), counts


def test_graph_replays_match_oracle(gpu, engine_lib, oracle):
    """Five replays of a four-group graph (launches of 2-3 sweeps, checks at
    the end of each replay): every replica's grid, magnetisation, fate and
    the counts equal the oracle's after every replay."""
    G = engine_lib
    groups = _groups()
    seeds = [11, 12, 13, 14]
    nsw, launches = [6, 4, 3, 2], [3, 1, 1, 2]
    checks = [g[5] for g in groups]
    print("[graph] capture + 5 replays vs oracle", flush=True)
    eg, rg = _make(G, groups, seeds)
    graph = G.Graph(eg, nsw, checks, [DN] * 4, [UP] * 4, launches)
    orc = [OracleGroup(oracle, L, R, b, h, sd, _start(L, R, sd), chk)
           for (L, R, m, b, h, chk), sd in zip(groups, seeds)]
    for rep in range(5):
        graph.launch(1)
        for i, e in enumerate(eg):
            m, counts = orc[i].sweep(nsw[i])
            assert np.array_equal(e.magnetisation(), m), (rep, i)
            assert np.array_equal(e.counts(), counts), (rep, i)
            if checks[i] == 2:
                assert np.array_equal(e.fates(), orc[i].fates), (rep, i)
            assert np.array_equal(e.get_grids(), orc[i].g), (rep, i)
    assert graph.ms() > 0
    graph.close()
    for e in eg:
        e.close()
    for r in rg:
        r.free()


def test_graph_after_direct_sweeps(gpu, engine_lib, oracle):
    """Direct sweeps of an engine between graph replays (left running on the
    engine's stream when the graph is launched) advance the same chain: the
    graph continues from where they stopped, so the result equals the oracle's
    one sequential stream."""
    G = engine_lib
    groups = _groups()
    seeds = [21, 22, 23, 24]
    nsw = [4, 2, 3, 2]
    checks = [g[5] for g in groups]
    print("[graph] replays interleaved with direct sweeps vs oracle", flush=True)
    eg, rg = _make(G, groups, seeds)
    graph = G.Graph(eg, nsw, checks, [DN] * 4, [UP] * 4)
    orc = [OracleGroup(oracle, L, R, b, h, sd, _start(L, R, sd), chk)
           for (L, R, m, b, h, chk), sd in zip(groups, seeds)]
    plan = ["graph", "direct", "graph", "graph", "direct", "direct", "graph"]
    for step in plan:
        if step == "graph":
            graph.launch(1)
            for i in range(4):
                orc[i].sweep(nsw[i])
        else:
            for i, e in enumerate(eg):
                e.sweep(i + 1, checks[i], DN, UP)  # asynchronous: not waited for here
                orc[i].sweep(i + 1)
    for i, e in enumerate(eg):
        assert np.array_equal(e.get_grids(), orc[i].g), i
        if checks[i] == 2:
            assert np.array_equal(e.fates(), orc[i].fates), i
    graph.close()
    for e in eg:
        e.close()
    for r in rg:
        r.free()


def test_batched_configuration_full_shape(gpu, engine_lib, oracle):
    """The batched configuration at its real shapes: L = 32, 48, 64, 96 (LDS
    multispin, the mark plane) with 8192, 4096, 4096, 2048 replicas and L = 256
    (method 0, its lattice staged in LDS as bits beside 8192-entry dependence
    tables) with 256, beta and h drawn per group, itask 0 and 1 alternating, launch shapes and
    sweeps per launch from the autotuner, three replays of 20 sweeps.  For 12
    replicas of every group (first, last and random ones) the oracle replays
    the same Philox streams: grids, magnetisations and fates must be equal; the
    counts must equal the engine's own fates and magnetisations."""
    from gpu_2dising_amd.batched import Batch

    G = engine_lib
    print("[batched] five groups at full shape, autotuned", flush=True)
    b = Batch(chk=20, seed=2026)
    # the mark plane for every method-2 group whose shape takes it (split words, 2 to 32
    # slots, quarter and eighth words from 8, a plane of at most 32 KB); the L = 256
    # lattice is staged in LDS as bits
    def plane(g):
        return (g.method == 2 and g.waves in (2, 4, 8) and 2 <= g.slots <= 32 and
                (g.waves == 2 or g.slots >= 8) and (8 // g.waves) * g.L * g.L <= 32768)
    want = [2 if plane(g) else 1 for g in b.groups]
    assert [g.engine.conflict_filter() for g in b.groups] == want
    assert [g.engine.lattice_staging() for g in b.groups if g.method == 0] == [3]
    replays = 3
    b.graph.launch(replays)
    rs = np.random.default_rng(5)
    for g in b.groups:
        print(f"[batched] oracle replay of L={g.L} R={g.R} method={g.method} slots={g.slots} "
              f"split/waves={g.waves} sweeps/launch={g.launch_sweeps}", flush=True)
        idx = np.unique(np.concatenate([[0, g.R - 1], rs.choice(g.R, 10, replace=False)]))
        grids = g.start[idx].copy()
        fates = np.full(idx.size, -1, dtype=np.int32)
        N = g.L * g.L
        for rep in range(replays):
            oracle.sweep_philox_replicas(g.L, grids, idx, g.seed, rep * b.chk * N, b.chk * N, g.beta, g.h)
            m = oracle.magnetisation_gpu(grids).astype(np.float64)
            und = fates == -1
            fates[und & (m > b.up)] = 1
            fates[und & (m < b.dn) & ~(m > b.up)] = 0
        got = g.engine.gather(idx).astype(np.int32)
        assert np.array_equal(got, grids), g.L
        mag = g.engine.magnetisation()
        assert np.array_equal(mag[idx], oracle.magnetisation_gpu(grids)), g.L
        counts = g.engine.counts()
        if g.itask == 1:
            f = g.engine.fates()
            assert np.array_equal(f[idx], fates), g.L
            assert counts[1] == int((f == 0).sum()) and counts[2] == int((f == 1).sum())
        else:
            assert counts[0] == int((mag.astype(np.float64) > b.up).sum())
    b.close()


def test_batched_refined_graph(gpu, engine_lib, oracle):
    """Batch with the coordinate search on the captured graph: the chosen
    shapes are valid and no slower than the tuner's by the search's own
    timings (the graph rebuilt at them is timed again, over two replays here,
    so that figure is only bounded loosely), and after the search every group
    starts again from attempt 0 of its streams: two replays, then for 6
    replicas of every group the oracle replays the same Philox streams from
    the start state."""
    from gpu_2dising_amd.batched import Batch

    G = engine_lib
    print("[batched] refined on the graph", flush=True)
    b = Batch(chk=20, seed=2026, refine=True, refine_replays=2)
    try:
        assert b.tuned_ms is not None and b.search_ms <= b.tuned_ms
        assert b.refined_ms < 2.0 * b.tuned_ms  # a slow graph instance replays 1.5 times slower
        for g in b.groups:
            assert g.engine.shape() == (g.slots, g.waves), (g.L, g.engine.shape(), g.slots, g.waves)
        b.graph.launch(2)
        rs = np.random.default_rng(11)
        for g in b.groups:
            print(f"[batched refined] L={g.L} shape=({g.slots}, {g.waves}, {g.launch_sweeps})", flush=True)
            idx = np.unique(np.concatenate([[0, g.R - 1], rs.choice(g.R, 4, replace=False)]))
            grids = g.start[idx].copy()
            N = g.L * g.L
            oracle.sweep_philox_replicas(g.L, grids, idx, g.seed, 0, 2 * b.chk * N, g.beta, g.h)
            assert np.array_equal(g.engine.gather(idx).astype(np.int32), grids), g.L
    finally:
        b.close()


def test_batched_fixed_shapes(gpu, engine_lib, oracle):
    """Batch with explicit (slots, waves, sweeps per launch) per group instead
    of the tuner's: each engine runs at the given shape, and its replicas
    follow the oracle."""
    from gpu_2dising_amd.batched import Batch

    G = engine_lib
    shapes = [(8, 2, 20), (32, 8, 20), (16, 4, 20), (32, 8, 20), (1024, 16, 10)]
    print("[batched] fixed shapes", shapes, flush=True)
    b = Batch(chk=20, seed=2026, shapes=shapes)
    rs = np.random.default_rng(9)
    try:
        for g, (s_, w_, _) in zip(b.groups, shapes):
            assert g.engine.shape() == (s_, w_), (g.L, g.engine.shape())
            g.engine.sweep(b.chk, G.CHECK_COUNT if g.itask == 0 else G.CHECK_FATES, b.dn, b.up)
            g.engine.sync()
            idx = np.unique(np.concatenate([[0, g.R - 1], rs.choice(g.R, 6, replace=False)]))
            grids = g.start[idx].copy()
            oracle.sweep_philox_replicas(g.L, grids, idx, g.seed, 0, b.chk * g.L * g.L, g.beta, g.h)
            assert np.array_equal(g.engine.gather(idx).astype(np.int32), grids), g.L
    finally:
        b.close()


def test_autotune_picks_valid_shape(gpu, engine_lib):
    G = engine_lib
    print("[autotune]", flush=True)
    for method, L, R in [(2, 32, 1024), (1, 64, 512), (0, 64, 64)]:
        s, w, k, rate = G.autotune(method, L, R, 2, 0.55, 0.05, 20)
        assert s in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024) and w in (1, 2, 4, 8, 16) and rate > 0
        assert k in (20, 10, 5) and 20 % k == 0
        assert G.get_tuning(method, L, R) == (s, w)
