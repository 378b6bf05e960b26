"""Batched execution: one captured hipGraph advancing several engines
(different L, layouts, temperatures and tasks) must reproduce the same engines
advanced one launch at a time, bit for bit (the random streams are
counter-based); and the autotuner's choice is a valid launch shape."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _groups():
    # (L, R, method, beta, h, check)
    return [(16, 128, 2, 0.52, 0.08, 1), (32, 64, 1, 0.58, 0.05, 2), (24, 96, 2, 0.55, 0.03, 2),
            (64, 32, 0, 0.50, 0.10, 1)]


def _make(G, groups, seeds):
    engines, rngs = [], []
    for (L, R, m, b, h, chk), sd in zip(groups, seeds):
        rng = G.RngState(R, sd)
        e = G.Engine(L, R, m, rng, b, h)
        rs = np.random.default_rng(sd)
        g = np.where(rs.random((R, L, L)) < 0.3, 1, -1).astype(np.int32)
        e.set_grids(g)
        if chk == 2:
            e.set_fates(np.full(R, -1, dtype=np.int32))
        engines.append(e)
        rngs.append(rng)
    return engines, rngs


def test_graph_replays_match_sequential(gpu, engine_lib):
    G = engine_lib
    groups = _groups()
    seeds = [11, 12, 13, 14]
    nsw, launches = [6, 4, 3, 2], [3, 1, 1, 2]
    dn, up = [-0.5] * 4, [0.4] * 4
    checks = [g[5] for g in groups]
    print("[graph] capture + 5 replays", flush=True)
    eg, rg = _make(G, groups, seeds)
    graph = G.Graph(eg, nsw, checks, dn, up, launches)
    es, rs_ = _make(G, groups, seeds)
    for rep in range(5):
        graph.launch(1)
        for i, e in enumerate(es):
            e.sweep(nsw[i], checks[i], dn[i], up[i])
        for i in range(len(groups)):
            assert np.array_equal(eg[i].counts(), es[i].counts()), (rep, i)
            assert np.array_equal(eg[i].magnetisation(), es[i].magnetisation()), (rep, i)
            if checks[i] == 2:
                assert np.array_equal(eg[i].fates(), es[i].fates())
    for i in range(len(groups)):
        assert np.array_equal(eg[i].get_grids(), es[i].get_grids())
    assert graph.ms() > 0
    graph.close()
    # after the graph, the engine continues the same stream on its own
    for i, (a, b) in enumerate(zip(eg, es)):
        a.sweep(2)
        b.sweep(2)
        assert np.array_equal(a.get_grids(), b.get_grids())
    for e in eg + es:
        e.close()
    for r in rg + rs_:
        r.free()


def test_autotune_picks_valid_shape(gpu, engine_lib):
    G = engine_lib
    print("[autotune]", flush=True)
    for method, L, R in [(2, 32, 1024), (1, 64, 512), (0, 64, 64)]:
        s, w, rate = G.autotune(method, L, R, 2)
        assert s in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512) and w in (1, 2, 4, 8) and rate > 0
        assert G.get_tuning(method, L, R) == (s, w)
