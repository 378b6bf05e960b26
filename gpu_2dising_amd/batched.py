"""Batched uneven workload: several engines (different L, layouts,
temperatures and tasks) advanced together by one captured hipGraph, with launch
shapes and sweeps per launch chosen by the engine's autotuner at each group's
own (beta, h).

The reference runs one lattice size, temperature and task per process
(src/mc_gpu.cu:530-831); this is the engine-side batching of many such runs.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import CHECK_COUNT, CHECK_FATES, Engine, Graph, RngState, autotune, set_conflict_filter, set_tuning

# L, replicas, gpu_method: four lattice sizes on the LDS multispin path and
# one on the HBM byte lattice
CONFIG4_GROUPS = [(32, 8192, 2), (48, 4096, 2), (64, 4096, 2), (96, 2048, 2), (256, 256, 0)]


def droplet_inputs(L: int, n: int = 16) -> np.ndarray:
    """n centred +1 droplets of radius L/10 .. L/5 in a -1 sea (int32 [n, L, L])."""
    r = np.arange(L)
    d2 = (r[:, None] - L // 2) ** 2 + (r[None, :] - L // 2) ** 2
    out = []
    for j in range(n):
        rad = L / 10 + (L / 10) * j / max(n - 1, 1)
        out.append(np.where(d2 <= rad * rad, 1, -1))
    return np.stack(out).astype(np.int32)


@dataclass
class Group:
    L: int
    R: int
    method: int
    beta: float
    h: float
    itask: int
    seed: int
    slots: int = 0
    waves: int = 0
    launch_sweeps: int = 0
    probe_rate: float = 0.0
    engine: Engine = None
    rng: RngState = None
    start: np.ndarray = field(default=None, repr=False)  # int32 [R, L, L] start grids
    ninputs: int = 1


class Batch:
    """Groups drawn as the batched configuration states: beta ~ U[0.50, 0.60] and
    h ~ U[0.03, 0.10] per group (seeded), itask 0 (count replicas above up_thr)
    and itask 1 (sticky fates from droplet inputs) alternating; one replay of the
    graph is `chk` sweeps of every group plus its threshold check."""

    def __init__(self, groups=CONFIG4_GROUPS, chk: int = 20, seed: int = 2026, tune: bool = True,
                 probe_sweeps: int = 4, dn: float = -0.9, up: float = 0.9, shapes=None,
                 conflict_filter=None):
        """`shapes`: optional (slots, waves, sweeps per launch) per group, used
        instead of the tuner's choice (for comparing builds at fixed shapes).
        `conflict_filter`: the method-2 filter of the groups' engines (as
        set_conflict_filter takes it), one for all or one per group; the
        process setting returns to automatic afterwards.  By default the
        shapes are tuned with the row/column masks, and only half-word groups
        whose mark plane is at most 4 KB (L <= 32) then take the plane: it
        doubles a workgroup's LDS, and in the graph fewer workgroups of the
        larger groups fit beside the large-L group's.  At the batched
        configuration's shapes replays took 5.16-5.24 ms with the plane for
        L = 32 only, 5.15-5.50 with none, 5.7-6.0 with L = 48 too, 6.5-7.1
        with L = 96 and 6.1-6.4 with every method-2 group on it, although
        each group alone ran faster with it."""
        auto = conflict_filter is None
        filters = ([1] * len(groups) if auto else
                   list(conflict_filter) if isinstance(conflict_filter, (list, tuple))
                   else [conflict_filter] * len(groups))
        if len(filters) != len(groups):
            raise ValueError("conflict_filter: one value, or one per group")
        try:
            self._build(groups, chk, seed, tune, probe_sweeps, dn, up, shapes, filters, auto)
        finally:
            set_conflict_filter(0)

    def _build(self, groups, chk, seed, tune, probe_sweeps, dn, up, shapes, filters, auto):
        rs = np.random.default_rng(seed)
        self.chk, self.dn, self.up = chk, dn, up
        self.groups: list[Group] = []
        for k, (L, R, m) in enumerate(groups):
            beta, h = float(rs.uniform(0.50, 0.60)), float(rs.uniform(0.03, 0.10))
            g = Group(L=L, R=R, method=m, beta=beta, h=h, itask=k % 2, seed=seed + k)
            set_conflict_filter(filters[k])
            if shapes is not None:
                g.slots, g.waves, g.launch_sweeps = shapes[k]
            elif tune:
                g.slots, g.waves, g.launch_sweeps, g.probe_rate = autotune(m, L, R, probe_sweeps, beta, h, chk)
            else:
                g.launch_sweeps = chk
            self.groups.append(g)
        if auto:  # the mark plane for half-word groups whose plane is at most 4 KB
            filters = [0 if g.method == 2 and g.waves == 2 and 4 * g.L * g.L <= 4096 else 1
                       for g in self.groups]
        for g, f in zip(self.groups, filters):
            set_conflict_filter(f)
            g.rng = RngState(g.R, g.seed)
            if shapes is not None:
                set_tuning(g.slots, g.waves)
            try:
                g.engine = Engine(g.L, g.R, g.method, g.rng, g.beta, g.h)
            finally:
                if shapes is not None:
                    set_tuning(0, 0)
            if shapes is not None:
                g.probe_rate = 1.0  # capture order by the groups' attempts
            if g.itask == 0:
                g.start = -np.ones((g.R, g.L, g.L), dtype=np.int32)
                g.engine.set_grids(g.start)
            else:
                ins = droplet_inputs(g.L)
                g.ninputs = ins.shape[0]
                g.start = np.repeat(ins, g.R // g.ninputs, axis=0)
                g.engine.set_inputs(ins)
                g.engine.set_fates(np.full(g.R, -1, dtype=np.int32))
        # branches captured longest first (each group's time alone estimated
        # from its tuner probe rate), so that the long groups start before the
        # short ones fill the CUs; except that an HBM-path group whose replicas
        # each take a workgroup with a large LDS conflict table does not go
        # first: dispatched first, those workgroups kept the next group's small
        # ones from spreading evenly over the CUs
        est = [g.L * g.L * g.R / max(g.probe_rate, 1.0) for g in self.groups]
        order = sorted(range(len(self.groups)), key=lambda i: -est[i])
        big_lds = [g.method == 0 and g.slots > 64 for g in self.groups]
        if len(order) > 1 and big_lds[order[0]] and not big_lds[order[1]]:
            order[0], order[1] = order[1], order[0]
        self.order = order
        gs = [self.groups[i] for i in order]
        self.graph = Graph([g.engine for g in gs], [chk] * len(gs),
                           [CHECK_COUNT if g.itask == 0 else CHECK_FATES for g in gs],
                           [dn] * len(gs), [up] * len(gs),
                           [chk // max(g.launch_sweeps, 1) for g in gs])

    @property
    def attempts_per_replay(self) -> int:
        return sum(g.L * g.L * g.R for g in self.groups) * self.chk

    def close(self) -> None:
        self.graph.close()
        for g in self.groups:
            g.engine.close()
            g.rng.free()
