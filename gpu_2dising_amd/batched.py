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

from . import (CHECK_COUNT, CHECK_FATES, Engine, Graph, RngState, autotune, set_conflict_filter,
               set_dependence_tables, set_lattice_staging, set_tuning)

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
                 conflict_filter=None, refine: bool = False, refine_replays: int = 10,
                 lattice_staging: int = 3, dependence_tables: int = 8192):
        """`shapes`: optional (slots, waves, sweeps per launch) per group, used
        instead of the tuner's choice (for comparing builds at fixed shapes).
        `conflict_filter`: the method-2 filter of the groups' engines (as
        set_conflict_filter takes it), one for all or one per group, for the
        tuner's probes and the engines; the process setting returns to
        automatic afterwards.  By default the mark plane for every method-2
        group (an engine takes the row/column masks where the plane does not
        fit or apply).
        `lattice_staging`, `dependence_tables`: the staging of multi-wave
        method-0 groups and the most entries of their dependence tables (as
        set_lattice_staging and set_dependence_tables take them); by default
        bits in LDS beside 8192-entry tables (74.6 KB per workgroup at
        L = 256).
        Together: the mark plane doubles a method-2 workgroup's LDS, and in
        the graph the groups share the CUs' LDS with the L = 256 group.  At
        the batched configuration's refined shapes, with the L = 256 group on
        bytes beside 8192-entry tables (129 KB per workgroup), replays took
        4.72-4.75 ms with the plane for L = 32 only and 6.6-7.2 ms with it on
        L = 48, 64 or 96 as well; with that group on bits beside 8192 entries,
        4.59-4.61 ms with the plane on every method-2 group, 4.70 on L = 32
        and 96, 4.90 on L = 32 only; on bits beside 16384 entries (the
        engine's own choice, 140 KB) 5.57 ms.
        `refine`: after the per-group tuning, a coordinate search over the
        groups' shapes on the captured graph itself (see `_refine`)."""
        filters = ([2 if m == 2 else 1 for (_, _, m) in groups] if conflict_filter is None else
                   list(conflict_filter) if isinstance(conflict_filter, (list, tuple))
                   else [conflict_filter] * len(groups))
        if len(filters) != len(groups):
            raise ValueError("conflict_filter: one value, or one per group")
        self.staging, self.tables = lattice_staging, dependence_tables
        self.refine_log, self.tuned_ms, self.search_ms, self.refined_ms = [], None, None, None
        set_lattice_staging(lattice_staging)
        set_dependence_tables(dependence_tables)
        try:
            self._build(groups, chk, seed, tune, probe_sweeps, dn, up, shapes, filters)
            if refine and shapes is None:
                self._refine(refine_replays)
        finally:
            set_conflict_filter(0)
            set_lattice_staging(0)
            set_dependence_tables(0)
            set_tuning(0, 0)

    def _build(self, groups, chk, seed, tune, probe_sweeps, dn, up, shapes, filters):
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
        self.filters = filters
        for k, g in enumerate(self.groups):
            if shapes is not None:
                g.probe_rate = 1.0  # capture order by the groups' attempts
            self._make_engine(k, fixed=shapes is not None)
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
        self._capture()


    def _make_engine(self, k: int, fixed: bool) -> None:
        """Group k's RNG streams (from attempt 0) and engine, at its shape if
        `fixed` (else the tuner's), loaded with its start state."""
        g = self.groups[k]
        set_conflict_filter(self.filters[k])
        set_lattice_staging(self.staging)
        set_dependence_tables(self.tables)
        g.rng = RngState(g.R, g.seed)
        if fixed:
            set_tuning(g.slots, g.waves)
        try:
            g.engine = Engine(g.L, g.R, g.method, g.rng, g.beta, g.h)
        finally:
            set_tuning(0, 0)
        if g.itask == 0:
            g.start = -np.ones((g.R, g.L, g.L), dtype=np.int32)
        else:
            ins = droplet_inputs(g.L)
            g.ninputs = ins.shape[0]
            g.start = np.repeat(ins, g.R // g.ninputs, axis=0)
        self._load_start(g)

    @staticmethod
    def _load_start(g: Group) -> None:
        if g.itask == 0:
            g.engine.set_grids(g.start)
        else:
            g.engine.set_inputs(droplet_inputs(g.L))
            g.engine.set_fates(np.full(g.R, -1, dtype=np.int32))

    def _capture(self) -> None:
        gs = [self.groups[i] for i in self.order]
        self.graph = Graph([g.engine for g in gs], [self.chk] * len(gs),
                           [CHECK_COUNT if g.itask == 0 else CHECK_FATES for g in gs],
                           [self.dn] * len(gs), [self.up] * len(gs),
                           [self.chk // max(g.launch_sweeps, 1) for g in gs])

    def _time_replay(self, replays: int) -> float:
        """ms per replay from the groups' start states (replays grow slower as
        the lattices nucleate), after one untimed replay."""
        for g in self.groups:
            self._load_start(g)
            g.engine.sync()
        self.graph.launch(1)
        ms0 = self.graph.ms()
        self.graph.launch(replays)
        return (self.graph.ms() - ms0) / replays

    def _candidates(self, g: Group) -> list:
        if g.method == 0 and g.slots > 64:
            return [(s_, s_ // 64, k) for s_ in (256, 512) for k in (self.chk // 4, self.chk // 2, self.chk)
                    if k >= 1 and self.chk % k == 0]
        if g.method != 2:
            return []
        slots = [s_ for s_ in (g.slots // 2, g.slots, 2 * g.slots) if 2 <= s_ <= 32]
        return [(s_, w_, self.chk) for s_ in slots for w_ in (2, 4, 8)]

    def _refine(self, replays: int, tol: float = 0.02, passes: int = 3) -> None:
        """Coordinate search over the groups' shapes on the captured graph: the
        tuner probes each group alone, where wide speculation fills the device,
        while in the graph the groups share it.  For each group in capture
        order, each candidate (method 2: half, the same and twice the slots on
        half, quarter or eighth words, all sweeps in one launch; the multi-wave
        method 0: 4 or 8 waves, launches of a quarter, half or all of the
        sweeps) is timed over `replays` replays from the start states and kept
        if it beats the best so far by `tol` ms; up to `passes` passes, until
        one changes nothing, every other pass visiting the groups in reverse
        order.  The groups' choices interact, so the search runs from two
        starts and keeps the faster end: the tuner's shapes, and the same with
        every method-2 group's sweeps in one launch and every multi-wave
        method-0 group on its widest candidate with the shortest launches.
        At the batched configuration the search from the tuner's shapes
        (L = 32 on 16 slots in two launches, L = 256 on 8 waves in one) ended,
        in three processes, twice at 4.57-4.58 ms (L = 32 on 8 slots, L = 256
        in launches of 5 sweeps) and once at 4.86 ms (L = 256 on 4 waves in
        launches of 10; neither move paid without the other).  Each timing
        uses a checked capture (`_capture_checked`).  The searches advance
        the RNG streams and the lattices, so every group ends at attempt 0 of
        its streams and at its start state."""
        shape = lambda g: (g.slots, g.waves, g.launch_sweeps)
        tuned = [shape(g) for g in self.groups]
        self.graph.close()
        best = self._capture_checked(replays)
        self.tuned_ms = best
        best = self._search(replays, tol, passes, best, 1)
        ends = [(best, [shape(g) for g in self.groups])]
        second = []
        for g, sh in zip(self.groups, tuned):
            if g.method == 2:
                second.append((sh[0], sh[1], self.chk))
            elif g.method == 0 and g.slots > 64:
                second.append(max(self._candidates(g), key=lambda c: (c[0], -c[2])))
            else:
                second.append(sh)
        if second not in (tuned, ends[0][1]):
            self.graph.close()
            for k, g in enumerate(self.groups):
                if shape(g) != second[k]:
                    g.slots, g.waves, g.launch_sweeps = second[k]
                    self._rebuild(k)
            t = self._capture_checked(replays)
            self.refine_log.append({"L": None, "from": "second start", "to": [list(x) for x in second], "ms": t})
            t = self._search(replays, tol, passes, t, 2)
            ends.append((t, [shape(g) for g in self.groups]))
        best, chosen = min(ends, key=lambda e: e[0])
        self.search_ms = best  # at most tuned_ms: each search only keeps faster shapes
        # fresh engines at the chosen shapes, an instance of the graph that
        # replays at the searched speed, then every group back to attempt 0
        # and its start state
        self.graph.close()
        for k, g in enumerate(self.groups):
            g.slots, g.waves, g.launch_sweeps = chosen[k]
            g.engine.close()
            g.rng.free()
            self._make_engine(k, fixed=True)
        self.refined_ms = self._capture_checked(replays, ref=best)
        for g in self.groups:
            g.engine.set_attempt(0)
            self._load_start(g)

    def _search(self, replays: int, tol: float, passes: int, best: float, start: int) -> float:
        """The coordinate passes of `_refine` from the groups' current shapes
        (the graph captured at them); returns the best ms per replay."""
        for p in range(passes):
            changed = False
            for k in (list(self.order) if p % 2 == 0 else list(reversed(self.order))):
                g = self.groups[k]
                for cand in self._candidates(g):
                    old = (g.slots, g.waves, g.launch_sweeps)
                    if cand == old:
                        continue
                    g.slots, g.waves, g.launch_sweeps = cand
                    try:
                        self._rebuild(k)
                        # one recapture at most: most candidates that look slow are
                        t = self._capture_checked(replays, ref=best, tries=2)
                    except RuntimeError:
                        t = None
                    if t is not None and t < best - tol:
                        self.refine_log.append({"L": g.L, "from": list(old), "to": list(cand), "ms": t,
                                                "start": start})
                        best, changed = t, True
                    else:
                        g.slots, g.waves, g.launch_sweeps = old
                        self._rebuild(k)
                        self._capture()
            if not changed:
                break
        return best

    def _capture_checked(self, replays: int, ref: float = None, tries: int = 4) -> float:
        """Captures the graph and times it from the start states.  Instances of
        one graph replay at one of two speeds (at the batched configuration
        4.75 or 7.3 ms, about one capture in five the slower; with 8 hardware
        queues 5.1 or 5.8 ms), so an instance more than 20% slower than `ref`
        (without `ref`: than another capture) is captured again, up to
        `tries` captures.  Without `ref` the reference is the fastest of three
        captures (two slow instances in a row are not rare enough).  Returns
        the kept instance's ms per replay."""
        self._capture()
        t = self._time_replay(replays)
        if ref is None:
            ref = t
            for _ in range(2):
                self.graph.close()
                self._capture()
                t = self._time_replay(replays)
                ref = min(ref, t)
        for _ in range(tries - 1):
            if t <= 1.2 * ref:
                break
            self.graph.close()
            self._capture()
            t = self._time_replay(replays)
        return t

    def _rebuild(self, k: int) -> None:
        """Group k's engine again at its current shape (the graph is closed,
        for the caller to capture again)."""
        g = self.groups[k]
        self.graph.close()
        g.engine.close()
        g.rng.free()
        self._make_engine(k, fixed=True)

    @property
    def attempts_per_replay(self) -> int:
        return sum(g.L * g.L * g.R for g in self.groups) * self.chk

    def close(self) -> None:
        self.graph.close()
        for g in self.groups:
            g.engine.close()
            g.rng.free()
