"""Online autotuner for the windowed update's launch parameters.

The search runs in stages over the parameters that change the launch but
never the results (trees and BC are bit-identical under every setting):

1. team size: threads per workgroup (1024, 512, 256: one root per workgroup
   team, or one per wavefront for the light phase);
2. heavy/light split, only when the caller gives a root order (heaviest
   first): how many roots run as whole-workgroup teams (all, or all but the
   last quarter of the queue, which then run one wavefront each, without hub
   lookups or hand-overs; in index order that quarter holds hub roots too,
   and one wavefront took 1.2 s over such a window on R-MAT scale 20);
3. hand-over thresholds: the descent frontier size (``park_min``) and the
   adjacency entries left to scan (``park_scan``) at which an item is
   finished by the whole-GPU level kernels.

Within a stage the candidates take the stream's leading windows in turn
(round-robin, ``reps`` windows each; the threshold stage takes one more);
each window's device time is the window-kernel launch plus the hand-over
kernels it caused.  Windows differ in their work, so a candidate replaces the
stage's first one (the current setting) only if its median is lower by more
than ``margin`` (the threshold stage: its mean, by more than 15%); otherwise
the current setting stays.  Tuning windows are
real work: nothing is thrown away or replayed.

Results are cached in a JSON file keyed by (vertices, edges, roots, window).
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass, field

import numpy as np

from . import KSTAT_UPDATE, KSTAT_WIDE, BatchConfig, Context, Forest, Graph

BLOCKS = (1024, 512, 256)
# (park_min, park_scan): the library default first; the last one is the
# serving loop's default (power-law graphs whose hubs make many items large)
HANDOVER = ((4096, 512 * 1024), (2048, 256 * 1024), (8192, 1024 * 1024), (4096, 2048 * 1024))

# kept for callers that size the tuning budget: the first stage's candidates
DEFAULT_CANDIDATES = tuple({"block_threads": b, "grid_blocks": 0} for b in BLOCKS)


@dataclass
class TuneResult:
    block_threads: int
    grid_blocks: int
    from_cache: bool
    windows_used: int
    table: list = field(default_factory=list)
    heavy_roots: int = -1
    park_min: int = 0
    park_scan: int = 0

    def config(self, window: int, **kw) -> BatchConfig:
        return BatchConfig(window=window, block_threads=self.block_threads, grid_blocks=self.grid_blocks,
                           heavy_roots=self.heavy_roots, park_min=self.park_min, park_scan=self.park_scan, **kw)


def cache_key(nv: int, ne: int, k: int, window: int) -> str:
    return f"nv={nv},ne={ne},k={k},w={window}"


def _load(path):
    if not path or not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def _store(path, key, entry):
    if not path:
        return
    data = _load(path)
    data[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tuning_windows(reps: int = 2, ordered: bool = True) -> int:
    """Windows a full search consumes at most (``ordered``: with a root
    order, which adds the heavy/light stage)."""
    return reps * (len(BLOCKS) + (2 if ordered else 0)) + (reps + 1) * len(HANDOVER)


def autotune(forest: Forest, g: Graph, ctx: Context, u, v, window: int, ne: int, reps: int = 2,
             cache: str | None = None, base: BatchConfig | None = None, candidates=None, margin: float = 0.07):
    """Tunes on the leading windows of (u, v), applying them.  Returns
    (TuneResult, number of insertions consumed)."""
    key = cache_key(forest.nv, ne, forest.k, window)
    hit = _load(cache).get(key)
    if hit is not None:
        return TuneResult(hit["block_threads"], hit.get("grid_blocks", 0), True, 0, hit.get("table", []),
                          hit.get("heavy_roots", -1), hit.get("park_min", 0), hit.get("park_scan", 0)), 0
    base = base or BatchConfig(window=window)
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    K = forest.k
    cur = {"block_threads": 1024, "grid_blocks": 0, "heavy_roots": -1, "park_min": 0, "park_scan": 0}
    table = []
    pos = 0
    used = 0

    def run_stage(stage, cands, nrep=reps, stat=np.median, need=margin):
        nonlocal pos, used
        times = [[] for _ in cands]
        for rep in range(nrep):
            for ci, cand in enumerate(cands):
                if pos + window > u.shape[0]:
                    break
                # a candidate several times slower than the best first window
                # is not run again (one-wavefront teams on hub-heavy graphs)
                firsts = [t[0] for t in times if t]
                if rep > 0 and times[ci] and times[ci][0] > 4.0 * min(firsts):
                    continue
                c = dict(cur, **cand)
                cfg = BatchConfig(window=window, block_threads=c["block_threads"], grid_blocks=c["grid_blocks"],
                                  use_graph=base.use_graph, recompute_bc=base.recompute_bc,
                                  heavy_roots=c["heavy_roots"], root_order=base.root_order,
                                  park_min=c["park_min"], park_scan=c["park_scan"])
                ctx.reset_stats(True)
                forest.insert_batched(g, ctx, u[pos:pos + window], v[pos:pos + window], cfg)
                ctx.synchronize()
                _, ms = ctx.kernel_stats(KSTAT_UPDATE)
                _, wms = ctx.kernel_stats(KSTAT_WIDE)
                times[ci].append(ms + wms)
                pos += window
                used += 1
        best, best_ms = None, None
        for ci, (cand, t) in enumerate(zip(cands, times)):
            if not t:
                continue
            med = float(stat(t))
            table.append(dict(cand, stage=stage, ms_per_window=round(med, 4), windows=len(t)))
            if best is None:
                best, best_ms = cand, med          # the current setting
            elif med < best_ms * (1.0 - need):
                best, best_ms = cand, med
        if best is not None:
            cur.update(best)

    if candidates is not None:
        run_stage("given", list(candidates))
    else:
        run_stage("team size", [{"block_threads": b} for b in BLOCKS])
        if cur["block_threads"] > 64:
            if base.root_order is not None:
                run_stage("heavy/light split", [{"heavy_roots": -1}, {"heavy_roots": K - K // 4}])
            # the thresholds matter in the few windows with a giant item, which
            # a median ignores: the mean, and a larger margin, since one such
            # window in a candidate's three moves its mean (on R-MAT scale 20
            # a 7% median rule once picked 8192/1M on tuning windows that then
            # ran the timed ones 8% slower)
            run_stage("hand-over thresholds", [{"park_min": a, "park_scan": b} for a, b in HANDOVER], reps + 1,
                      stat=np.mean, need=max(margin, 0.15))
    res = TuneResult(cur["block_threads"], cur["grid_blocks"], False, used, table, cur["heavy_roots"],
                     cur["park_min"], cur["park_scan"])
    _store(cache, key, {k: getattr(res, k) for k in ("block_threads", "grid_blocks", "heavy_roots", "park_min",
                                                     "park_scan", "table")})
    return res, pos


def result_dict(res: TuneResult) -> dict:
    return asdict(res)
