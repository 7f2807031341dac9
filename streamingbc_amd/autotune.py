"""Online autotuner for the windowed update's launch parameters.

The first windows of a stream are run under each candidate launch shape in
turn (round-robin, ``reps`` windows each), the device time of each window's
update kernel is recorded, and the candidate with the lowest median time per
window is kept for the rest of the stream.  Tuning windows are real work: the
launch shape never changes results (trees and BC are bit-identical under every
shape), so nothing is thrown away or replayed.

Results are cached in a JSON file keyed by (vertices, edges, roots, window)
so a later stream on the same graph shape skips the sweep.
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass, field

import numpy as np

from . import KSTAT_UPDATE, KSTAT_WIDE, BatchConfig, Context, Forest, Graph

DEFAULT_CANDIDATES = (
    {"block_threads": 1024, "grid_blocks": 0},
    {"block_threads": 512, "grid_blocks": 0},
    {"block_threads": 256, "grid_blocks": 0},
)


@dataclass
class TuneResult:
    block_threads: int
    grid_blocks: int
    from_cache: bool
    windows_used: int
    table: list = field(default_factory=list)

    def config(self, window: int, **kw) -> BatchConfig:
        return BatchConfig(window=window, block_threads=self.block_threads, grid_blocks=self.grid_blocks, **kw)


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


def autotune(forest: Forest, g: Graph, ctx: Context, u, v, window: int, ne: int, candidates=DEFAULT_CANDIDATES,
             reps: int = 2, cache: str | None = None, base: BatchConfig | None = None):
    """Tunes on the leading windows of (u, v), applying them.  Returns
    (TuneResult, number of insertions consumed)."""
    key = cache_key(forest.nv, ne, forest.k, window)
    hit = _load(cache).get(key)
    if hit is not None:
        return TuneResult(hit["block_threads"], hit["grid_blocks"], True, 0, hit.get("table", [])), 0
    base = base or BatchConfig(window=window)
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    times = {i: [] for i in range(len(candidates))}
    pos = 0
    used = 0
    for rep in range(reps):
        for ci, cand in enumerate(candidates):
            if pos + window > u.shape[0]:
                break
            cfg = BatchConfig(window=window, block_threads=cand["block_threads"], grid_blocks=cand["grid_blocks"],
                              use_graph=base.use_graph, recompute_bc=base.recompute_bc,
                              heavy_roots=base.heavy_roots, root_order=base.root_order)
            ctx.reset_stats(True)
            forest.insert_batched(g, ctx, u[pos:pos + window], v[pos:pos + window], cfg)
            ctx.synchronize()
            _, ms = ctx.kernel_stats(KSTAT_UPDATE)
            _, wms = ctx.kernel_stats(KSTAT_WIDE)
            times[ci].append(ms + wms)
            pos += window
            used += 1
    table = []
    for ci, cand in enumerate(candidates):
        if times[ci]:
            table.append(dict(cand, ms_per_window=float(np.median(times[ci])), windows=len(times[ci])))
    if not table:
        c = candidates[0]
        return TuneResult(c["block_threads"], c["grid_blocks"], False, 0, []), 0
    best = min(table, key=lambda r: r["ms_per_window"])
    res = TuneResult(best["block_threads"], best["grid_blocks"], False, used, table)
    _store(cache, key, {"block_threads": res.block_threads, "grid_blocks": res.grid_blocks, "table": table})
    return res, pos


def result_dict(res: TuneResult) -> dict:
    return asdict(res)
