"""Serving loop: several graphs resident in HBM at once, each with its own
forest, context and HIP stream, fed by a stream of edge insertions.

* Admission: a graph is admitted only while the sum of resident footprints
  (``streamingBCResidentFootprint``: forest + device CSR + update scratch)
  stays within the HBM budget; otherwise ``admit`` raises ``SbcError`` with
  status ``SBC_ERR_BUDGET`` before touching the device, and ``submit`` queues
  the graph instead: queued graphs are admitted, first come first served,
  while ``serve`` runs, as soon as graphs whose streams have ended are evicted
  (``evict_when_done``) and their memory is free; their insertions wait
  meanwhile and the wait counts in their latency.
* Continuous batching: one host worker per resident graph takes every
  insertion that has arrived and not yet been applied (up to ``max_window``)
  and applies them as one window; work for different graphs runs
  concurrently on their own streams (the library's calls release the GIL).
* Latency of an insertion = time its window finished on the device (BC of
  the graph includes it from then on) minus its arrival time.

Results do not depend on how insertions were batched or what else was
resident: per root the update applies insertions in stream order with the
same arithmetic, so a graph served in a mix ends bit-identical to the same
graph served alone (BC read with the deterministic recompute).

Run ``python -m streamingbc_amd.serve`` for the Chung-Lu serving workload.
"""
from __future__ import annotations

import argparse
import collections
import json
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import (SBC_ERR_ARG, SBC_ERR_BUDGET, SBC_ERR_EXISTS, BatchConfig, Context, Forest, Graph, HostGraph, SbcError,
               gen_chung_lu, resident_footprint)


# An item goes to the whole-GPU kernels once its ascent level has this many
# adjacency entries left to scan (the library's default is 512K).  Measured on
# the Chung-Lu serving set (2 s per rate): 512K gives 17.2K/s sustained at an
# offered 50K/s and p99 7.0 ms at 10K/s; 2M gives 25.9K/s and p99 5.3 ms
# (1M with park_min 8192: 25.9K/s, 7.3 ms; no hand-overs: 19.1K/s, 97 ms).  Every hand-over runs
# one launch per level from the graph's worker while the other graphs'
# windows wait for the GPU, so fewer, larger ones pay off here.
SERVE_PARK_SCAN = 2048 * 1024


@dataclass
class Pending:
    name: str
    hg: HostGraph
    roots: np.ndarray
    footprint: int


@dataclass
class Resident:
    name: str
    hg: HostGraph
    roots: np.ndarray
    footprint: int
    ctx: Context
    g: Graph
    f: Forest
    applied: int = 0
    windows: list = field(default_factory=list)


class Server:
    def __init__(self, budget_bytes: int, max_window: int = 128, block_threads: int = 1024, use_graph: bool = True,
                 num_cu: int = 0, park_min: int = 0, park_scan: int = SERVE_PARK_SCAN):
        self.budget = int(budget_bytes)
        self.max_window = int(max_window)
        self.park_min = int(park_min)
        self.park_scan = int(park_scan)
        self.block_threads = block_threads
        self.use_graph = use_graph
        self.num_cu = num_cu
        self.resident: dict[str, Resident] = {}
        self.pending: collections.deque = collections.deque()
        self.admitted_at: dict[str, float] = {}   # seconds into serve() when a queued graph was admitted

    @property
    def used(self) -> int:
        return sum(r.footprint for r in self.resident.values())

    def footprint(self, hg: HostGraph, k: int) -> int:
        return resident_footprint(hg, k, self.num_cu)

    def admit(self, name: str, hg: HostGraph, roots) -> Resident:
        if name in self.resident:
            raise SbcError(SBC_ERR_EXISTS, f"graph {name!r} is already resident")
        roots = np.ascontiguousarray(roots, dtype=np.int64)
        if roots.ndim != 1 or roots.size == 0:
            raise SbcError(SBC_ERR_ARG, "roots must be a non-empty 1-d array")
        fp = self.footprint(hg, roots.size)
        if self.used + fp > self.budget:
            raise SbcError(SBC_ERR_BUDGET, f"graph {name!r} needs {fp} B; {self.budget - self.used} B of "
                                           f"{self.budget} B free")
        ctx = Context(hg.nv)
        g = Graph(hg)
        f = Forest(hg.nv, roots)
        f.init_static(g, ctx)
        r = Resident(name, hg, roots, fp, ctx, g, f)
        self.resident[name] = r
        return r

    def submit(self, name: str, hg: HostGraph, roots) -> bool:
        """Admits the graph now if it fits, else queues it for admission
        during ``serve``.  Returns True if admitted now.  A graph larger than
        the whole budget is refused (SBC_ERR_BUDGET)."""
        roots = np.ascontiguousarray(roots, dtype=np.int64)
        if roots.ndim != 1 or roots.size == 0:
            raise SbcError(SBC_ERR_ARG, "roots must be a non-empty 1-d array")
        if name in self.resident or any(p.name == name for p in self.pending):
            raise SbcError(SBC_ERR_EXISTS, f"graph {name!r} is already resident or queued")
        fp = self.footprint(hg, roots.size)
        if fp > self.budget:
            raise SbcError(SBC_ERR_BUDGET, f"graph {name!r} needs {fp} B; the budget is {self.budget} B")
        if not self.pending and self.used + fp <= self.budget:
            self.admit(name, hg, roots)
            return True
        self.pending.append(Pending(name, hg, roots, fp))
        return False

    def evict(self, name: str) -> None:
        r = self.resident.pop(name)
        r.f.close()
        r.g.close()
        r.ctx.close()

    def _worker(self, r: Resident, t, u, v, t0, lat, err):
        try:
            n = t.shape[0]
            i = 0
            cfg = BatchConfig(window=self.max_window, block_threads=self.block_threads, use_graph=self.use_graph,
                              park_min=self.park_min, park_scan=self.park_scan)
            while i < n:
                now = time.perf_counter() - t0
                if t[i] > now:
                    time.sleep(min(t[i] - now, 0.01))
                    continue
                j = int(np.searchsorted(t, now, side="right"))
                j = min(j, i + self.max_window, n)
                r.f.insert_batched(r.g, r.ctx, u[i:j], v[i:j], cfg)
                r.ctx.synchronize()
                done = time.perf_counter() - t0
                lat[i:j] = done - t[i:j]
                r.windows.append(j - i)
                r.applied += j - i
                i = j
        except BaseException as e:  # reported by serve()
            err.append((r.name, e))

    def serve(self, arrivals: dict, evict_when_done: bool = False, results: dict | None = None) -> dict:
        """arrivals: name -> (t_seconds sorted ascending, u, v) for resident
        or queued graphs.  Blocks until every insertion is applied.  With
        evict_when_done a graph is evicted as soon as its stream is applied
        (its deterministic BC kept in ``results`` when given), which admits
        queued graphs that then fit."""
        queued = {p.name for p in self.pending}
        for name in arrivals:
            if name not in self.resident and name not in queued:
                raise SbcError(SBC_ERR_ARG, f"graph {name!r} is neither resident nor queued")
        lats = {}
        err = []
        threads = []
        windows = {}
        lock = threading.Lock()
        streams = {}
        for name, (t, u, v) in arrivals.items():
            t = np.asarray(t, dtype=np.float64)
            lats[name] = np.zeros(t.shape[0])
            streams[name] = (t, np.asarray(u, np.int64), np.asarray(v, np.int64))
        t0 = time.perf_counter()

        def run(name):
            r = self.resident[name]
            t, u, v = streams[name]
            self._worker(r, t, u, v, t0, lats[name], err)
            with lock:
                windows[name] = list(r.windows)
                if results is not None:
                    results[name] = self.bc(name)
                if evict_when_done:
                    self.evict(name)
                    admit_ready()

        def start(name):
            th = threading.Thread(target=run, args=(name,))
            th.start()
            threads.append(th)

        def admit_ready():
            # first come, first served: the head of the queue waits for room
            while self.pending and self.used + self.pending[0].footprint <= self.budget:
                p = self.pending.popleft()
                try:
                    self.admit(p.name, p.hg, p.roots)
                except BaseException as e:  # reported after the join
                    err.append((p.name, e))
                    continue
                self.admitted_at[p.name] = time.perf_counter() - t0
                if p.name in streams:
                    start(p.name)

        with lock:
            for name in arrivals:
                if name in self.resident:
                    start(name)
        while True:
            with lock:
                live = [th for th in threads if th.is_alive()]
                if not live:
                    break
            live[0].join()
        wall = time.perf_counter() - t0
        if err:
            raise err[0][1]
        left = [n for n in arrivals if n not in windows]
        if left:
            raise SbcError(SBC_ERR_BUDGET, f"graphs {left} were never admitted (evict_when_done frees memory)")
        return summarize(lats, wall, windows)

    def bc(self, name: str) -> np.ndarray:
        r = self.resident[name]
        r.f.recompute_bc(r.ctx)
        return r.f.bc()


def summarize(lats: dict, wall: float, windows: dict) -> dict:
    allv = np.concatenate([a for a in lats.values()]) if lats else np.zeros(0)
    out = {
        "insertions": int(allv.size),
        "wall_s": round(wall, 4),
        "sustained_ins_per_s": round(allv.size / wall, 2) if wall > 0 else None,
        "p50_ms": round(float(np.percentile(allv, 50)) * 1e3, 3) if allv.size else None,
        "p99_ms": round(float(np.percentile(allv, 99)) * 1e3, 3) if allv.size else None,
        "graphs": {},
    }
    for name, a in lats.items():
        w = windows[name]
        out["graphs"][name] = {
            "insertions": int(a.size),
            "p50_ms": round(float(np.percentile(a, 50)) * 1e3, 3) if a.size else None,
            "p99_ms": round(float(np.percentile(a, 99)) * 1e3, 3) if a.size else None,
            "windows": len(w),
            "mean_window": round(float(np.mean(w)), 2) if w else 0.0,
        }
    return out


def poisson_arrivals(rate: float, count: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


# Chung-Lu serving set: (log2 vertices, roots).  gamma 2.3, mean degree 8.
SERVE_GRAPHS = ((14, 128), (15, 128), (16, 256), (17, 256), (18, 512), (20, 512))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Serving loop over resident Chung-Lu graphs")
    ap.add_argument("--rates", default="1000,10000,50000", help="aggregate arrival rates (insertions/s)")
    ap.add_argument("--seconds", type=float, default=4.0, help="arrival span per rate")
    ap.add_argument("--max-window", type=int, default=128)
    ap.add_argument("--park-min", type=int, default=0,
                    help="sbcBatchConfig.park_min (0: the library default, < 0: never hand items to the wide kernels)")
    ap.add_argument("--park-scan", type=int, default=SERVE_PARK_SCAN,
                    help="sbcBatchConfig.park_scan (0: the library default of 512K adjacency entries)")
    ap.add_argument("--budget-gib", type=float, default=200.0)
    ap.add_argument("--graphs", default=",".join(f"{s}:{k}" for s, k in SERVE_GRAPHS))
    ap.add_argument("--seed", type=int, default=11)
    a = ap.parse_args(argv)
    srv = Server(int(a.budget_gib * (1 << 30)), max_window=a.max_window, park_min=a.park_min, park_scan=a.park_scan)
    specs = [tuple(int(x) for x in s.split(":")) for s in a.graphs.split(",")]
    t0 = time.perf_counter()
    hgs = {}
    for sc, k in specs:
        name = f"chunglu{sc}_k{k}"
        hg = gen_chung_lu(1 << sc, 2.3, 8.0, seed=sc, keep_lcc=True)
        srv.admit(name, hg, hg.pick_roots(min(k, hg.nv), sc))
        hgs[name] = hg
        print(f"[serve] admitted {name} nv={hg.nv} ne={hg.ne} footprint={srv.resident[name].footprint / 2**30:.2f} GiB",
              flush=True)
    setup_s = time.perf_counter() - t0
    results = []
    offset = {n: 0 for n in hgs}
    rates = [float(x) for x in a.rates.split(",")]
    need = int(sum(r * a.seconds for r in rates) / len(hgs) * 1.5) + 1000
    pools = {n: hgs[n].insertions(need, a.seed) for n in hgs}
    for rate in rates:
        count = int(rate * a.seconds)
        t = poisson_arrivals(rate, count, a.seed)
        rng = np.random.default_rng(a.seed)
        owner = rng.integers(0, len(hgs), size=count)
        arr = {}
        for gi, name in enumerate(hgs):
            sel = owner == gi
            m = int(sel.sum())
            pu, pv = pools[name]
            o = offset[name]
            if o + m > pu.shape[0]:
                raise SystemExit(f"insertion pool of {name} exhausted")
            arr[name] = (t[sel], pu[o:o + m], pv[o:o + m])
            offset[name] = o + m
        res = srv.serve(arr)
        res["rate"] = rate
        print(f"[serve] rate {rate:.0f}/s: p50 {res['p50_ms']} ms p99 {res['p99_ms']} ms "
              f"sustained {res['sustained_ins_per_s']}/s", flush=True)
        results.append(res)
    print(json.dumps({"workload": "serve_chunglu", "setup_s": round(setup_s, 2), "max_window": a.max_window,
                      "resident_gib": round(srv.used / 2**30, 2), "results": results}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
