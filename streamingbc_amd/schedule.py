"""Scheduling of the windowed update: which roots a window runs first and
which run in the light (one wavefront per root) phase.

The reference orders roots per insertion by its BALANCE estimate
(src/streamingbc.c:98-122: 2*edgesBelow + edgesAbove of the deeper endpoint,
sorted descending, src/streamingbc.c:80-86).  Here the estimate is a learned
one: the cost model (``CostModel``) predicts log1p(affected vertices) of each
(root, insertion) item from features read on the device before the window
runs; a root's window cost is the sum of its items' predicted work.  Roots are
ordered by descending cost (longest first) and those predicted cheap are
binned into the light phase.  The schedule changes timing only: every
schedule gives bit-identical trees and BC.

``StaticSchedule`` (identity order, every root heavy) is the fallback used
until the model is trained and whenever it is unavailable.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import (BatchConfig, CostModel, Context, Forest, Graph, schedule_window, set_root_history, update_root_history,
               window_features)

# Fixed per-item work (in affected-vertex units) added to each prediction:
# even an item that changes nothing reads its endpoints and syncs its team.
ITEM_OVERHEAD = 4.0


@dataclass
class WindowPlan:
    root_order: np.ndarray | None
    heavy_roots: int
    predicted: np.ndarray | None = None  # (n, K) predicted log1p(affected)
    device_order: bool = False           # order left on the device by schedule_window
    records: bool = False                # the schedule reads this window's item records


class StaticSchedule:
    """Identity root order, every root in the heavy phase."""

    name = "static"
    needs_records = False

    def plan(self, forest, g, ctx, u, v):
        return WindowPlan(None, -1), None

    def observe(self, plan, X, records, ctx=None, k=None) -> None:
        pass


@dataclass
class LearnedSchedule:
    """Cost-model schedule.  Collects (features, label) pairs from the
    windows it sees; after ``train_after`` windows it trains the model and
    from then on orders and bins roots by predicted cost.

    light_max: roots whose predicted window work (affected vertices summed
    over the window's items) is below this run in the light phase; 0 keeps
    every root heavy (ordering only)."""

    train_after: int = 8
    epochs: int = 12
    batch: int = 1024
    lr: float = 2e-3
    seed: int = 7
    light_max: float = 0.0
    hist_decay: float = 0.5
    mfma: bool = True
    name: str = "mlp"
    # the host path reads every window's item records (root history feature);
    # the device path keeps that feature on the device once trained and reads
    # records only for its training windows and the report windows
    needs_records: bool = True
    model: CostModel | None = None
    _X: list = field(default_factory=list)
    _y: list = field(default_factory=list)
    _hist: np.ndarray | None = None
    _windows: int = 0
    trained: bool = False
    train_loss: float | None = None
    train_seconds: float = 0.0
    eval_pred: list = field(default_factory=list)
    eval_true: list = field(default_factory=list)
    device: bool = True     # features, predictions and the root sort on the device
    eval_every: int = 4     # trained windows whose predictions are copied back for the report

    def __post_init__(self):
        if self.model is None:
            self.model = CostModel(self.seed)

    def _features(self, forest, g, ctx, u, v):
        if self._hist is None:
            self._hist = np.zeros(forest.k, dtype=np.float32)
        return window_features(forest, g, ctx, u, v, self._hist)

    def _predict_only(self, forest, g, ctx, u, v):
        _, pred = schedule_window(forest, g, ctx, self.model, u, v, None, 0.0, ITEM_OVERHEAD, predictions=True)
        return pred

    def _device(self, forest) -> bool:
        return self.trained and self.device and self.mfma and forest.k <= 4096

    def plan(self, forest: Forest, g: Graph, ctx: Context, u, v):
        """Returns (WindowPlan, features of the window)."""
        if self._hist is None:
            self._hist = np.zeros(forest.k, dtype=np.float32)
        self._ctx = ctx
        if self._device(forest):
            self._sched_windows = getattr(self, "_sched_windows", 0) + 1
            want = (self._sched_windows - 1) % self.eval_every == 0
            if self.light_max <= 0 and forest.k <= ctx.num_cu:
                # every root gets a team of its own at once (a 1024-thread team
                # takes a whole CU's LDS): the order cannot change the window
                # time, so only report windows are predicted
                if not want:
                    return WindowPlan(None, -1, None), None
                return WindowPlan(None, -1, self._predict_only(forest, g, ctx, u, v), records=True), None
            heavy, pred = schedule_window(forest, g, ctx, self.model, u, v, None, self.light_max,
                                          ITEM_OVERHEAD, predictions=want)
            return WindowPlan(None, heavy, pred, device_order=True, records=want), None
        X = self._features(forest, g, ctx, u, v)
        if not self.trained:
            return WindowPlan(None, -1, None, records=True), X
        pred = self.model.predict(X, mfma=self.mfma)  # (n, K)
        work = np.expm1(np.maximum(pred, 0.0)).sum(axis=0) + ITEM_OVERHEAD * pred.shape[0]
        order = np.argsort(-work, kind="stable").astype(np.int32)
        heavy = int(np.count_nonzero(work >= self.light_max)) if self.light_max > 0 else -1
        return WindowPlan(order, heavy, pred, records=True), X

    def observe(self, plan: WindowPlan, X: np.ndarray, records: np.ndarray | None, ctx=None, k=None) -> None:
        """records: (n, K) item records of the window just run (may be None
        on the device path when the plan did not ask for them)."""
        if self.trained and getattr(self, "_hist_on_device", False):
            # the history feature lives on the device: fold the window in there
            update_root_history(ctx if ctx is not None else self._ctx, k if k is not None else self._hist.shape[0],
                                self.hist_decay)
            if records is not None and plan.predicted is not None:
                y = np.log1p(records["affected"].astype(np.float64)).astype(np.float32)
                self.eval_pred.append(plan.predicted.copy())
                self.eval_true.append(y.copy())
            self._windows += 1
            return
        y = np.log1p(records["affected"].astype(np.float64)).astype(np.float32)
        if plan.predicted is not None:
            self.eval_pred.append(plan.predicted.copy())
            self.eval_true.append(y.copy())
        if X is not None and not self.trained:
            self._X.append(X.reshape(-1, X.shape[-1]))
            self._y.append(y.reshape(-1))
        # per-root history feature: decayed mean label of the root's items
        self._hist = (self.hist_decay * self._hist + (1.0 - self.hist_decay) * y.mean(axis=0)).astype(np.float32)
        self._windows += 1
        if not self.trained and self._windows >= self.train_after:
            self.fit()
            if self.device and self.mfma and self._hist.shape[0] <= 4096 and getattr(self, "_ctx", None) is not None:
                set_root_history(self._ctx, self._hist)
                self._hist_on_device = True

    def fit(self) -> float:
        X = np.concatenate(self._X)
        y = np.concatenate(self._y)
        t0 = time.perf_counter()
        self.train_loss = self.model.train(X, y, epochs=self.epochs, batch=self.batch, lr=self.lr, seed=self.seed)
        self.train_seconds = time.perf_counter() - t0
        self.trained = True
        self.train_samples = int(X.shape[0])
        return self.train_loss

    def report(self) -> dict:
        """Prediction error on windows scheduled after training (held out:
        never used for training)."""
        if not self.eval_pred:
            return {"trained": self.trained, "eval_items": 0}
        p = np.concatenate([a.reshape(-1) for a in self.eval_pred]).astype(np.float64)
        t = np.concatenate([a.reshape(-1) for a in self.eval_true]).astype(np.float64)
        rw_p = np.concatenate([np.expm1(np.maximum(a, 0)).sum(axis=0) for a in self.eval_pred])
        rw_t = np.concatenate([np.expm1(a).sum(axis=0) for a in self.eval_true])
        return {
            "trained": True,
            "train_samples": getattr(self, "train_samples", 0),
            "train_loss": self.train_loss,
            "train_s": round(self.train_seconds, 3),
            "eval_items": int(p.size),
            "rmse_log1p": float(np.sqrt(np.mean((p - t) ** 2))),
            "mae_log1p": float(np.mean(np.abs(p - t))),
            "root_window_spearman": spearman(rw_p, rw_t),
        }


def spearman(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size < 2:
        return float("nan")
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    den = np.sqrt((ra * ra).sum() * (rb * rb).sum())
    return float((ra * rb).sum() / den) if den > 0 else float("nan")


def run_stream(forest: Forest, g: Graph, ctx: Context, u, v, window: int, schedule, base: BatchConfig | None = None):
    """Applies insertions (u[i], v[i]) window by window under ``schedule``.
    Returns per-window wall seconds (host call, including the plan) and the
    item records (n, K)."""
    base = base or BatchConfig(window=window, block_threads=1024)
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    n = u.shape[0]
    times = []
    recs = []
    for w0 in range(0, n, window):
        uu, vv = u[w0:w0 + window], v[w0:w0 + window]
        t0 = time.perf_counter()
        plan, X = schedule.plan(forest, g, ctx, uu, vv)
        cfg = BatchConfig(window=window, block_threads=base.block_threads, grid_blocks=base.grid_blocks,
                          use_graph=base.use_graph, recompute_bc=base.recompute_bc, heavy_roots=plan.heavy_roots,
                          root_order=plan.root_order, park_min=base.park_min, use_context_order=plan.device_order)
        rec = forest.insert_batched(g, ctx, uu, vv, cfg, records=True)
        times.append(time.perf_counter() - t0)
        recs.append(rec)
        schedule.observe(plan, X, rec, ctx, forest.k)
    return np.array(times), np.concatenate(recs) if recs else None
