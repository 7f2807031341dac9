"""What the learned cost model gains over scheduling without it.

R-MAT scale 15 (16 draws/vertex) with four roots per compute unit (K = 4 x
CUs) and windows of 32 insertions: a 1024-thread team takes a whole CU, so a
window runs its roots in four waves of teams and the root order decides which
roots wait.  Three forests start from the same state (the same leading
windows, identity order; the cost model trains on the first forest's item
records) and then run the same timed windows, each under one order:

  static   identity root order;
  balance  roots by descending BALANCE estimate over the window (the
           reference's per-insertion load-balancing formula,
           src/streamingbc.c:98-122, summed per root; computed on the host
           from each tree's levels);
  mlp      roots by descending predicted window cost, on the device
           (streamingBCScheduleWindow, BF16-MFMA inference).

The three end with bit-identical trees and BC; the device time of the timed
windows is reported for each, and the learned order must not lose to the
identity order by more than timing noise.
"""
import numpy as np
import pytest

from oracle_py import balance_work

pytestmark = pytest.mark.gpu

sbc = pytest.importorskip("streamingbc_amd")
from streamingbc_amd.schedule import LearnedSchedule  # noqa: E402


def _csr(hg, u, v):
    """CSR of hg plus the undirected edges (u[i], v[i])."""
    src = np.repeat(np.arange(hg.nv, dtype=np.int64), np.diff(hg.rowptr))
    a = np.concatenate([src, u, v])
    b = np.concatenate([hg.col, v, u])
    order = np.argsort(a, kind="stable")
    a, b = a[order], b[order]
    rowptr = np.zeros(hg.nv + 1, dtype=np.int64)
    np.add.at(rowptr, a + 1, 1)
    return np.cumsum(rowptr), b


def _window_ms(ctx):
    _, ms = ctx.kernel_stats(sbc.KSTAT_UPDATE)
    _, wms = ctx.kernel_stats(sbc.KSTAT_WIDE)
    return ms + wms


def test_learned_order_against_static_and_balance():
    hg = sbc.gen_rmat(15, 16 << 15, seed=4, keep_lcc=False)
    ctx0 = sbc.Context(hg.nv)
    ncu = ctx0.num_cu
    K, W, n_train, n_timed = 4 * ncu, 32, 8, 8
    roots = hg.pick_roots(K, 3)
    u, v = hg.insertions((n_train + n_timed) * W, 19)
    print(f"[gain] rmat15 nv={hg.nv} slots={hg.ne} K={K} W={W} train={n_train} timed={n_timed}", flush=True)
    runs = {}
    for name in ("static", "balance", "mlp"):
        ctx = sbc.Context(hg.nv)
        g = sbc.Graph(hg)
        f = sbc.Forest(hg.nv, roots)
        f.init_static(g, ctx)
        runs[name] = (ctx, g, f)
    sched = LearnedSchedule(train_after=n_train, epochs=12, batch=1024, lr=2e-3, eval_every=1)
    base = sbc.BatchConfig(window=W, block_threads=1024)
    # leading windows: identity order everywhere; the model learns from them
    for w in range(n_train):
        a, b = w * W, (w + 1) * W
        for name, (ctx, g, f) in runs.items():
            if name == "mlp":
                plan, X = sched.plan(f, g, ctx, u[a:b], v[a:b])
                rec = f.insert_batched(g, ctx, u[a:b], v[a:b], base, records=True)
                sched.observe(plan, X, rec, ctx, K)
            else:
                f.insert_batched(g, ctx, u[a:b], v[a:b], base)
    assert sched.trained
    ms = {name: 0.0 for name in runs}
    inserted_u, inserted_v = list(u[:n_train * W]), list(v[:n_train * W])
    for w in range(n_train, n_train + n_timed):
        a, b = w * W, (w + 1) * W
        for name, (ctx, g, f) in runs.items():
            cfg = sbc.BatchConfig(window=W, block_threads=1024)
            plan = None
            if name == "balance":
                rowptr, col = _csr(hg, np.array(inserted_u), np.array(inserted_v))
                work = balance_work(rowptr, col, f.levels(), u[a:b], v[a:b])
                cfg.root_order = np.argsort(-work, kind="stable").astype(np.int32)
            elif name == "mlp":
                plan, X = sched.plan(f, g, ctx, u[a:b], v[a:b])
                cfg.use_context_order = plan.device_order
                cfg.root_order = plan.root_order
            ctx.synchronize()
            ctx.reset_stats(True)
            rec = f.insert_batched(g, ctx, u[a:b], v[a:b], cfg, records=plan is not None and plan.records)
            ctx.synchronize()
            ms[name] += _window_ms(ctx)
            if plan is not None:
                sched.observe(plan, None, rec, ctx, K)
        inserted_u += list(u[a:b])
        inserted_v += list(v[a:b])
    rep = sched.report()
    print(f"[gain] device ms over {n_timed} windows: {ms}; static/mlp {ms['static'] / ms['mlp']:.3f} "
          f"balance/mlp {ms['balance'] / ms['mlp']:.3f}; prediction {rep}", flush=True)
    # the order never changes results
    ref = runs["static"][2]
    runs["static"][2].recompute_bc(runs["static"][0])
    for name in ("balance", "mlp"):
        ctx, g, f = runs[name]
        f.recompute_bc(ctx)
        for x, y in zip(f.trees(), ref.trees()):
            assert np.array_equal(x, y), name
        assert np.array_equal(f.bc(), ref.bc()), name
    assert rep["root_window_spearman"] > 0.5
    assert ms["mlp"] <= 1.05 * ms["static"], "the learned order lost to the identity order"
