#!/usr/bin/env python3
"""Benchmark: edge insertions per second with BC updated for all K roots.

Default workload (the headline configuration): undirected R-MAT scale 20
(2^20 vertices, 16 edge draws per vertex, a/b/c/d = .57/.19/.19/.05,
deduplicated, seed 1), K = 256 roots drawn from the largest component (seed
7), random non-adjacent insertions inside it (seed 11), applied in windows of
128 by the batched persistent kernel.  One step = one window.  The stream is
16,384 insertions: 30 untimed leading windows (the cost model's training
data under --schedule mlp), 2 warmup windows, 96 timed windows (12,288
insertions), then 64 insertions in the plain form.

Printed JSON line (rank 0):
  value         insertions/s of the batched form over the timed steps (inputs
                resident in HBM; the window's edge list is copied in per step)
  plain_gpu     insertions/s of the plain form (one launch sequence per insertion)
  roofline      algorithmic bytes of the update kernel (per-item formula below)
                / its average launch time from HIP events on the context stream
  cpu_baseline  the reference's own OpenMP COARSE update (compiled from its
                sources into oracle/_ref) timed on a bounded prefix of the
                same stream on this host's cores

Algorithmic bytes per (root, insertion) item:
  8 (classify: dist[u], dist[v])
  + 8 * processed (row begin/end per vertex visit, descent and ascent)
  + 8 * scanned   (4 B neighbour id + 4 B dist per adjacency entry)
  + 48 * affected (read sigma,delta 16 + write sigma,delta 16 + BC RMW 16)
  + 4 * moved     (dist write)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "edge insertions/sec with BC updated (K roots) vs plain GPU; p99 update latency"

WORKLOADS = {
    # name: (generator, args, K, window, description)
    "rmat20_k256_w128": ("rmat", dict(scale=20, draws=16 << 20, seed=1, keep_lcc=False), 256, 128,
                         "undirected R-MAT scale 20, 16 draws/vertex, deduplicated, K=256, windows of 128"),
    "rmat18_k512_w64": ("rmat", dict(scale=18, draws=16 << 18, seed=1, keep_lcc=False), 512, 64,
                        "undirected R-MAT scale 18, 16 draws/vertex, deduplicated, K=512, windows of 64"),
    "chunglu15_k256": ("chunglu", dict(n=1 << 15, gamma=2.3, mean_deg=8.0, seed=1, keep_lcc=True), 256, 64,
                       "Chung-Lu 32768 vertices, exponent 2.3, mean degree 8, LCC, K=256"),
}


def item_bytes(rec) -> float:
    r = rec
    return float(8 * r.size + 8 * r["processed"].sum() + 8 * r["scanned"].sum() + 48 * r["affected"].sum()
                 + 4 * r["moved"].sum())


def make_workload(name, n_ins):
    import streamingbc_amd as sbc
    kind, kw, K, W, desc = WORKLOADS[name]
    t0 = time.time()
    if kind == "rmat":
        hg = sbc.gen_rmat(kw["scale"], kw["draws"], seed=kw["seed"], keep_lcc=kw["keep_lcc"])
    else:
        hg = sbc.gen_chung_lu(kw["n"], kw["gamma"], kw["mean_deg"], seed=kw["seed"], keep_lcc=kw["keep_lcc"])
    roots = hg.pick_roots(K, 7)
    u, v = hg.insertions(n_ins, 11)
    return hg, roots, u, v, K, W, desc, time.time() - t0


def cpu_baseline(hg, roots, u, v, budget_s, threads, gran, dels=()):
    """The reference's update timed on the host cores; then, if dels is given,
    its deleteEdgeStreamingBC on those edges (at most budget_s / 2 of it)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_py import Reference, ref_available
    if not ref_available():
        return None, None
    ref = Reference(hg.rowptr, hg.col, roots, nt=threads)
    t_static = ref.static()
    spent, n = 0.0, 0
    while n < len(u) and (spent < budget_s or n == 0):
        spent += ref.insert(int(u[n]), int(v[n]), 1, gran)  # BALANCE, as every reference run script
        n += 1
    ins = {"value": n / spent, "unit": "insertions/s", "cores": threads, "kind": "reference",
           "granularity": "COARSE" if gran == 0 else "FINE",
           "sample": f"first {n} insertions of the same stream, {spent:.1f} s inside insertEdgeStreamingBC "
                     f"(static pass {t_static:.1f} s untimed)"}
    dl = None
    if len(dels):
        from oracle_py import ref_lib
        L = ref_lib()
        cnt = np.zeros(3, dtype=np.int64)
        dspent, nd = 0.0, 0
        for a, b in dels:
            if nd and dspent >= budget_s / 2:
                break
            t = L.ref_delete(ref.h, int(a), int(b), 1, gran, cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
            if t < 0:
                continue
            dspent += t
            nd += 1
        if nd:
            dl = {"value": nd / dspent, "unit": "deletions/s", "cores": threads, "kind": "reference",
                  "sample": f"first {nd} of the same deletions after the {n} insertions above, {dspent:.1f} s "
                            f"inside deleteEdgeStreamingBC"}
    ref.close()
    return ins, dl


def pick_deletions(hg, n, seed=13):
    """n distinct edges of the initial graph, as (u, v) with u < v."""
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(hg.nv, dtype=np.int64), np.diff(hg.rowptr))
    out, seen = [], set()
    while len(out) < n:
        e = int(rng.integers(hg.ne))
        a, b = int(src[e]), int(hg.col[e])
        key = (min(a, b), max(a, b))
        if a != b and key not in seen:
            seen.add(key)
            out.append(key)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=96)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat20_k256_w128", choices=sorted(WORKLOADS))
    ap.add_argument("--window", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--heavy", type=int, default=-2, help="-2: autotuned/default, -1: all heavy")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--plain", type=int, default=64, help="insertions timed in the plain form")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of reference update time")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--schedule", default="static", choices=["static", "mlp"],
                    help="root order per window: identity, or by cost-model prediction (trained on --train-windows)")
    ap.add_argument("--train-windows", type=int, default=30,
                    help="untimed leading windows (the cost model's training data under --schedule mlp)")
    ap.add_argument("--light-max", type=float, default=0.0,
                    help="--schedule mlp: roots predicted below this many affected vertices per window run light")
    ap.add_argument("--autotune", action="store_true", help="tune block size on extra warmup windows first")
    ap.add_argument("--park-min", type=int, default=0,
                    help="descent frontier size at which an item is finished by whole-GPU kernels "
                         "(0: library default, <0: never)")
    ap.add_argument("--deletions", type=int, default=16,
                    help="edges of the initial graph deleted (and timed) after the insertion legs")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (PMC traffic calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 4, "calibrate_copy_bytes_written": n * 4,
                          "copies": 3}))
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(roots would be sharded across GPUs with one all-reduce of BC per readout)", file=sys.stderr)
        sys.exit(2)

    import streamingbc_amd as sbc

    kind, kw, K, W, desc = WORKLOADS[args.workload]
    if args.window:
        W = args.window
    from streamingbc_amd.autotune import DEFAULT_CANDIDATES
    n_tune = 2 * len(DEFAULT_CANDIDATES) if args.autotune else 0
    n_train = args.train_windows
    n_ins = (n_tune + n_train + args.warmup + args.steps) * W + args.plain
    hg, roots, u, v, K, _, desc, t_gen = make_workload(args.workload, n_ins)
    print(f"[bench] {args.workload}: nv={hg.nv} directed slots={hg.ne} K={K} W={W} "
          f"insertions={len(u)} gen={t_gen:.1f}s", file=sys.stderr, flush=True)

    ctx = sbc.Context(hg.nv)
    g = sbc.Graph(hg)
    f = sbc.Forest(hg.nv, roots)
    t0 = time.time()
    f.init_static(g, ctx)
    t_static = time.time() - t0
    print(f"[bench] static pass {t_static:.2f}s", file=sys.stderr, flush=True)

    cfg = sbc.BatchConfig(window=W, block_threads=args.block or 1024, use_graph=not args.no_graph,
                          park_min=args.park_min)
    if args.heavy >= -1:
        cfg.heavy_roots = args.heavy
    pos = 0
    tune = None
    if args.autotune:
        from streamingbc_amd.autotune import autotune
        res, used = autotune(f, g, ctx, u, v, W, hg.ne, base=cfg)
        pos += used
        cfg.block_threads, cfg.grid_blocks = res.block_threads, res.grid_blocks
        tune = {"block_threads": res.block_threads, "grid_blocks": res.grid_blocks, "table": res.table}
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)

    from streamingbc_amd.schedule import LearnedSchedule, StaticSchedule
    if args.schedule == "mlp":
        sched = LearnedSchedule(train_after=max(1, n_train), light_max=args.light_max)
    else:
        sched = StaticSchedule()

    def window(a, b, records=False):
        plan, X = sched.plan(f, g, ctx, u[a:b], v[a:b])
        c = sbc.BatchConfig(window=W, block_threads=cfg.block_threads, grid_blocks=cfg.grid_blocks,
                            use_graph=cfg.use_graph, heavy_roots=cfg.heavy_roots if plan.heavy_roots < 0
                            else plan.heavy_roots, root_order=plan.root_order, park_min=cfg.park_min,
                            use_context_order=plan.device_order)
        rec = f.insert_batched(g, ctx, u[a:b], v[a:b], c, records=records or sched.needs_records)
        parked[0] += ctx.last_parked()
        sched.observe(plan, X, rec)
        return rec

    parked = [0]
    # training windows (cost model) and warmup windows, untimed
    for _ in range(n_train + args.warmup):
        window(pos, pos + W)
        pos += W
    ctx.synchronize()

    # timed windows
    parked[0] = 0
    ctx.reset_stats(timing=True)
    lat = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        window(pos, pos + W)
        lat.append(time.perf_counter() - ts)
        pos += W
    ctx.synchronize()
    t_batched = time.perf_counter() - t0
    n_launch, k_ms = ctx.kernel_stats(sbc.KSTAT_UPDATE)
    # work of the timed windows, summed per root on the device (the item
    # records' fields without copying the records in the timed region)
    tot = ctx.work_totals(K)
    n_items = args.steps * W * K
    alg_bytes = float(8 * n_items + 8 * tot["processed"].sum() + 8 * tot["scanned"].sum()
                      + 48 * tot["affected"].sum() + 4 * tot["moved"].sum())
    per_root = tot["scanned"].astype(np.float64)
    value = args.steps * W / t_batched
    print(f"[bench] batched: {value:.1f} ins/s, window {1e3 * t_batched / args.steps:.2f} ms, "
          f"kernel avg {k_ms / max(n_launch, 1):.2f} ms", file=sys.stderr, flush=True)

    # plain form on the continuation of the stream
    plain = None
    if args.plain > 0:
        ctx.reset_stats(timing=True)
        pr = []
        ctx.synchronize()
        tp = time.perf_counter()
        for i in range(args.plain):
            f.insert_edge(g, ctx, int(u[pos + i]), int(v[pos + i]), sbc.FINE)
            pr.append(ctx.last_records(K))
        ctx.synchronize()
        t_plain = time.perf_counter() - tp
        pn, pms = ctx.kernel_stats(sbc.KSTAT_UPDATE)
        prec = np.concatenate(pr)
        plain = {"value": args.plain / t_plain, "unit": "insertions/s", "insertions": args.plain,
                 "kernel_ms_avg": pms / max(pn, 1),
                 "achieved_GBs": item_bytes(prec) / (pms * 1e-3) / 1e9 if pms > 0 else None}
        print(f"[bench] plain: {plain['value']:.1f} ins/s", file=sys.stderr, flush=True)

    # deletions of edges of the initial graph, after every insertion leg
    deletion = None
    dels = pick_deletions(hg, args.deletions) if args.deletions > 0 else []
    if dels:
        ctx.reset_stats(timing=True)
        cases = np.zeros(3, dtype=np.int64)
        dlat = []
        ctx.synchronize()
        td = time.perf_counter()
        for a, b in dels:
            ts = time.perf_counter()
            c = f.delete_edge(g, ctx, a, b)
            dlat.append(time.perf_counter() - ts)
            cases += [c["sameLevel"], c["adjacent"], c["movement"]]
        ctx.synchronize()
        t_del = time.perf_counter() - td
        sn, sms = ctx.kernel_stats(sbc.KSTAT_STATIC)
        un, ums = ctx.kernel_stats(sbc.KSTAT_UPDATE)
        rebuilt = int(cases[2])
        # the static pass of a rebuilt (movement-case) root reads every
        # adjacency entry with the level of its far end (8 B) and writes
        # level, state and record of every vertex (4 + 1 + 16 B)
        sbytes = rebuilt * (8.0 * hg.ne + 21.0 * hg.nv)
        deletion = {"value": round(len(dels) / t_del, 3), "unit": "deletions/s", "deletions": len(dels),
                    "ms_p50": round(1e3 * float(np.percentile(dlat, 50)), 3),
                    "cases": {"sameLevel": int(cases[0]), "adjacent": int(cases[1]), "movement": int(cases[2])},
                    "update_kernel_ms_avg": round(ums / max(un, 1), 3),
                    "static_kernel_ms_avg": round(sms / max(sn, 1), 3),
                    "static_GBs": round(sbytes / (sms * 1e-3) / 1e9, 1) if sms > 0 else None,
                    "method": "adjacent-case roots: update kernel with negated added paths; movement-case "
                              "roots: rebuilt by the static pass; BC updated by the same changes"}
        print(f"[bench] deletions: {deletion['value']:.2f} del/s, cases {deletion['cases']}",
              file=sys.stderr, flush=True)

    launch_ms = k_ms / max(n_launch, 1)
    achieved = alg_bytes / max(n_launch, 1) / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "sbc_window_kernel", "launch_ms": round(launch_ms, 4),
                "alg_bytes_per_launch": alg_bytes / max(n_launch, 1)}
    # measured HBM bytes per launch of this kernel on this workload: the
    # calibrated FETCH_SIZE + WRITE_SIZE of the rocprofv3 PMC passes kept
    # beside the kernel-trace summary
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath) and cfg.block_threads == 1024 and W == WORKLOADS[args.workload][3]:
        with open(tpath) as fh:
            t = json.load(fh)
        roofline["traffic"] = t["hbm_bytes_per_launch"]
        roofline["traffic_source"] = f"profiles/traffic_{args.workload}.json ({t['windows']})"
        # the bytes the kernel actually moves, per second of launch: far below
        # `achieved`, whose formula counts every adjacency entry of every
        # processed vertex (most of which lookups and source sets never read)
        if launch_ms > 0:
            roofline["measured_GBs"] = round(t["hbm_bytes_per_launch"] / (launch_ms * 1e-3) / 1e9, 2)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "insertions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_batched / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "nv": hg.nv, "directed_slots": hg.ne,
                   "roots": K, "window": W, "block_threads": cfg.block_threads, "hip_graph": cfg.use_graph,
                   "parallelism": "single GPU, roots across workgroups"},
        "latency_ms": {"window_p50": round(1e3 * float(np.percentile(lat, 50)), 4),
                       "window_p99": round(1e3 * float(np.percentile(lat, 99)), 4)},
        "plain_gpu": plain,
        "speedup_vs_plain": round(value / plain["value"], 3) if plain else None,
        "roofline": roofline,
        "items": {"count": n_items, "nonzero_gap": int(tot["nonzero"].sum()),
                  "root_scan_max_over_mean": float(per_root.max() / max(per_root.mean(), 1)),
                  "mean_affected": float(tot["affected"].sum() / n_items), "max_affected": int(tot["max_affected"].max()),
                  "mean_scanned": float(tot["scanned"].sum() / n_items),
                  "parked": int(parked[0])},
        "static_pass_s": round(t_static, 3),
        "deletion": deletion,
        "schedule": {"name": sched.name, **(sched.report() if args.schedule == "mlp" else {})},
        "autotune": tune,
    }

    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        f.close()
        g.close()
        cb, cdl = cpu_baseline(hg, roots, u, v, args.cpu_budget, threads, 0, dels)
        if cb:
            out["cpu_baseline"] = cb
            out["speedup_vs_cpu"] = round(value / cb["value"], 2)
        if cdl and deletion:
            deletion["cpu_baseline"] = cdl
            deletion["speedup_vs_cpu"] = round(deletion["value"] / cdl["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
