#!/usr/bin/env python3
"""Benchmark: edge insertions per second with BC updated for all K roots.

Default workload (the headline configuration): undirected R-MAT scale 20
(2^20 vertices, 16 edge draws per vertex, a/b/c/d = .57/.19/.19/.05,
deduplicated, seed 1), K = 256 roots drawn from the largest component (seed
7), random non-adjacent insertions inside it (seed 11), applied in windows of
128 by the batched persistent kernel.  One step = one window.  The stream is
cut into 30 untimed leading windows (on at most the first 18 the autotuner
picks team size and hand-over thresholds; the rest are the cost
model's training data under --schedule mlp), the warmup windows and the
timed windows; the plain form is then timed on exactly the insertions of the
timed windows, on a second forest brought to the same state.

Printed JSON line (rank 0):
  value         insertions/s of the batched form over the timed steps (inputs
                resident in HBM; the window's edge list is copied in per step)
  plain_gpu     insertions/s of the plain form (one launch sequence per
                insertion) on the same insertions, from the same state
  roofline      HBM bytes of the window-kernel launch (rocprofv3 FETCH_SIZE /
                WRITE_SIZE, calibrated by a device copy of known size, on the
                same window indices of the same stream; profiles/) divided by
                the launch's duration from HIP events on the context stream in
                this run; beside it the algorithmic rate (formula below)
  cpu_baseline  the reference's own OpenMP update (compiled from its sources
                into oracle/_ref), COARSE and FINE, timed on a bounded prefix
                of the same stream on this host's CPU share

Algorithmic bytes per (root, insertion) item, from the counts the update
records (the reference's traversal counters):
  8 (classify: dist[u], dist[v])
  + 8 * processed (row begin/end per vertex visit, descent and ascent)
  + 8 * scanned   (4 B neighbour id + 4 B dist per adjacency entry)
  + 48 * affected (read sigma,delta 16 + write sigma,delta 16 + BC RMW 16)
  + 4 * moved     (dist write)
scanned counts every entry of every visited list, including lists whose
entries the kernel resolves by table lookups without reading them, so the
algorithmic rate is not a bandwidth and can exceed the HBM peak.

`python bench.py --traffic-from DIR` turns rocprofv3 runs of this script
(DIR/trace, DIR/fetch, DIR/write, DIR/cfetch, DIR/cwrite) into
profiles/traffic_<workload>.json and profiles/window_kernel_<workload>.md.
"""
from __future__ import annotations

import argparse
import csv
import ctypes
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "edge insertions/sec with BC updated (K roots) vs plain GPU; p99 update latency"
TRAIN_WINDOWS = 30

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


# ------------------------------------------------------------------ host CPU

def host_info() -> dict:
    """CPU model and core counts of this host, and the share this process may
    use (its affinity mask, capped by OMP_NUM_THREADS when that is set)."""
    model, phys = None, set()
    pid = core = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, val = line.partition(":")
                k, val = k.strip(), val.strip()
                if k == "model name" and model is None:
                    model = val
                elif k == "physical id":
                    pid = val
                elif k == "core id":
                    core = val
                elif not k and pid is not None:
                    phys.add((pid, core))
                    pid = core = None
        if pid is not None:
            phys.add((pid, core))
    except OSError:
        pass
    share = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(share, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else share
    return {"cpu_model": model, "physical_cores_machine": len(phys) or None, "logical_cpus_machine": os.cpu_count(),
            "cpu_share": share, "omp_num_threads": omp, "threads": threads}


def cpu_baseline(hg, roots, u, v, budget_s, threads, dels=()):
    """The reference's insertEdgeStreamingBC (BALANCE, as its run scripts
    use), COARSE on the stream's first insertions for about budget_s, then
    FINE on the next ones for about budget_s, each on `threads` OpenMP
    threads; then, if dels is given, its deleteEdgeStreamingBC (COARSE) on
    those edges for at most budget_s / 2.  Returns (insertion dict, deletion
    dict); the insertion dict says why when the reference is not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_py import Reference, ref_available
    if not ref_available():
        return {"status": "unavailable",
                "reason": "oracle/_ref/libsbcref.so is not built (it is compiled from the reference's sources "
                          "by oracle/build_ref.sh where they are present)"}, None
    ref = Reference(hg.rowptr, hg.col, roots, nt=threads)
    t_static = ref.static()
    legs = {}
    n = 0
    # COARSE may take at most half of the stream, so that FINE has insertions
    # left on a stream that the host gets through within the budget
    for gran, name, end in ((0, "COARSE", max(1, len(u) // 2)), (1, "FINE", len(u))):
        spent, a = 0.0, n
        while n < end and (spent < budget_s or n == a):
            spent += ref.insert(int(u[n]), int(v[n]), 1, gran)
            n += 1
        legs[name] = {"value": (n - a) / spent if spent > 0 else None, "insertions": n - a,
                      "seconds": round(spent, 2),
                      "sample": f"insertions {a}..{n - 1} of the same stream, {spent:.1f} s inside "
                                f"insertEdgeStreamingBC ({name}, BALANCE)"}
    best = max((k for k in legs if legs[k]["value"]), key=lambda k: legs[k]["value"])
    ins = {"value": legs[best]["value"], "unit": "insertions/s", "cores": threads, "kind": "reference",
           "granularity": best,
           "sample": f"{legs[best]['sample']}; the faster of COARSE and FINE; static pass {t_static:.1f} s untimed",
           "coarse": legs["COARSE"], "fine": legs["FINE"]}
    dl = None
    if len(dels):
        from oracle_py import ref_lib
        L = ref_lib()
        cnt = np.zeros(3, dtype=np.int64)
        dspent, nd = 0.0, 0
        for a, b in dels:
            if nd and dspent >= budget_s / 2:
                break
            t = L.ref_delete(ref.h, int(a), int(b), 1, 0, cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
            if t < 0:
                continue
            dspent += t
            nd += 1
        if nd:
            dl = {"value": nd / dspent, "unit": "deletions/s", "cores": threads, "kind": "reference",
                  "sample": f"first {nd} of the same deletions after the {n} insertions above, {dspent:.1f} s "
                            f"inside deleteEdgeStreamingBC (COARSE, BALANCE)"}
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


# ------------------------------------------------------------------ rocprofv3 outputs

def _rows(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    out = []
    for p in files:
        with open(p, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def _short(name):
    return name.split("(")[0].replace("void ", "").strip()


def _dispatches(rows):
    """Per dispatch (in submission order): kernel, grid size, and the counter
    values / timestamps of the rows."""
    by = {}
    for r in rows:
        did = int(r["Dispatch_Id"])
        d = by.setdefault(did, {"kernel": _short(r["Kernel_Name"]), "grid": int(r.get("Grid_Size", 0) or 0)})
        if "Counter_Name" in r:
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        if "Start_Timestamp" in r and r["Start_Timestamp"]:
            d["ns"] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    return [by[k] for k in sorted(by)]


def _window_launches(ds):
    """The window kernel's main launch of each window, in stream order: the
    dispatch that follows a window graph's append and queue reset (resumed
    roots' launches follow the whole-GPU kernels instead)."""
    return [d for i, d in enumerate(ds)
            if d["kernel"].startswith("sbc_window_kernel") and i >= 2
            and ds[i - 1]["kernel"] == "sbc_reset_kernel" and ds[i - 2]["kernel"] == "sbc_append_kernel"]


def traffic_from(d, workload, bench_json, commands):
    """Writes profiles/traffic_<workload>.json and the markdown summary from
    rocprofv3 runs under d: trace (kernel trace + stats), fetch / write (one
    PMC counter each, same bench command), cfetch / cwrite (a 1 GiB copy)."""
    trace = _dispatches(_rows(os.path.join(d, "trace"), "kernel_trace.csv"))
    fetch = _window_launches(_dispatches(_rows(os.path.join(d, "fetch"), "counter_collection.csv")))
    write = _window_launches(_dispatches(_rows(os.path.join(d, "write"), "counter_collection.csv")))
    tw = _window_launches(trace)
    with open(bench_json) as fh:
        line = [l for l in fh if l.startswith("{")][-1]
    b = json.loads(line)
    # the main forest's windows come first; later legs (the plain form's
    # replay and insertions) may match the same launch pattern
    n_main = b["timed_windows"][1] + 1
    fetch, write, tw = fetch[:n_main], write[:n_main], tw[:n_main]
    if not (len(fetch) == len(write) == len(tw) == n_main):
        raise SystemExit(f"window launches differ between runs: trace {len(tw)} fetch {len(fetch)} write {len(write)}")

    def copy_kib(sub, counter):
        ds = _dispatches(_rows(os.path.join(d, sub), "counter_collection.csv"))
        vals = [x[counter] for x in ds if counter in x and "copy" in x["kernel"].lower()]
        return float(np.median(vals))

    nbytes_copy = 1 << 30
    ff = nbytes_copy / 1024.0 / copy_kib("cfetch", "FETCH_SIZE")
    wf = nbytes_copy / 1024.0 / copy_kib("cwrite", "WRITE_SIZE")
    fk = [x["FETCH_SIZE"] for x in fetch]
    wk = [x["WRITE_SIZE"] for x in write]
    ns = [x["ns"] for x in tw]
    t0 = b["timed_windows"][0]
    t1 = t0 + b["steps"]
    out = {
        "kernel": "sbc_window_kernel<16, false>",
        "workload": workload,
        "windows": f"window indices 0..{len(fk) - 1} of the stream (main launch of each window)",
        "fetch_factor": round(ff, 4),
        "write_factor": round(wf, 4),
        "per_window": {"fetch_kib": [round(x, 1) for x in fk], "write_kib": [round(x, 1) for x in wk],
                       "trace_us": [round(x / 1e3, 1) for x in ns]},
        "commands": commands,
    }
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    # summary
    stats = {}
    for x in trace:
        s = stats.setdefault(x["kernel"], [0, 0])
        s[0] += 1
        s[1] += x.get("ns", 0)
    top = sorted(stats.items(), key=lambda kv: -kv[1][1])[:10]
    tb = np.array([(fk[i] * ff + wk[i] * wf) * 1024.0 for i in range(t0, min(t1, len(fk)))])
    tus = np.array(ns[t0:t1]) / 1e3
    md = [f"# sbc_window_kernel on {WORKLOADS[workload][4]} (MI355X, 1 GPU)", "",
          "Commands run from the repository root with TMPDIR=/tmp and SBC_AUTOTUNE_CACHE pointing at one",
          "file, so every run uses the launch parameters the first run's autotuner chose.  The runs use the",
          "identity root order (`--schedule static`): with K <= compute units every order runs the same",
          "kernel work, and the default learned order only adds report predictions.", "",
          "## Kernel trace", "", "    " + commands[0], "",
          "| kernel | dispatches | total ms | avg us |", "|---|---|---|---|"]
    for k, (c, t) in top:
        md.append(f"| {k} | {c} | {t / 1e6:.2f} | {t / 1e3 / c:,.1f} |")
    main_us = np.array(ns) / 1e3
    md.append(f"| sbc_window_kernel<16, false>, window launches of the main forest | {len(ns)} | "
              f"{main_us.sum() / 1e3:.2f} | {main_us.mean():,.1f} |")
    md += ["", "(The other window-kernel dispatches are the plain form's insertions, its replay of the leading",
           "windows and the launches that resume roots after a hand-over.)"]
    md += ["",
           f"Window-kernel main launches of the timed windows {t0}..{t1 - 1} (the bench's own timed range in "
           f"the same command): rocprof average {tus.mean():,.1f} us; the bench's HIP events around each "
           f"window graph (append, queue reset, window kernel) in that run: "
           f"{1e3 * b['roofline']['launch_ms']:,.1f} us.",
           f"That run: {b['value']:,.1f} insertions/s, {b['ms_per_step']:.3f} ms per window.", "",
           "## HBM traffic (PMC, one counter per pass, same command and stream)", ""]
    md += ["    " + c for c in commands[1:]]
    md += ["",
           f"Calibration: a 1 GiB device copy reads FETCH_SIZE = 1/{ff:.3f} of its bytes and WRITE_SIZE = "
           f"1/{wf:.3f}; those factors turn the counters into bytes.",
           f"Timed windows {t0}..{t1 - 1}: {tb.mean() / 1e9:.3f} GB of HBM traffic per window-kernel launch "
           f"(fetch {np.mean(fk[t0:t1]) * ff / 2**20:.3f} GiB, write {np.mean(wk[t0:t1]) * wf / 2**20:.3f} GiB), "
           f"{tb.mean() / (tus.mean() * 1e-6) / 1e9:,.1f} GB/s over the rocprof launch time = "
           f"{tb.mean() / (tus.mean() * 1e-6) / 1e9 / HBM_PEAK_GBS:.3f} of the 8 TB/s peak.",
           f"All {len(fk)} windows: {np.mean([(a * ff + c * wf) * 1024 for a, c in zip(fk, wk)]) / 1e9:.3f} GB per "
           f"launch on average.", ""]
    mpath = os.path.join(ROOT, "profiles", f"window_kernel_{workload}.md")
    with open(mpath, "w") as fh:
        fh.write("\n".join(md))
    print(f"wrote {path} and {mpath}")


def roofline_from_profile(workload, t0, steps, launch_ms, block_threads, W):
    """HBM bytes per window-kernel launch over window indices t0..t0+steps-1,
    from profiles/traffic_<workload>.json, and the rate over launch_ms."""
    tpath = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(tpath) or block_threads != 1024 or W != WORKLOADS[workload][3] or launch_ms <= 0:
        return None
    with open(tpath) as fh:
        t = json.load(fh)
    if "per_window" not in t:
        return None
    pw = t["per_window"]
    n = len(pw["fetch_kib"])
    b = [(pw["fetch_kib"][i] * t["fetch_factor"] + pw["write_kib"][i] * t["write_factor"]) * 1024.0
         for i in range(n)]
    if t0 + steps <= n:
        traffic = float(np.mean(b[t0:t0 + steps]))
        basis = f"windows {t0}..{t0 + steps - 1}, the ones timed here"
    else:
        traffic = float(np.mean(b))
        basis = f"mean over windows 0..{n - 1} (the timed range {t0}..{t0 + steps - 1} was not profiled)"
    return traffic, basis, f"profiles/traffic_{workload}.json"


# ------------------------------------------------------------------ main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="rmat20_k256_w128", choices=sorted(WORKLOADS))
    ap.add_argument("--window", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--heavy", type=int, default=-2, help="-2: autotuned/default, -1: all heavy")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--plain", type=int, default=-1,
                    help="insertions timed in the plain form (-1: every insertion of the timed windows, 0: none)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of reference update time per granularity")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--schedule", default="mlp", choices=["static", "mlp"],
                    help="root order per window: by cost-model prediction (trained on the leading windows; with at "
                         "most one team per compute unit every order is equivalent and only report windows are "
                         "predicted), or identity")
    ap.add_argument("--train-windows", type=int, default=TRAIN_WINDOWS,
                    help="untimed leading windows (the cost model's training data under --schedule mlp)")
    ap.add_argument("--light-max", type=float, default=0.0,
                    help="--schedule mlp: roots predicted below this many affected vertices per window run light")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-parameter search (team size, hand-over thresholds) "
                         "that otherwise runs on the first untimed leading windows")
    ap.add_argument("--park-min", type=int, default=0,
                    help="descent frontier size at which an item is finished by whole-GPU kernels "
                         "(0: library default, <0: never)")
    ap.add_argument("--deletions", type=int, default=16,
                    help="edges of the initial graph deleted (and timed) after the insertion legs")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (PMC traffic calibration)")
    ap.add_argument("--traffic-from", metavar="DIR",
                    help="only write profiles/ from rocprofv3 runs under DIR (see the module docstring)")
    ap.add_argument("--bench-json", metavar="FILE", help="with --traffic-from: this script's output in DIR/trace's run")
    ap.add_argument("--commands", metavar="FILE", help="with --traffic-from: the profiling commands, one per line")
    args = ap.parse_args()

    if args.traffic_from:
        with open(args.commands) as fh:
            cmds = [l.strip() for l in fh if l.strip()]
        traffic_from(args.traffic_from, args.workload, args.bench_json, cmds)
        return

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
    from streamingbc_amd.autotune import tuning_windows
    # the search takes its windows from the untimed leading ones
    n_train = max(args.train_windows, 0 if args.no_autotune else tuning_windows(ordered=False))
    n_ins = (n_train + args.warmup + args.steps) * W
    hg, roots, u, v, K, _, desc, t_gen = make_workload(args.workload, n_ins)
    print(f"[bench] {args.workload}: nv={hg.nv} directed slots={hg.ne} K={K} W={W} "
          f"insertions={len(u)} gen={t_gen:.1f}s", file=sys.stderr, flush=True)

    ctx = sbc.Context(hg.nv)
    g = sbc.Graph(hg)
    f = sbc.Forest(hg.nv, roots)
    ctx.reset_stats(True)
    t0 = time.time()
    f.init_static(g, ctx)
    t_static = time.time() - t0
    _, static_ms = ctx.kernel_stats(sbc.KSTAT_STATIC)
    # two passes over every adjacency entry per root: neighbour id and its level, 4 B each
    static_alg = 2.0 * hg.ne * 8.0 * K
    static_pass = {"s": round(t_static, 3), "kernel_ms": round(static_ms, 3),
                   "algorithmic_GBs": round(static_alg / (static_ms * 1e-3) / 1e9, 1) if static_ms > 0 else None,
                   "algorithmic_frac": round(static_alg / (static_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                   if static_ms > 0 else None}
    print(f"[bench] static pass {t_static:.2f}s (kernel {static_ms:.1f} ms)", file=sys.stderr, flush=True)

    cfg = sbc.BatchConfig(window=W, block_threads=args.block or 1024, use_graph=not args.no_graph,
                          park_min=args.park_min)
    if args.heavy >= -1:
        cfg.heavy_roots = args.heavy
    pos = 0
    tune = None
    if not args.no_autotune:
        from streamingbc_amd.autotune import autotune
        res, used = autotune(f, g, ctx, u, v, W, hg.ne, base=cfg, cache=os.environ.get("SBC_AUTOTUNE_CACHE"))
        pos += used
        cfg.block_threads, cfg.grid_blocks = res.block_threads, res.grid_blocks
        cfg.park_min, cfg.park_scan = res.park_min, res.park_scan
        if args.heavy < -1:
            cfg.heavy_roots = res.heavy_roots
        tune = {"block_threads": res.block_threads, "heavy_roots": res.heavy_roots, "park_min": res.park_min,
                "park_scan": res.park_scan, "windows": used // W, "from_cache": res.from_cache, "table": res.table}
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)

    from streamingbc_amd.schedule import LearnedSchedule, StaticSchedule
    if args.schedule == "mlp":
        # it learns from the leading windows the autotuner left
        sched = LearnedSchedule(train_after=max(1, n_train - pos // W), light_max=args.light_max, eval_every=8)
    else:
        sched = StaticSchedule()

    def window(a, b, records=False):
        plan, X = sched.plan(f, g, ctx, u[a:b], v[a:b])
        c = sbc.BatchConfig(window=W, block_threads=cfg.block_threads, grid_blocks=cfg.grid_blocks,
                            use_graph=cfg.use_graph, heavy_roots=cfg.heavy_roots if plan.heavy_roots < 0
                            else plan.heavy_roots, root_order=plan.root_order, park_min=cfg.park_min,
                            use_context_order=plan.device_order, park_scan=cfg.park_scan)
        rec = f.insert_batched(g, ctx, u[a:b], v[a:b], c, records=records or plan.records)
        parked[0] += ctx.last_parked()
        sched.observe(plan, X, rec, ctx, K)
        return rec

    parked = [0]
    # the rest of the leading windows (the cost model's training data) and the
    # warmup windows, untimed
    for _ in range(n_train - pos // W + args.warmup):
        window(pos, pos + W)
        pos += W
    ctx.synchronize()

    # timed windows
    parked[0] = 0
    timed_first_window = pos // W
    timed_first_ins = pos
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
    n_wide, w_ms = ctx.kernel_stats(sbc.KSTAT_WIDE)
    # work of the timed windows, summed per root on the device (the item
    # records' fields without copying the records in the timed region)
    tot = ctx.work_totals(K)
    n_items = args.steps * W * K
    alg_bytes = float(8 * n_items + 8 * tot["processed"].sum() + 8 * tot["scanned"].sum()
                      + 48 * tot["affected"].sum() + 4 * tot["moved"].sum())
    per_root = tot["scanned"].astype(np.float64)
    value = args.steps * W / t_batched
    launch_ms = k_ms / max(n_launch, 1)
    print(f"[bench] batched: {value:.1f} ins/s, window {1e3 * t_batched / args.steps:.2f} ms, "
          f"window-kernel launch avg {launch_ms:.3f} ms, hand-over kernels {w_ms / max(args.steps, 1):.3f} ms "
          f"per window", file=sys.stderr, flush=True)

    # plain form on the same insertions, from the same state: a second forest
    # replays the untimed leading windows with the batched form, then takes
    # the timed windows' insertions one at a time
    plain = None
    n_plain = args.steps * W if args.plain < 0 else min(args.plain, args.steps * W)
    if n_plain > 0:
        ctx2 = sbc.Context(hg.nv)
        g2 = sbc.Graph(hg)
        f2 = sbc.Forest(hg.nv, roots)
        f2.init_static(g2, ctx2)
        f2.insert_batched(g2, ctx2, u[:timed_first_ins], v[:timed_first_ins],
                          sbc.BatchConfig(window=W, block_threads=1024))
        ctx2.synchronize()
        ctx2.reset_stats(timing=True)
        pparked = 0
        tp = time.perf_counter()
        for i in range(timed_first_ins, timed_first_ins + n_plain):
            f2.insert_edge(g2, ctx2, int(u[i]), int(v[i]), sbc.FINE)
            pparked += ctx2.last_parked()
        ctx2.synchronize()
        t_plain = time.perf_counter() - tp
        pn, pms = ctx2.kernel_stats(sbc.KSTAT_UPDATE)
        _, pwms = ctx2.kernel_stats(sbc.KSTAT_WIDE)
        plain = {"value": n_plain / t_plain, "unit": "insertions/s", "insertions": n_plain,
                 "same_insertions_as_timed_windows": True,
                 "kernel_ms_avg": pms / max(pn, 1), "handover_ms_total": round(pwms, 3), "handed_over": pparked}
        if n_plain == args.steps * W:
            # both forests now hold the same stream: their trees must agree
            same = all(np.array_equal(a, b) for ri in (0, K // 2, K - 1)
                       for a, b in zip(f.tree(ri), f2.tree(ri)))
            plain["trees_equal_batched"] = bool(same)
        f2.close()
        g2.close()
        ctx2.close()
        print(f"[bench] plain: {plain['value']:.1f} ins/s on {n_plain} insertions", file=sys.stderr, flush=True)

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
        # movement-case roots are updated in place: one workgroup each below
        # 4M adjacency slots, the delta pass over the whole GPU above
        wide_delta = hg.ne >= (4 << 20)
        deletion = {"value": round(len(dels) / t_del, 3), "unit": "deletions/s", "deletions": len(dels),
                    "ms_p50": round(1e3 * float(np.percentile(dlat, 50)), 3),
                    "cases": {"sameLevel": int(cases[0]), "adjacent": int(cases[1]), "movement": int(cases[2])},
                    "update_kernel_ms_avg": round(ums / max(un, 1), 3),
                    "movement_kernel_ms_avg": round(sms / max(sn, 1), 3),
                    "method": "adjacent-case roots: update kernel with negated added paths; movement-case "
                              "roots: updated in place (" + ("delta pass over the whole GPU, one launch per level"
                                                            if wide_delta else "one workgroup each")
                              + "); BC updated by the same changes"}
        print(f"[bench] deletions: {deletion['value']:.2f} del/s, cases {deletion['cases']}",
              file=sys.stderr, flush=True)

    alg_rate = alg_bytes / max(n_launch, 1) / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    prof = roofline_from_profile(args.workload, timed_first_window, args.steps, launch_ms, cfg.block_threads, W)
    roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": "sbc_window_kernel", "launch_ms": round(launch_ms, 4),
                "algorithmic": {"GBs": round(alg_rate, 2), "bytes_per_launch": alg_bytes / max(n_launch, 1),
                                "note": "8 + 8*processed + 8*scanned + 48*affected + 4*moved per item; scanned "
                                        "counts list entries the kernel resolves by table lookups without reading "
                                        "them, so this rate is not a bandwidth"}}
    if prof:
        traffic, basis, src = prof
        ach = traffic / (launch_ms * 1e-3) / 1e9
        if ach > HBM_PEAK_GBS:
            raise SystemExit(f"bench.py: measured HBM rate {ach:.0f} GB/s exceeds the peak: stale profile {src}?")
        roofline.update({"achieved": round(ach, 2), "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": round(traffic),
                         "basis": "HBM bytes per window-kernel launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated, "
                                  f"{basis}) / the launch time from HIP events in this run",
                         "traffic_source": src})
    else:
        roofline["basis"] = "no PMC profile of this workload and launch shape in profiles/: traffic unmeasured"

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
        "timed_windows": [timed_first_window, timed_first_window + args.steps - 1],
        "latency_ms": {"window_p50": round(1e3 * float(np.percentile(lat, 50)), 4),
                       "window_p99": round(1e3 * float(np.percentile(lat, 99)), 4)},
        "plain_gpu": plain,
        "speedup_vs_plain": round(value / plain["value"], 3) if plain else None,
        "roofline": roofline,
        "handover_ms_per_window": round(w_ms / max(args.steps, 1), 4),
        "items": {"count": n_items, "nonzero_gap": int(tot["nonzero"].sum()),
                  "root_scan_max_over_mean": float(per_root.max() / max(per_root.mean(), 1)),
                  "mean_affected": float(tot["affected"].sum() / n_items), "max_affected": int(tot["max_affected"].max()),
                  "mean_scanned": float(tot["scanned"].sum() / n_items),
                  "parked": int(parked[0])},
        "static_pass": static_pass,
        "deletion": deletion,
        "schedule": {"name": sched.name, **(sched.report() if args.schedule == "mlp" else {})},
        "autotune": tune,
    }

    if not args.no_cpu_baseline:
        hi = host_info()
        f.close()
        g.close()
        cb, cdl = cpu_baseline(hg, roots, u, v, args.cpu_budget, hi["threads"], dels)
        cb["host"] = hi
        out["cpu_baseline"] = cb
        if "value" in cb:
            out["speedup_vs_cpu"] = round(value / cb["value"], 2)
        if cdl and deletion:
            deletion["cpu_baseline"] = cdl
            deletion["speedup_vs_cpu"] = round(deletion["value"] / cdl["value"], 2)
    else:
        out["cpu_baseline"] = {"status": "skipped", "reason": "--no-cpu-baseline"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
