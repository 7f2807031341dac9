"""Parity at the full sizes of the benchmark configurations.

  C2  Chung-Lu 32,768 vertices (exponent 2.3, mean degree 8, largest
      component), K=256 roots (seed 7), 1,024 insertions applied one at a time
      in the plain form: every root's level, sigma and delta bit-exact against
      the oracle's FINE order, item records equal, BC within 1e-9 after every
      64 insertions.
  C3  R-MAT scale 18 (16 draws/vertex, deduplicated), K=512, windows of 64.
  C4  R-MAT scale 20 (16 draws/vertex, deduplicated), K=256, windows of 128.
      For both: the full graph and root set on the device, 2,560 insertions in
      batched windows with the library's default hand-over thresholds; the
      oracle follows a seeded sample of 8 roots (the static pass, every item
      record and the final trees bit-exact), BC from the device's atomics
      agrees with the deterministic recompute within 1e-9, and no sigma
      overflows.
  C5  the six serving graphs (Chung-Lu 2^14..2^20, 128-512 roots): each served
      alone and all six in one mix at an aggregate 1K insertions/s end
      bit-identical.

The oracle runs on host threads (one forest per group of roots sharing the
graph); every root's tree is the single-forest oracle's.
"""
import numpy as np
import pytest

from oracle_py import FINE as O_FINE, PULL, OracleGroup
from sbc_testutil import assert_bc_close

pytestmark = pytest.mark.gpu

sbc = pytest.importorskip("streamingbc_amd")

KEYS = ("kase", "affected", "moved", "processed", "scanned")


def _engine(hg, roots):
    ctx = sbc.Context(hg.nv)
    g = sbc.Graph(hg)
    f = sbc.Forest(hg.nv, roots)
    f.init_static(g, ctx)
    return ctx, g, f


def _same_tree(f, o, ri, oi, what):
    lv, sg, dl = f.tree(ri)
    lo, so, do = o.tree(oi)
    assert np.array_equal(lv, lo), f"{what}: root index {ri} level differs at {np.flatnonzero(lv != lo)[:5]}"
    assert np.array_equal(sg, so), f"{what}: root index {ri} sigma differs at {np.flatnonzero(sg != so)[:5]}"
    bad = np.flatnonzero(dl != do)
    assert bad.size == 0, f"{what}: root index {ri} delta differs at {bad[:5]}"
    # sigma of a reachable vertex is a positive path count: a wrapped int64
    # shows up negative (or zero)
    reach = lv != np.iinfo(np.int64).max
    assert sg[reach].min() >= 1, f"{what}: root index {ri} sigma wrapped"


def test_c2_chunglu15_plain_all_roots():
    hg = sbc.gen_chung_lu(1 << 15, 2.3, 8.0, seed=1, keep_lcc=True)
    roots = hg.pick_roots(256, 7)
    u, v = hg.insertions(1024, 11)
    assert len(u) == 1024
    print(f"[C2] chunglu15 nv={hg.nv} slots={hg.ne} K={len(roots)} insertions={len(u)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = OracleGroup(hg.rowptr, hg.col, roots)
    o.static(PULL)
    for ri in range(len(roots)):
        _same_tree(f, o, ri, ri, "C2 static pass")
    assert_bc_close(f.bc(), o.bc(), what="C2 static BC")
    for i, (a, b) in enumerate(zip(u, v)):
        f.insert_edge(g, ctx, int(a), int(b), sbc.FINE)
        rec_o = o.insert(int(a), int(b), O_FINE, items=True)
        rec_g = ctx.last_records(f.k)
        for key in KEYS:
            assert np.array_equal(rec_g[key], rec_o[key]), f"C2 insertion {i}: {key}"
        if (i + 1) % 64 == 0:
            assert_bc_close(f.bc(), o.bc(), what=f"C2 BC after {i + 1} insertions")
        if (i + 1) % 256 == 0:
            for ri in range(len(roots)):
                _same_tree(f, o, ri, ri, f"C2 after {i + 1} insertions")
            print(f"[C2] {i + 1} insertions: trees of all {len(roots)} roots bit-exact", flush=True)
    assert not o.sigma_overflowed()


BATCHED = {
    # name: generator kwargs, K, window
    "C3_rmat18_k512_w64": (dict(scale=18, draws=16 << 18), 512, 64),
    "C4_rmat20_k256_w128": (dict(scale=20, draws=16 << 20), 256, 128),
}


@pytest.mark.parametrize("name", sorted(BATCHED))
def test_batched_full_size_sampled_roots(name):
    kw, K, W = BATCHED[name]
    n = 2560
    hg = sbc.gen_rmat(kw["scale"], kw["draws"], seed=1, keep_lcc=False)
    roots = hg.pick_roots(K, 7)
    u, v = hg.insertions(n, 11)
    sample = np.sort(np.random.default_rng(5).choice(K, 8, replace=False))
    print(f"[{name}] nv={hg.nv} slots={hg.ne} max degree={hg.degree().max()} K={K} W={W} insertions={n} "
          f"sampled root indices={sample.tolist()}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = OracleGroup(hg.rowptr, hg.col, roots[sample])
    o.static(PULL)
    for j, ri in enumerate(sample):
        _same_tree(f, o, int(ri), j, f"{name} static pass")
    cfg = sbc.BatchConfig(window=W, block_threads=1024)
    parked = 0
    for w0 in range(0, n, W):
        rec = f.insert_batched(g, ctx, u[w0:w0 + W], v[w0:w0 + W], cfg, records=True)
        parked += ctx.last_parked()
        for i in range(rec.shape[0]):
            ro = o.insert(int(u[w0 + i]), int(v[w0 + i]), O_FINE, items=True)
            for key in KEYS:
                assert np.array_equal(rec[i][sample][key], ro[key]), f"{name} insertion {w0 + i}: {key}"
    for j, ri in enumerate(sample):
        _same_tree(f, o, int(ri), j, f"{name} after {n} insertions")
    assert not o.sigma_overflowed(), f"{name}: sigma overflowed int64 in the oracle"
    bc_atomic = f.bc()
    f.recompute_bc(ctx)
    assert_bc_close(bc_atomic, f.bc(), what=f"{name} BC from atomics vs recompute")
    print(f"[{name}] {n} insertions: sampled trees bit-exact, items handed to the wide kernels: {parked}",
          flush=True)


def test_c5_six_graphs_alone_vs_mix():
    from streamingbc_amd.serve import SERVE_GRAPHS, Server, poisson_arrivals
    rate, seconds = 1000.0, 2.0
    count = int(rate * seconds)
    gs = {}
    for sc, k in SERVE_GRAPHS:
        hg = sbc.gen_chung_lu(1 << sc, 2.3, 8.0, seed=sc, keep_lcc=True)
        gs[f"chunglu{sc}_k{k}"] = (hg, hg.pick_roots(min(k, hg.nv), sc))
    t = poisson_arrivals(rate, count, 11)
    owner = np.random.default_rng(11).integers(0, len(gs), size=count)
    arr = {}
    for gi, (name, (hg, _)) in enumerate(gs.items()):
        sel = owner == gi
        pu, pv = hg.insertions(int(sel.sum()), 11)
        arr[name] = (t[sel], pu, pv)
    print("[C5] graphs " + ", ".join(f"{n} nv={h.nv} K={len(r)}" for n, (h, r) in gs.items()), flush=True)
    alone = {}
    for name, (hg, roots) in gs.items():
        srv = Server(budget_bytes=200 << 30)
        srv.admit(name, hg, roots)
        srv.serve({name: arr[name]})
        r = srv.resident[name]
        assert r.ctx.graph_counts()[1] == len(r.windows), f"{name}: a window ran without its graph"
        alone[name] = (srv.bc(name), [r.f.tree(i) for i in range(0, len(roots), 37)])
        srv.evict(name)
        print(f"[C5] {name} alone: {len(arr[name][0])} insertions", flush=True)
    srv = Server(budget_bytes=200 << 30)
    for name, (hg, roots) in gs.items():
        srv.admit(name, hg, roots)
    res = srv.serve(arr)
    print(f"[C5] mix at {rate:.0f}/s: p50 {res['p50_ms']} ms p99 {res['p99_ms']} ms "
          f"sustained {res['sustained_ins_per_s']}/s", flush=True)
    assert res["insertions"] == count
    for name, (hg, roots) in gs.items():
        r = srv.resident[name]
        assert r.ctx.graph_counts()[1] == len(r.windows), f"{name}: a window ran without its graph"
        assert r.applied == len(arr[name][0])
        assert np.array_equal(srv.bc(name), alone[name][0]), name
        for j, ri in enumerate(range(0, len(roots), 37)):
            for a, b in zip(r.f.tree(ri), alone[name][1][j]):
                assert np.array_equal(a, b), f"{name} root index {ri}"
