"""GPU parity: the HIP engine against the CPU oracle on the same seeded inputs.

Bar: dist (level) and sigma bit-exact; per-root delta bit-exact against the
oracle's FINE-order restatement (the engine sums each vertex's terms in
adjacency order, with FP contraction off); BC within 1e-9 relative (floor
1e-6 x max|BC|) because the device adds per-root BC changes with atomics in
arbitrary order.
"""
import os

import numpy as np
import pytest

from oracle_py import COARSE as O_COARSE, FINE as O_FINE, MOVE_IN_PLACE, MOVE_REBUILD, PULL, PUSH, Oracle
from sbc_testutil import LONG_MAX, assert_bc_close, karate_path

pytestmark = pytest.mark.gpu

sbc = pytest.importorskip("streamingbc_amd")


def _engine(hg, roots):
    ctx = sbc.Context(hg.nv)
    g = sbc.Graph(hg)
    f = sbc.Forest(hg.nv, roots)
    f.init_static(g, ctx)
    return ctx, g, f


def _compare_trees(f, o, what, delta_exact=True):
    lv, sg, dl = f.trees()
    lo, so, do = o.trees()
    assert np.array_equal(lv, lo), f"{what}: level mismatch at {np.argwhere(lv != lo)[:5]}"
    assert np.array_equal(sg, so), f"{what}: sigma mismatch at {np.argwhere(sg != so)[:5]}"
    if delta_exact:
        bad = np.argwhere(dl != do)
        assert bad.size == 0, f"{what}: delta differs at {bad[:5]} gpu={dl[tuple(bad[0])]!r} orc={do[tuple(bad[0])]!r}"
    else:
        assert_bc_close(dl, do, what=f"{what} delta")


def _odel(o, a, b):
    """The oracle's deletion with the movement case the device runs: in place
    by default (every graph here is below the size from which the device
    rebuilds instead), rebuilt under SBC_DELETE_MOVE=rebuild."""
    move = MOVE_REBUILD if os.environ.get("SBC_DELETE_MOVE") == "rebuild" else MOVE_IN_PLACE
    return o.delete(a, b, move)


def _rmat_c1():
    hg = sbc.gen_rmat(10, 8192, seed=1, keep_lcc=True)
    u, v = hg.insertions(256, 1)
    return hg, np.arange(hg.nv), u, v


@pytest.mark.parametrize("name", ["karate", "rmat10"])
def test_static_matches_oracle(name):
    if name == "karate":
        hg = sbc.read_dimacs(karate_path())
        roots = np.arange(hg.nv)
    else:
        hg, roots, _, _ = _rmat_c1()
    print(f"[static] {name} nv={hg.nv} k={len(roots)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    _compare_trees(f, o, f"static {name}")
    # same fold order as the oracle's (root order): BC is bit-identical
    assert np.array_equal(f.bc(), o.bc())
    # and the reference's push order agrees within tolerance
    op = Oracle(hg.rowptr, hg.col, roots)
    op.static(PUSH)
    assert_bc_close(f.bc(), op.bc())


def _static_trees(hg, roots, form, monkeypatch):
    if form == "vertex":
        monkeypatch.setenv("SBC_STATIC_FORM", "vertex")
    else:
        monkeypatch.delenv("SBC_STATIC_FORM", raising=False)
    ctx, g, f = _engine(hg, roots)
    out = f.trees(), f.bc()
    f.close()
    g.close()
    ctx.close()
    return out


# nv 16K and 2.5K (small LDS bitmap), 524K (large bitmap), 1.2M (past the
# bitmap: both forms are the per-vertex kernel)
@pytest.mark.parametrize("name", ["rmat14", "chunglu_multi", "rmat19", "sparse_1_2m"])
def test_static_bitmap_form_matches_vertex_form(name, monkeypatch):
    if name == "rmat14":
        hg = sbc.gen_rmat(14, 16 << 14, seed=2, keep_lcc=False)
        roots = hg.pick_roots(64, 3)
    elif name == "chunglu_multi":
        # several components; roots also outside the largest one
        hg = sbc.gen_chung_lu(2500, 2.3, 2.0, seed=4, keep_lcc=False)
        roots = hg.pick_roots(48, 5, from_lcc=False)
    elif name == "rmat19":
        hg = sbc.gen_rmat(19, 16 << 19, seed=1, keep_lcc=False)
        roots = hg.pick_roots(16, 7)
    else:
        n = 1_200_000
        rng = np.random.default_rng(9)
        eu = rng.integers(1, 2000, 6000)
        ev = rng.integers(1, 2000, 6000)
        keep = eu != ev
        hg = sbc.graph_from_edges(n, eu[keep], ev[keep])
        roots = hg.pick_roots(8, 1)
    print(f"[static forms] {name} nv={hg.nv} ne={hg.ne} k={len(roots)}", flush=True)
    (lb, sb_, db), bcb = _static_trees(hg, roots, "bitmap", monkeypatch)
    (lv, sv, dv), bcv = _static_trees(hg, roots, "vertex", monkeypatch)
    assert np.array_equal(lb, lv) and np.array_equal(sb_, sv)
    bad = np.argwhere(db != dv)
    assert bad.size == 0, f"delta differs at {bad[:5]}"
    assert np.array_equal(bcb, bcv)
    if hg.nv < 20000:
        o = Oracle(hg.rowptr, hg.col, roots)
        o.static(PULL)
        lo, so, do = o.trees()
        assert np.array_equal(lb, lo) and np.array_equal(sb_, so) and np.array_equal(db, do)


@pytest.mark.parametrize("gran", [sbc.FINE, sbc.COARSE])
def test_plain_insert_karate(gran):
    hg = sbc.read_dimacs(karate_path())
    roots = np.arange(hg.nv)
    u, v = hg.insertions(20, 5)
    print(f"[plain] karate gran={gran} n={len(u)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    for a, b in zip(u, v):
        f.insert_edge(g, ctx, int(a), int(b), gran)
        o.insert(int(a), int(b), O_FINE)
    _compare_trees(f, o, "karate plain")
    assert_bc_close(f.bc(), o.bc())


def test_plain_insert_rmat10_exact():
    hg, roots, u, v = _rmat_c1()
    print(f"[plain] rmat10 exact nv={hg.nv} n={len(u)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    for i, (a, b) in enumerate(zip(u, v)):
        f.insert_edge(g, ctx, int(a), int(b), sbc.FINE)
        rec_o = o.insert(int(a), int(b), O_FINE, items=True)
        if i % 64 == 0:
            rec_g = ctx.last_records(f.k)
            assert np.array_equal(rec_g["kase"], rec_o["kase"])
            assert np.array_equal(rec_g["affected"], rec_o["affected"])
            assert np.array_equal(rec_g["moved"], rec_o["moved"])
            assert np.array_equal(rec_g["processed"], rec_o["processed"])
            assert np.array_equal(rec_g["scanned"], rec_o["scanned"])
    _compare_trees(f, o, "rmat10 plain")
    assert_bc_close(f.bc(), o.bc())


def test_update_equals_static_recompute_chunglu():
    hg = sbc.gen_chung_lu(4096, 2.3, 8.0, seed=3, keep_lcc=True)
    roots = hg.pick_roots(64, 7)
    u, v = hg.insertions(200, 9)
    print(f"[recompute] chunglu nv={hg.nv} k=64 n={len(u)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    for a, b in zip(u, v):
        f.insert_edge(g, ctx, int(a), int(b), sbc.FINE)
    # static pass of the oracle on the post-insertion graph
    o = Oracle(hg.rowptr, hg.col, roots)
    for a, b in zip(u, v):
        o.add_edge(int(a), int(b))
    o.static(PULL)
    lv, sg, dl = f.trees()
    lo, so, do = o.trees()
    assert np.array_equal(lv, lo)
    assert np.array_equal(sg, so)
    assert_bc_close(dl, do, what="delta vs recompute")
    assert_bc_close(f.bc(), o.bc(), what="bc vs recompute")


@pytest.mark.parametrize("window,block,heavy", [(8, 256, -1), (16, 64, 0), (5, 512, 10)])
def test_batched_matches_plain(window, block, heavy):
    hg = sbc.gen_chung_lu(4096, 2.3, 8.0, seed=5, keep_lcc=True)
    roots = hg.pick_roots(48, 7)
    u, v = hg.insertions(64, 11)
    print(f"[batched] w={window} bt={block} heavy={heavy} nv={hg.nv}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    cfg = sbc.BatchConfig(window=window, block_threads=block, heavy_roots=heavy, use_graph=True)
    rec = f.insert_batched(g, ctx, u, v, cfg, records=True)
    for i, (a, b) in enumerate(zip(u, v)):
        ro = o.insert(int(a), int(b), O_FINE, items=True)
        assert np.array_equal(rec[i]["kase"], ro["kase"]), f"insertion {i}"
        assert np.array_equal(rec[i]["affected"], ro["affected"]), f"insertion {i}"
    _compare_trees(f, o, "batched")
    assert_bc_close(f.bc(), o.bc())
    # deterministic BC readout equals the oracle's root-order sum within tolerance
    f.recompute_bc(ctx)
    assert_bc_close(f.bc(), o.bc())


def test_batched_schedule_invariance():
    hg = sbc.gen_chung_lu(4096, 2.3, 8.0, seed=6, keep_lcc=True)
    roots = hg.pick_roots(40, 2)
    u, v = hg.insertions(48, 4)
    print("[invariance] two schedules", flush=True)
    outs = []
    for order, heavy, bt in [(None, -1, 256), (np.arange(40)[::-1].copy(), 7, 1024)]:
        ctx, g, f = _engine(hg, roots)
        cfg = sbc.BatchConfig(window=12, block_threads=bt, heavy_roots=heavy, root_order=order,
                              recompute_bc=True)
        f.insert_batched(g, ctx, u, v, cfg)
        outs.append((f.trees(), f.bc()))
    (t1, b1), (t2, b2) = outs
    for a, b in zip(t1, t2):
        assert np.array_equal(a, b)
    assert np.array_equal(b1, b2)  # deterministic BC: bit-identical across schedules


def test_errors_are_reported():
    hg = sbc.read_dimacs(karate_path())
    ctx, g, f = _engine(hg, np.arange(hg.nv))
    with pytest.raises(sbc.SbcError) as e:
        f.insert_edge(g, ctx, 1, 2, sbc.FINE)  # already an edge
    assert e.value.status == sbc.SBC_ERR_EXISTS
    with pytest.raises(sbc.SbcError) as e:
        f.insert_batched(g, ctx, [3, 5], [3, 6], sbc.BatchConfig(window=2))
    assert e.value.status == sbc.SBC_ERR_EXISTS
    with pytest.raises(sbc.SbcError):
        f.insert_edge(g, ctx, 1, 10_000, sbc.FINE)


def test_batch_validation_reports_the_first_offender():
    """A batched call is checked whole before anything changes: the first
    insertion in batch order that is out of range, a self loop, a repeat of
    an earlier one (either direction) or an existing edge fails it, and a
    refused batch leaves graph and trees as they were."""
    hg = sbc.read_dimacs(karate_path())
    ctx, g, f = _engine(hg, np.arange(hg.nv))
    before = f.trees()
    deg = [g.degree(x) for x in range(hg.nv)]
    cases = [
        ([(3, 6), (6, 3)], sbc.SBC_ERR_EXISTS, 1),            # repeat, reversed
        ([(3, 6), (1, 2)], sbc.SBC_ERR_EXISTS, 1),            # existing edge
        ([(3, 6), (5, 10_000)], sbc.SBC_ERR_ARG, 1),          # out of range
        ([(3, 6), (1, 2), (5, 10_000)], sbc.SBC_ERR_EXISTS, 1),
        ([(3, 6), (5, 10_000), (1, 2)], sbc.SBC_ERR_ARG, 1),
        ([(3, 6), (4, 7), (7, 4), (1, 2)], sbc.SBC_ERR_EXISTS, 2),
        ([(3, 6), (4, 7), (1, 2), (7, 4)], sbc.SBC_ERR_EXISTS, 2),
        ([(3, 6), (9, 9), (4, 7), (7, 4)], sbc.SBC_ERR_EXISTS, 1),  # self loop first
        ([(3, 6), (4, 7), (7, 4), (-1, 2)], sbc.SBC_ERR_EXISTS, 2),
    ]
    for pairs, status, pos in cases:
        u = [a for a, _ in pairs]
        v = [b for _, b in pairs]
        with pytest.raises(sbc.SbcError) as e:
            f.insert_batched(g, ctx, u, v, sbc.BatchConfig(window=2, block_threads=1024))
        assert e.value.status == status, pairs
        assert f"insertion {pos} " in str(e.value), (pairs, str(e.value))
    assert [g.degree(x) for x in range(hg.nv)] == deg
    for p, q in zip(before, f.trees()):
        assert np.array_equal(p, q)
    # the batch goes through once it is clean
    f.insert_batched(g, ctx, [3, 4], [6, 7], sbc.BatchConfig(window=2, block_threads=1024))
    assert g.degree(3) == deg[3] + 1 and g.degree(7) == deg[7] + 1


@pytest.mark.parametrize("name", ["karate", "dolphins", "jazz"])
def test_gpu_matches_reference_golden(name):
    """Device results against the reference's own outputs (fixtures written
    by the reference compiled from its sources): level/sigma bit-exact,
    delta/BC within tolerance (the device sums delta in pull order), and the
    device's per-item visit/scan counts equal the reference's counters."""
    import os
    from sbc_testutil import GOLDEN
    g = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    hg = sbc.read_dimacs(os.path.join(GOLDEN, "graphs", f"{name}.graph"))
    print(f"[golden] {name}", flush=True)
    ctx, gr, f = _engine(hg, np.arange(hg.nv))
    assert_bc_close(f.bc(), g["coarse_bc"][0])
    for i, (a, b) in enumerate(zip(g["u"], g["v"])):
        f.insert_edge(gr, ctx, int(a), int(b), sbc.FINE)
        rec = ctx.last_records(f.k)
        c = g["fine_counters"][i]
        assert rec["processed"].sum() == c[0] and rec["scanned"].sum() == c[1]
        assert (rec["kase"] == 1).sum() == c[2] and (rec["kase"] == 2).sum() == c[3]
    lv, sg, dl = f.trees()
    assert np.array_equal(lv, g["fine_level"])
    assert np.array_equal(sg, g["fine_sigma"])
    assert_bc_close(dl, g["fine_delta"], what="delta")
    assert_bc_close(f.bc(), g["fine_bc"][-1])


def _hub_stream(hg, n, seed):
    """Random insertions plus some at the highest-degree vertices, so the
    device's hub lookup path (tables over the initial lists, appended slots
    checked one by one) is exercised."""
    u, v = hg.insertions(n, seed)
    deg = hg.degree()
    hubs = np.argsort(deg)[::-1][:4]
    rng = np.random.default_rng(seed)
    lcc = np.where(hg.largest_component())[0]
    extra_u, extra_v = [], []
    seen = set(zip(u.tolist(), v.tolist())) | set(zip(v.tolist(), u.tolist()))
    for h in hubs:
        nb = set(hg.col[hg.rowptr[h]:hg.rowptr[h + 1]].tolist())
        for _ in range(3):
            x = int(rng.choice(lcc))
            if x != h and x not in nb and (h, x) not in seen:
                extra_u.append(int(h)); extra_v.append(x); seen.add((h, x)); seen.add((x, int(h)))
    uu = np.concatenate([u[: n // 2], extra_u, u[n // 2:]]).astype(np.int64)
    vv = np.concatenate([v[: n // 2], extra_v, v[n // 2:]]).astype(np.int64)
    return uu, vv


@pytest.mark.parametrize("batched,run_cap,hub_batch", [(False, None, None), (True, None, None), (True, "3", None),
                                                       (True, None, "3"), (False, None, "2")])
def test_hub_lookup_path_rmat14(batched, run_cap, hub_batch, monkeypatch):
    # run_cap "3": the flat hub scan keeps 3 terms per wavefront, so nearly
    # every hub level takes the overflow path (whole-team rescans);
    # hub_batch: hubs per pass, so levels with more hubs run several passes
    # through the global overflow queue
    if run_cap:
        monkeypatch.setenv("SBC_HUB_RUN_CAP", run_cap)
    if hub_batch:
        monkeypatch.setenv("SBC_HUB_BATCH", hub_batch)
    hg = sbc.gen_rmat(14, 16 << 14, seed=3, keep_lcc=True)
    assert hg.degree().max() > 1000, "need vertices above the lookup-table threshold"
    roots = hg.pick_roots(32, 5)
    u, v = _hub_stream(hg, 48, 6)
    print(f"[hubs] rmat14 nv={hg.nv} maxdeg={hg.degree().max()} n={len(u)} batched={batched}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    if batched:
        rec = f.insert_batched(g, ctx, u, v, sbc.BatchConfig(window=16, block_threads=1024), records=True)
    for i, (a, b) in enumerate(zip(u, v)):
        if not batched:
            f.insert_edge(g, ctx, int(a), int(b), sbc.FINE)
            rec_g = ctx.last_records(f.k)
        else:
            rec_g = rec[i]
        rec_o = o.insert(int(a), int(b), O_FINE, items=True)
        assert np.array_equal(rec_g["affected"], rec_o["affected"]), f"insertion {i}"
        assert np.array_equal(rec_g["scanned"], rec_o["scanned"]), f"insertion {i}"
    _compare_trees(f, o, "rmat14 hubs")
    assert_bc_close(f.bc(), o.bc())


@pytest.mark.parametrize("window", [0, 4])
def test_c_driver_matches_oracle(window, tmp_path):
    """The C driver (a reference-style caller of include/streamingbc_hip.h)
    gives the oracle's BC vector and trees on the same graph, roots and
    stream.  With the reference's calls (window 0) it reads BC straight from
    forest->totalBC after the last insertEdgeStreamingBC, as the reference's
    callers do; the batched form reads it through streamingBCGetBC."""
    import json
    import os
    import subprocess
    from sbc_testutil import GOLDEN, ROOT
    exe = os.path.join(ROOT, "streamingbc_amd", "lib", "sbc_stream")
    path = os.path.join(GOLDEN, "graphs", "jazz.graph")
    dump = str(tmp_path / "bc_tree.bin")
    ri = 17
    out = subprocess.run([exe, "-g", path, "-e", "24", "-s", "3", "-w", str(window), "-o", dump, "-t", str(ri)],
                         check=True, capture_output=True, text=True, timeout=120)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    hg = sbc.read_dimacs(path)
    nv = hg.nv
    raw = np.fromfile(dump, dtype=np.uint8)
    assert raw.size == 32 * nv
    bc = raw[:8 * nv].view(np.float64)
    lv = raw[8 * nv:16 * nv].view(np.int64)
    sg = raw[16 * nv:24 * nv].view(np.int64)
    dl = raw[24 * nv:].view(np.float64)
    roots = np.arange(nv)
    u, v = hg.insertions(24, 13)  # the driver draws its stream with seed + 10
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    for a, b in zip(u, v):
        o.insert(int(a), int(b), O_FINE)
    assert res["insertions"] == 24
    assert_bc_close(bc, o.bc(), what="BC vector read by the C driver")
    lo, so, do = o.tree(ri)
    assert np.array_equal(lv, lo) and np.array_equal(sg, so)
    assert np.array_equal(dl, do)


def test_c_driver_deletes_back_to_the_static_bc():
    """The C driver inserts 24 edges and deletes them again (newest first)
    through deleteEdgeStreamingBC: the case counts equal the oracle's and BC
    returns to the static pass's within rounding."""
    import json
    import os
    import subprocess
    from sbc_testutil import GOLDEN, ROOT
    exe = os.path.join(ROOT, "streamingbc_amd", "lib", "sbc_stream")
    path = os.path.join(GOLDEN, "graphs", "jazz.graph")
    out = subprocess.run([exe, "-g", path, "-e", "24", "-s", "3", "-d", "24"], check=True,
                         capture_output=True, text=True, timeout=120)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    hg = sbc.read_dimacs(path)
    roots = np.arange(hg.nv)
    u, v = hg.insertions(24, 13)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    for a, b in zip(u, v):
        o.insert(int(a), int(b), O_FINE)
    tot = np.zeros(3, dtype=np.int64)
    for a, b in zip(u[::-1], v[::-1]):
        tot += _odel(o, int(a), int(b))
    assert res["deletions"] == 24
    assert (res["same_level"], res["adjacent"], res["movement"]) == tuple(int(t) for t in tot)
    assert abs(res["bc_sum"] - res["bc_sum_static"]) <= 1e-9 * abs(res["bc_sum_static"])
    assert abs(res["bc_max"] - o.bc().max()) <= 1e-9 * abs(o.bc().max())


@pytest.mark.skip(reason="kept as a record: the shared-levels path (idle workgroups taking chunks of another root's "
                         "large BFS levels) faulted the GPU twice with its cause unfound; its code is no longer in "
                         "the library and the SBC_ASSIST_* switches it set do nothing, so this test exercises only "
                         "the default path")
@pytest.mark.parametrize("knobs", [("16", "8", "32"), ("96", "64", "256")])
def test_shared_levels_match_oracle(knobs, monkeypatch):
    """Levels above a small size are split into chunks taken by idle
    workgroups (the owner and helpers append to shared list counters);
    trees, records and BC stay exactly the oracle's."""
    mn, dc, ac = knobs
    monkeypatch.setenv("SBC_ASSIST_MIN", mn)
    monkeypatch.setenv("SBC_ASSIST_DCHUNK", dc)
    monkeypatch.setenv("SBC_ASSIST_ACHUNK", ac)
    hg = sbc.gen_rmat(14, 16 << 14, seed=3, keep_lcc=True)
    roots = hg.pick_roots(24, 5)
    u, v = _hub_stream(hg, 48, 6)
    print(f"[shared] rmat14 knobs={knobs} k={len(roots)} n={len(u)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    half = len(u) // 2
    rec = f.insert_batched(g, ctx, u[:half], v[:half], sbc.BatchConfig(window=8, block_threads=1024), records=True)
    for i in range(len(u)):
        if i >= half:
            f.insert_edge(g, ctx, int(u[i]), int(v[i]), sbc.FINE)
            rec_g = ctx.last_records(f.k)
        else:
            rec_g = rec[i]
        rec_o = o.insert(int(u[i]), int(v[i]), O_FINE, items=True)
        for key in ("affected", "scanned", "processed", "moved"):
            assert np.array_equal(rec_g[key], rec_o[key]), f"insertion {i} {key}"
    _compare_trees(f, o, "shared levels")
    assert_bc_close(f.bc(), o.bc())


@pytest.mark.parametrize("recompute", [False, True])
def test_parked_items_match_oracle(recompute):
    """Items whose descent frontier reaches park_min vertices are parked by
    their team and finished by the whole-GPU level kernels, then the root
    resumes; with a tiny threshold nearly every nonzero item takes that path.
    Trees, records and BC stay exactly the oracle's."""
    hg = sbc.gen_rmat(12, 16 << 12, seed=9, keep_lcc=True)
    roots = hg.pick_roots(48, 4)
    u, v = hg.insertions(64, 21)
    print(f"[parked] rmat12 nv={hg.nv} k={len(roots)} n={len(u)} recompute={recompute}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    cfg = sbc.BatchConfig(window=16, block_threads=1024, park_min=4, recompute_bc=recompute)
    rec = f.insert_batched(g, ctx, u, v, cfg, records=True)
    assert ctx.last_parked() > 0
    for i in range(len(u)):
        rec_o = o.insert(int(u[i]), int(v[i]), O_FINE, items=True)
        for key in ("kase", "affected", "scanned", "processed", "moved"):
            assert np.array_equal(rec[i][key], rec_o[key]), f"insertion {i} {key}"
    _compare_trees(f, o, "parked items")
    if recompute:
        f.recompute_bc(ctx)
    assert_bc_close(f.bc(), o.bc())
    # and never parking gives the same trees
    ctx2, g2, f2 = _engine(hg, roots)
    f2.insert_batched(g2, ctx2, u, v, sbc.BatchConfig(window=16, block_threads=1024, park_min=-1))
    assert ctx2.last_parked() == 0
    for a, b in zip(f.trees(), f2.trees()):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("window", [1, 16])
def test_items_handed_over_in_the_ascent_match_oracle(window, monkeypatch):
    """An ascent level whose hubs leave more entries to scan than the
    threshold than lookups can serve is handed over to the whole-GPU level
    kernels after the team's light-vertex pass (they redo the level, counting
    only its hubs).  With the frontier threshold out of reach and a tiny scan
    threshold every such item takes that path; trees, records and BC stay
    exactly the oracle's."""
    monkeypatch.setenv("SBC_PARK_SCAN", "600")
    hg = sbc.gen_rmat(12, 16 << 12, seed=9, keep_lcc=True)
    roots = hg.pick_roots(48, 4)
    u, v = hg.insertions(64, 21)
    print(f"[ascent handover] rmat12 nv={hg.nv} k={len(roots)} n={len(u)} window={window}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    cfg = sbc.BatchConfig(window=window, block_threads=1024, park_min=1 << 20, recompute_bc=True)
    rec = f.insert_batched(g, ctx, u, v, cfg, records=True)
    parked = ctx.last_parked()
    print(f"[ascent handover] items handed over in the last window: {parked}", flush=True)
    for i in range(len(u)):
        rec_o = o.insert(int(u[i]), int(v[i]), O_FINE, items=True)
        for key in ("kase", "affected", "scanned", "processed", "moved"):
            assert np.array_equal(rec[i][key], rec_o[key]), f"insertion {i} {key}"
    _compare_trees(f, o, "ascent handover")
    f.recompute_bc(ctx)
    assert_bc_close(f.bc(), o.bc())


@pytest.mark.parametrize("knob,value", [("SBC_PARK_MIN", "4"), ("SBC_PARK_SCAN", "600")])
def test_plain_items_handed_over_match_oracle(knob, value, monkeypatch):
    """The plain FINE form hands large items to the whole-GPU level kernels
    with the batched form's thresholds; with a tiny frontier threshold nearly
    every nonzero item takes that path (descent hand-overs), with a tiny scan
    threshold the items whose hub levels scan more than it (ascent
    hand-overs).  Records, trees and BC stay the oracle's."""
    monkeypatch.setenv(knob, value)
    hg = sbc.gen_rmat(12, 16 << 12, seed=9, keep_lcc=True)
    roots = hg.pick_roots(48, 4)
    u, v = hg.insertions(64, 21)
    print(f"[plain handover] rmat12 nv={hg.nv} k={len(roots)} n={len(u)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    parked = 0
    for i, (a, b) in enumerate(zip(u, v)):
        f.insert_edge(g, ctx, int(a), int(b), sbc.FINE)
        parked += ctx.last_parked()
        rec_g = ctx.last_records(f.k)
        rec_o = o.insert(int(a), int(b), O_FINE, items=True)
        for key in ("kase", "affected", "scanned", "processed", "moved"):
            assert np.array_equal(rec_g[key], rec_o[key]), f"insertion {i} {key}"
    print(f"[plain handover] {knob}={value}: {parked} items handed over", flush=True)
    if knob == "SBC_PARK_MIN":
        assert parked > 0, "no plain item was handed over"
    _compare_trees(f, o, "plain hand-over")
    assert_bc_close(f.bc(), o.bc())


@pytest.mark.gpu
def test_work_totals_equal_record_sums(monkeypatch):
    """The per-root work totals kept on the device equal the sums of the item
    records (including items finished by the wide kernels)."""
    hg = sbc.gen_rmat(12, 16 << 12, seed=9, keep_lcc=True)
    roots = hg.pick_roots(48, 4)
    u, v = hg.insertions(64, 21)
    ctx, g, f = _engine(hg, roots)
    ctx.reset_stats(False)
    rec = f.insert_batched(g, ctx, u, v, sbc.BatchConfig(window=16, block_threads=1024, park_min=16), records=True)
    tot = ctx.work_totals(len(roots))
    r = rec.reshape(-1)
    print(f"[totals] parked {ctx.last_parked()} nonzero {int(tot['nonzero'].sum())}", flush=True)
    for key in ("processed", "scanned", "affected", "moved"):
        want = np.bincount(r["root_index"], weights=r[key].astype(np.float64), minlength=len(roots))
        assert np.array_equal(tot[key], want.astype(np.int64)), key
    nz = np.bincount(r["root_index"], weights=(r["kase"] > 0).astype(np.float64), minlength=len(roots))
    assert np.array_equal(tot["nonzero"], nz.astype(np.int64))
    mx = np.zeros(len(roots), dtype=np.int64)
    np.maximum.at(mx, r["root_index"], r["affected"])
    assert np.array_equal(tot["max_affected"], mx)
    # a reset starts them again
    ctx.reset_stats(False)
    assert ctx.work_totals(len(roots))["scanned"].sum() == 0


def test_hub_lookup_level3_leaf_roots():
    """Roots that are leaves hanging off a low-degree vertex see the hubs on
    level 3; their ascent-only hubs there are resolved by table lookups with
    the root's two-hop set as parent candidates.  Exact against the oracle."""
    hg = sbc.gen_rmat(14, 16 << 14, seed=3, keep_lcc=True)
    deg = hg.degree()
    leaves = []
    for x in np.flatnonzero(deg == 1):
        nb = hg.col[hg.rowptr[x]]
        if 2 <= deg[nb] <= 24:
            leaves.append(x)
        if len(leaves) == 24:
            break
    assert len(leaves) >= 8
    roots = np.array(leaves, dtype=np.int64)
    u, v = _hub_stream(hg, 48, 8)
    print(f"[level3] rmat14 k={len(roots)} n={len(u)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    lv, _, _ = f.tree(0)
    assert (deg[lv == 3] > 512).any(), "expected hubs on level 3"
    rec = f.insert_batched(g, ctx, u, v, sbc.BatchConfig(window=16, block_threads=1024), records=True)
    for i in range(len(u)):
        rec_o = o.insert(int(u[i]), int(v[i]), O_FINE, items=True)
        for key in ("affected", "scanned", "processed", "moved"):
            assert np.array_equal(rec[i][key], rec_o[key]), f"insertion {i} {key}"
    _compare_trees(f, o, "level-3 lookups")
    assert_bc_close(f.bc(), o.bc())


# ------------------------------------------------------------ edge deletion
# The device deletion brings each movement-case root to the tree the static
# pass gives on the new graph (updated in place, or rebuilt), so after any mix
# of insertions and deletions the per-root trees equal the oracle's bit for
# bit (same adjacency order on both sides: removal keeps the order, insertion
# appends); BC agrees within 1e-9 with the oracle's incremental BC and the
# reference's fixtures.

import hashlib

from sbc_testutil import GOLDEN


def _apply_ops(f, g, ctx, o, ops):
    counts = []
    for op, a, b in ops:
        a, b = int(a), int(b)
        if op == 1:
            f.insert_edge(g, ctx, a, b, sbc.FINE)
            o.insert(a, b, O_FINE)
            counts.append((-1, -1, -1))
        else:
            c = f.delete_edge(g, ctx, a, b)
            co = _odel(o, a, b)
            got = (c["sameLevel"], c["adjacent"], c["movement"])
            assert got == tuple(int(x) for x in co), f"case counts of deleting ({a},{b}): {got} vs {tuple(co)}"
            counts.append(got)
    return np.array(counts, dtype=np.int64)


# movement-case roots: updated in place (the default below 4M adjacency
# slots; "incremental_wide": the delta pass over the whole GPU, "..._giant":
# with every vertex of 8 or more neighbours summed by a workgroup), or, with
# SBC_DELETE_MOVE=rebuild, rebuilt by the static pass: SBC_WIDE_REBUILD_MAX =
# 1000 rebuilds them one root at a time over the whole GPU; 0 with
# SBC_WIDE_GRAPH_MIN = 0 all together over the whole GPU; 0 with the default
# graph-size threshold (above these graphs) one workgroup per root
REBUILD_MODES = {"incremental": (None, None), "incremental_wide": (None, None),
                 "incremental_wide_giant": (None, None), "single": ("1000", None),
                 "multi": ("0", "0"), "workgroup": ("0", None)}


def _rebuild_mode(monkeypatch, mode):
    wide, gmin = REBUILD_MODES[mode]
    if wide is None:
        for k in ("SBC_WIDE_REBUILD_MAX", "SBC_WIDE_GRAPH_MIN"):
            monkeypatch.delenv(k, raising=False)
        # one workgroup per root throughout, or the delta pass over the GPU
        monkeypatch.setenv("SBC_DELETE_MOVE", "in_place_wide" if mode.startswith("incremental_wide") else "in_place")
        if mode == "incremental_wide_giant":
            monkeypatch.setenv("SBC_WIDE_GIANT", "8")
        else:
            monkeypatch.delenv("SBC_WIDE_GIANT", raising=False)
        return
    monkeypatch.setenv("SBC_DELETE_MOVE", "rebuild")
    monkeypatch.setenv("SBC_WIDE_REBUILD_MAX", wide)
    if gmin is not None:
        monkeypatch.setenv("SBC_WIDE_GRAPH_MIN", gmin)
    else:
        monkeypatch.delenv("SBC_WIDE_GRAPH_MIN", raising=False)


@pytest.mark.parametrize("wide", sorted(REBUILD_MODES))
@pytest.mark.parametrize("name", ["karate", "dolphins"])
def test_delete_matches_oracle_and_reference(name, wide, monkeypatch):
    _rebuild_mode(monkeypatch, wide)
    fx = np.load(os.path.join(GOLDEN, f"refdel_{name}.npz"))
    hg = sbc.read_dimacs(os.path.join(GOLDEN, "graphs", f"{name}.graph"))
    roots = np.arange(hg.nv)
    print(f"[delete] {name} ops={len(fx['ops'])} wide={wide}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    for i, op in enumerate(fx["ops"]):
        c = _apply_ops(f, g, ctx, o, [op])[0]
        if op[0] == 0:
            assert np.array_equal(c, fx["fine_counts"][i]), f"counts of operation {i} vs the reference"
        assert_bc_close(f.bc(), fx["fine_bc"][i + 1], what=f"BC after operation {i} vs the reference")
        assert_bc_close(f.bc(), o.bc(), what=f"BC after operation {i} vs the oracle")
    _compare_trees(f, o, f"{name} after deletions")
    lv, sg, _ = f.trees()
    assert np.array_equal(lv, fx["fine_level"]) and np.array_equal(sg, fx["fine_sigma"])


def _graph_edges(hg, n, seed):
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(hg.nv, dtype=np.int64), np.diff(hg.rowptr))
    out, seen = [], set()
    while len(out) < n:
        e = int(rng.integers(hg.ne))
        a, b = int(src[e]), int(hg.col[e])
        if (min(a, b), max(a, b)) not in seen:
            seen.add((min(a, b), max(a, b)))
            out.append((a, b))
    return out


@pytest.mark.parametrize("mode", ["incremental", "incremental_wide", "incremental_wide_giant"])
@pytest.mark.parametrize("name", ["chunglu_sparse", "rmat12"])
def test_delete_movement_in_place_matches_oracle(name, mode, monkeypatch):
    """Many movement-case roots (subtrees cut off, moved down, hubs above
    them), updated in place: trees bit-identical to the oracle's static
    rebuild after every deletion batch, and the same trees as the rebuild."""
    if name == "chunglu_sparse":
        hg = sbc.gen_chung_lu(4096, 2.3, 3.0, seed=6, keep_lcc=True)
        roots = hg.pick_roots(64, 2)
    else:
        hg = sbc.gen_rmat(12, 8 << 12, seed=4, keep_lcc=True)
        roots = hg.pick_roots(48, 3)
    dels = _graph_edges(hg, 48, 21)
    print(f"[delete] in place {name} {mode} nv={hg.nv} k={len(roots)} deletions={len(dels)}", flush=True)
    _rebuild_mode(monkeypatch, mode)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    moved = 0
    for i in range(0, len(dels), 8):
        c = _apply_ops(f, g, ctx, o, [(0, a, b) for a, b in dels[i:i + 8]])
        moved += int(c[:, 2].sum())
        _compare_trees(f, o, f"{name} after deletion {i + 7}")
        assert_bc_close(f.bc(), o.bc(), what=f"BC after deletion {i + 7}")
    assert moved > 20, f"only {moved} movement-case roots"
    f.recompute_bc(ctx)
    assert_bc_close(f.bc(), o.bc(), what="folded BC")
    # the rebuild path on the same deletions: the same levels and path
    # counts; deltas that no input of theirs touched keep the rounding of
    # earlier updates in place, so they agree within rounding
    _rebuild_mode(monkeypatch, "workgroup")
    ctx2, g2, f2 = _engine(hg, roots)
    for a, b in dels:
        f2.delete_edge(g2, ctx2, a, b)
    (l1, s1, d1), (l2, s2, d2) = f.trees(), f2.trees()
    assert np.array_equal(l1, l2) and np.array_equal(s1, s2)
    assert_bc_close(d1, d2, what="delta in place vs rebuilt")
    assert_bc_close(f.bc(), f2.bc(), what="BC in place vs rebuilt")


@pytest.mark.parametrize("wide", sorted(REBUILD_MODES))
def test_delete_rmat10_matches_reference_hashes(wide, monkeypatch):
    _rebuild_mode(monkeypatch, wide)
    fx = np.load(os.path.join(GOLDEN, "refdel_rmat10.npz"))
    hg = sbc.gen_rmat(10, 8192, seed=1, keep_lcc=True)
    assert hg.nv == int(fx["nv"])
    roots = np.arange(hg.nv)
    print(f"[delete] rmat10 deletions={len(fx['ops'])} wide={wide}", flush=True)
    ctx, g, f = _engine(hg, roots)
    for i, (op, a, b) in enumerate(fx["ops"]):
        c = f.delete_edge(g, ctx, int(a), int(b))
        assert (c["sameLevel"], c["adjacent"], c["movement"]) == tuple(int(x) for x in fx["coarse_counts"][i])
        if i % 8 == 0 or i == len(fx["ops"]) - 1:
            lv, sg, _ = f.trees()
            h = hashlib.sha256(lv.tobytes() + sg.tobytes()).digest()
            assert h == fx["coarse_tree_sha256"][i].tobytes(), f"trees differ after deletion {i}"
        if (i + 1) % 8 == 0:
            assert_bc_close(f.bc(), fx["coarse_bc"][(i + 1) // 8], what=f"BC after deletion {i}")


def _hub_graph():
    hg = sbc.gen_rmat(13, 16 * 8192, seed=2, keep_lcc=True)
    deg = np.diff(hg.rowptr)
    assert deg.max() > 700, "the graph must hold vertices with lookup tables (more than 512 neighbours)"
    return hg, deg


def test_delete_hub_neighbours_then_insert_matches_oracle():
    """Deleting entries of a hub's list rebuilds its lookup table; later
    insertions, plain and batched, look neighbours up in it."""
    hg, deg = _hub_graph()
    roots = hg.pick_roots(24, 4)
    hub = int(np.argmax(deg))
    nb = hg.col[hg.rowptr[hub]:hg.rowptr[hub + 1]]
    print(f"[delete] hub {hub} degree {deg[hub]} nv={hg.nv} k={len(roots)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    picks = [0, 1, 7, len(nb) // 2, len(nb) - 2]
    _apply_ops(f, g, ctx, o, [(0, hub, int(nb[p])) for p in picks])
    _compare_trees(f, o, "after hub deletions")
    u, v = hg.insertions(40, 6)
    far = next(x for x in range(1, hg.nv) if x != hub and x not in set(nb.tolist()) and deg[x] > 0)
    _apply_ops(f, g, ctx, o, [(1, hub, far)] + [(1, int(a), int(b)) for a, b in zip(u[:10], v[:10])])
    _compare_trees(f, o, "plain insertions after hub deletions")
    f.insert_batched(g, ctx, u[10:30], v[10:30], sbc.BatchConfig(window=8, block_threads=1024))
    for a, b in zip(u[10:30], v[10:30]):
        o.insert(int(a), int(b), O_FINE)
    _compare_trees(f, o, "batched insertions after hub deletions")
    # an edge appended to the hub after its table was built, one in the
    # table's part of the list, and three inserted edges elsewhere
    later = [(0, int(a), int(b)) for a, b in zip(u[:3], v[:3])]
    _apply_ops(f, g, ctx, o, [(0, hub, far), (0, hub, int(nb[3]))] + later)
    f.insert_batched(g, ctx, u[30:], v[30:], sbc.BatchConfig(window=5, block_threads=1024))
    for a, b in zip(u[30:], v[30:]):
        o.insert(int(a), int(b), O_FINE)
    _compare_trees(f, o, "after the second round")
    assert_bc_close(f.bc(), o.bc())


def test_delete_disconnects_and_legacy_form():
    """A bridge deletion leaves vertices unreachable (LONG_MAX sentinels); the
    legacy two-step form gives the same trees; deleting an absent edge is
    refused before anything changes."""
    a = [(i, i + 1) for i in range(1, 8)]
    hg = sbc.graph_from_edges(9, np.array([x for x, y in a]), np.array([y for x, y in a]))
    roots = np.arange(hg.nv)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    _apply_ops(f, g, ctx, o, [(0, 4, 5)])
    _compare_trees(f, o, "path cut")
    lv, sg, _ = f.tree(1)
    assert np.all(lv[5:9] == LONG_MAX) and np.all(sg[5:9] == LONG_MAX)
    bc = f.bc()
    for k in range(1, 5):
        assert bc[k] == 2 * (k - 1) * (4 - k)
    assert g.remove_edge(2, 3) == 1 and g.remove_edge(3, 2) == 1
    c = f.update_deleted(g, ctx, 2, 3)
    co = _odel(o, 2, 3)
    assert (c["sameLevel"], c["adjacent"], c["movement"]) == tuple(int(x) for x in co)
    _compare_trees(f, o, "legacy deletion")
    before = f.trees()
    with pytest.raises(sbc.SbcError) as ei:
        f.delete_edge(g, ctx, 2, 3)
    assert ei.value.status == sbc.SBC_ERR_ABSENT
    after = f.trees()
    assert all(np.array_equal(x, y) for x, y in zip(before, after))
    assert g.degree(2) == 1 and g.degree(3) == 1


@pytest.mark.parametrize("park_scan", ["64", "2000"])
def test_delete_items_handed_to_wide_kernels_match_oracle(park_scan, monkeypatch):
    """Deletion items whose descent or ascent scans more than SBC_PARK_SCAN
    entries are finished by the wide kernels, which mark the shallower
    endpoint as a parent once the deeper one's level is done."""
    monkeypatch.setenv("SBC_PARK_SCAN", park_scan)
    hg, deg = _hub_graph()
    roots = hg.pick_roots(24, 7)
    print(f"[delete] parked items, park_scan={park_scan} nv={hg.nv}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    rng = np.random.default_rng(11)
    src = np.repeat(np.arange(hg.nv), np.diff(hg.rowptr))
    picks = rng.choice(hg.ne, 12, replace=False)
    ops = []
    for e in picks:
        a, b = int(src[e]), int(hg.col[e])
        if (0, min(a, b), max(a, b)) not in ops:
            ops.append((0, min(a, b), max(a, b)))
    parked = 0
    adjacent = 0
    for op in ops:
        c = _apply_ops(f, g, ctx, o, [op])[0]
        parked += ctx.last_parked()
        adjacent += int(c[1])
    assert adjacent > 0 and parked > 0, f"no deletion item was handed over (adjacent roots {adjacent})"
    _compare_trees(f, o, f"deletions with park_scan {park_scan}")
    assert_bc_close(f.bc(), o.bc())


def _assert_fold_exact(f, ctx, o, roots):
    """streamingBCRecomputeBC folds the trees in root order: equal bit for bit
    to the same fold of the oracle's trees."""
    f.recompute_bc(ctx)
    _, _, dl = o.trees()
    acc = np.zeros(dl.shape[1])
    for ri, r in enumerate(roots):
        d = dl[ri].copy()
        d[int(r)] = 0.0
        acc = acc + d
    assert np.array_equal(f.bc(), 0.0 + acc)


def test_vertex_delete_and_insert_match_oracle():
    """deleteVertexStreamingBC removes every edge of a vertex in list order,
    one deletion update each; insertVertexStreamingBC connects it again (one
    batched call, equal to one insertion per edge); duplicates and self loops
    in the list are skipped."""
    hg = sbc.read_dimacs(karate_path())
    roots = np.arange(hg.nv)
    print("[vertex] karate delete/insert vertex 1", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    nb = [int(x) for x in hg.col[hg.rowptr[1]:hg.rowptr[2]]]
    removed, c = f.delete_vertex(g, ctx, 1)
    assert list(removed) == nb and g.degree(1) == 0
    tot = np.zeros(3, dtype=np.int64)
    for x in nb:
        tot += _odel(o, 1, x)
    assert (c["sameLevel"], c["adjacent"], c["movement"]) == tuple(int(t) for t in tot)
    _compare_trees(f, o, "after deleting vertex 1")
    # 16 consecutive deletions take the most central vertex's BC (and its
    # neighbours' shares) to 0: device and oracle each keep rounding residues
    # of ~1e-12 there, above 1e-9 x the 1e-6 x max|BC| floor
    assert_bc_close(f.bc(), o.bc(), rel=1e-8)
    _assert_fold_exact(f, ctx, o, roots)
    again = [34, 2, 3, 3, 1, 12, 33, 2]
    f.insert_vertex(g, ctx, 1, again)
    seen = set()
    for x in again:
        if x != 1 and x not in seen:
            seen.add(x)
            o.insert(1, x, O_FINE)
    assert g.degree(1) == len(seen)
    _compare_trees(f, o, "after inserting vertex 1 again")
    _assert_fold_exact(f, ctx, o, roots)
    removed, _ = f.delete_vertex(g, ctx, 1)
    assert list(removed) == [34, 2, 3, 12, 33]


def test_checkpoint_resume_is_bit_identical(tmp_path):
    """A checkpoint taken after insertions, deletions (one rebuilding a hub's
    lookup table) and a batched window restores trees and BC exactly; the
    resumed forest stays bit-identical to the original under further updates,
    and both equal the oracle, which runs the same operations from the start
    without any checkpoint."""
    hg, deg = _hub_graph()
    roots = hg.pick_roots(24, 9)
    hub = int(np.argmax(deg))
    nb = hg.col[hg.rowptr[hub]:hg.rowptr[hub + 1]]
    print(f"[checkpoint] nv={hg.nv} k={len(roots)}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    u, v = hg.insertions(48, 12)
    for a, b in zip(u[:4], v[:4]):
        f.insert_edge(g, ctx, int(a), int(b), sbc.FINE)
        o.insert(int(a), int(b), O_FINE)
    for a, b in ((hub, int(nb[2])), (int(u[0]), int(v[0]))):
        f.delete_edge(g, ctx, a, b)
        _odel(o, a, b)
    f.insert_batched(g, ctx, u[4:20], v[4:20], sbc.BatchConfig(window=8, block_threads=1024))
    for a, b in zip(u[4:20], v[4:20]):
        o.insert(int(a), int(b), O_FINE)
    path = str(tmp_path / "state.ckpt")
    sbc.checkpoint_save(f, g, path)
    f2, g2 = sbc.checkpoint_load(path)
    ctx2 = sbc.Context(hg.nv)
    for x, y in zip(f.trees(), f2.trees()):
        assert np.array_equal(x, y)
    assert np.array_equal(f.bc(), f2.bc())
    assert all(g.degree(x) == g2.degree(x) for x in range(hg.nv))
    for ff, gg, cc in ((f, g, ctx), (f2, g2, ctx2)):
        ff.insert_batched(gg, cc, u[20:40], v[20:40], sbc.BatchConfig(window=8, block_threads=1024))
        ff.delete_edge(gg, cc, hub, int(nb[5]))
        ff.insert_edge(gg, cc, int(u[40]), int(v[40]), sbc.FINE)
        ff.recompute_bc(cc)
    for x, y in zip(f.trees(), f2.trees()):
        assert np.array_equal(x, y)
    assert np.array_equal(f.bc(), f2.bc())
    for a, b in zip(u[20:40], v[20:40]):
        o.insert(int(a), int(b), O_FINE)
    _odel(o, hub, int(nb[5]))
    o.insert(int(u[40]), int(v[40]), O_FINE)
    _compare_trees(f2, o, "resumed forest vs the oracle")
    assert_bc_close(f2.bc(), o.bc(), what="resumed BC vs the oracle")
    f2.close()
    g2.close()


def test_batched_relayout_inside_one_window():
    """One window appends more edges to a vertex than its append slack holds:
    the batched call lays the graph out again up front (fresh slack for every
    vertex, the device image uploaded again) and the results stay the
    oracle's; later windows and plain insertions use the new layout."""
    hg = sbc.gen_chung_lu(4096, 2.3, 8.0, seed=7, keep_lcc=True)
    roots = hg.pick_roots(32, 3)
    deg = hg.degree()
    x = int(np.flatnonzero(deg == 3)[0])
    nbx = set(hg.col[hg.rowptr[x]:hg.rowptr[x + 1]].tolist())
    lcc = np.flatnonzero(hg.largest_component())
    far = [int(y) for y in lcc if y != x and int(y) not in nbx][:12]
    room = 4 + 3 // 16
    assert len(far) > room
    u = np.array([x] * len(far), dtype=np.int64)
    v = np.array(far, dtype=np.int64)
    print(f"[relayout] vertex {x} degree 3, slack {room}, {len(far)} insertions in one window", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    rec = f.insert_batched(g, ctx, u, v, sbc.BatchConfig(window=16, block_threads=1024), records=True)
    assert g.degree(x) == 3 + len(far)
    for i, (a, b) in enumerate(zip(u, v)):
        ro = o.insert(int(a), int(b), O_FINE, items=True)
        for key in ("kase", "affected", "scanned", "processed", "moved"):
            assert np.array_equal(rec[i][key], ro[key]), f"insertion {i} {key}"
    _compare_trees(f, o, "after the relayout window")
    u2, v2 = hg.insertions(40, 9)
    keep = [(a, b) for a, b in zip(u2, v2) if a != x and b != x][:24]
    f.insert_batched(g, ctx, [a for a, _ in keep[:16]], [b for _, b in keep[:16]],
                     sbc.BatchConfig(window=8, block_threads=1024))
    for a, b in keep[:16]:
        o.insert(int(a), int(b), O_FINE)
    for a, b in keep[16:]:
        f.insert_edge(g, ctx, int(a), int(b), sbc.FINE)
        o.insert(int(a), int(b), O_FINE)
    _compare_trees(f, o, "after later windows")
    assert_bc_close(f.bc(), o.bc())


@pytest.mark.parametrize("mode", ["single", "incremental", "incremental_wide"])
def test_wide_rebuild_of_a_root_with_a_big_hub(mode, monkeypatch):
    """Movement-case roots when a level holds a vertex with more than 16384
    neighbours: rebuilt over the whole GPU (its terms computed across the GPU
    and summed in order by one wavefront), or updated in place (the hub's
    delta recomputed by one wavefront, or, with the delta pass over the whole
    GPU, by a workgroup over six LDS slabs); bit-identical to the oracle."""
    _rebuild_mode(monkeypatch, mode)
    n = 20100
    edges = [(1, i) for i in range(2, 20002)] + [(i, i + 1) for i in range(2, 300)] + [(20002, 20003), (20003, 5)]
    hg = sbc.graph_from_edges(n, np.array([a for a, b in edges]), np.array([b for a, b in edges]))
    roots = np.array([3, 7, 150, 299, 20003, 1, 9000, 12000], dtype=np.int64)
    print(f"[delete] big hub nv={hg.nv}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    c = _apply_ops(f, g, ctx, o, [(0, 1, 9000), (0, 1, 250), (0, 20003, 5)])
    assert c[:, 2].sum() > 0, "no movement-case root"
    _compare_trees(f, o, "wide rebuild with a big hub")
    assert_bc_close(f.bc(), o.bc())


def test_hub_roots_match_oracle():
    """Roots that are themselves hubs with lookup tables: their own level-0
    update is resolved by table lookups over the processed level-1 vertices."""
    hg, deg = _hub_graph()
    top = np.argsort(-deg)[:6]
    roots = np.concatenate([top, hg.pick_roots(10, 5)]).astype(np.int64)
    print(f"[hub roots] nv={hg.nv} roots={len(roots)} top degree {deg[top[0]]}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    u, v = hg.insertions(40, 21)
    for a, b in zip(u[:12], v[:12]):
        f.insert_edge(g, ctx, int(a), int(b), sbc.FINE)
        o.insert(int(a), int(b), O_FINE)
    _compare_trees(f, o, "plain insertions, hub roots")
    f.insert_batched(g, ctx, u[12:], v[12:], sbc.BatchConfig(window=7, block_threads=1024))
    for a, b in zip(u[12:], v[12:]):
        o.insert(int(a), int(b), O_FINE)
    _compare_trees(f, o, "batched insertions, hub roots")
    assert_bc_close(f.bc(), o.bc())


def _bucket_of(w, log_t):
    """The bucket of neighbour w in a hub table of 2^log_t slots (8 slots a
    bucket; the multiplicative hash the library uses)."""
    return ((int(w) * 0x9E3779B1) & 0xFFFFFFFF) >> (32 - (log_t - 3))


@pytest.mark.parametrize("batched", [False, True])
def test_hub_table_overflow_chains(batched):
    """A hub whose 700 neighbours all hash to 8 of its table's 512 buckets:
    lookups of its processed children walk chains of ~90 full buckets.
    Trees, item records and BC stay the oracle's."""
    rng = np.random.default_rng(21)
    n = 60000
    log_t = 12  # 4 x 701 slots -> 4096
    pool = [w for w in range(3, n + 1) if _bucket_of(w, log_t) < 8]
    leaves = np.array(rng.choice(pool, size=700, replace=False), dtype=np.int64)
    rest = np.array([w for w in range(3, n + 1) if _bucket_of(w, log_t) >= 8], dtype=np.int64)
    lower = rng.choice(rest, size=300, replace=False).astype(np.int64)
    eu, ev = [1], [2]  # root 1 - hub 2
    eu += [2] * len(leaves)
    ev += leaves.tolist()
    for x in lower:  # level-3 vertices under 2-3 leaves each
        for w in rng.choice(leaves, size=int(rng.integers(2, 4)), replace=False):
            eu.append(int(x)); ev.append(int(w))
    hg = sbc.graph_from_edges(n, eu, ev)
    assert hg.degree()[2] == 701
    # the premise: a long overflow chain in the hub's table
    filled = np.bincount([_bucket_of(w, log_t) for w in hg.col[hg.rowptr[2]:hg.rowptr[3]]], minlength=512)
    assert filled[:8].sum() == 700 and filled.max() > 8
    roots = np.array([1, 2, int(leaves[0]), int(lower[0])], dtype=np.int64)
    cand = np.concatenate([leaves, lower])
    seen = set(zip(eu, ev)) | set(zip(ev, eu))
    u, v = [], []
    while len(u) < 64:
        a, b = (int(x) for x in rng.choice(cand, size=2, replace=False))
        if (a, b) not in seen:
            seen.add((a, b)); seen.add((b, a))
            u.append(a); v.append(b)
    print(f"[hub-chains] nv={hg.nv} n={len(u)} batched={batched}", flush=True)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    _compare_trees(f, o, "hub chains static")
    if batched:
        rec = f.insert_batched(g, ctx, u, v, sbc.BatchConfig(window=16, block_threads=1024), records=True)
    for i, (a, b) in enumerate(zip(u, v)):
        if not batched:
            f.insert_edge(g, ctx, a, b, sbc.FINE)
            rec_g = ctx.last_records(f.k)
        else:
            rec_g = rec[i]
        rec_o = o.insert(a, b, O_FINE, items=True)
        assert np.array_equal(rec_g["affected"], rec_o["affected"]), f"insertion {i}"
        assert np.array_equal(rec_g["scanned"], rec_o["scanned"]), f"insertion {i}"
    _compare_trees(f, o, "hub chains")
    assert_bc_close(f.bc(), o.bc())


@pytest.mark.parametrize("giant", [None, "64"])
def test_handed_over_giant_hub_matches_oracle(giant, monkeypatch):
    """A hub of 12,000 neighbours inside items finished by the whole-GPU level
    kernels.  In an ascent level it takes a workgroup whose wavefronts place
    its terms in LDS in list order, slab after slab (3840 neighbours a slab,
    two buffers), while one lane adds them; in a descent frontier its list is
    split into work groups that any wavefront takes.  With SBC_WIDE_GIANT=64
    every vertex past 64 neighbours takes the workgroup path.  Trees, item
    records and BC stay the oracle's."""
    if giant:
        monkeypatch.setenv("SBC_WIDE_GIANT", giant)
    rng = np.random.default_rng(31)
    n = 20000
    hub_nb = rng.choice(np.arange(2, n + 1), size=12000, replace=False)
    eu = rng.integers(2, n + 1, size=3 * n).tolist() + [1] * len(hub_nb)
    ev = rng.integers(2, n + 1, size=3 * n).tolist() + hub_nb.tolist()
    hg = sbc.graph_from_edges(n, eu, ev)
    assert hg.degree()[1] == 12000
    roots = hg.pick_roots(32, 5)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    lv0 = o.trees()[0]
    # roots two levels above the hub: an edge from such a root to a hub
    # neighbour three or more levels down puts the hub into a descent frontier
    pairs = []
    for ri in np.flatnonzero(lv0[:, 1] == 2)[:6]:
        far = hub_nb[lv0[ri, hub_nb] >= 3]
        if far.size:
            pairs.append((int(roots[ri]), int(far[0])))
    assert pairs, "no root two levels above the hub"
    seen = {(min(a, b), max(a, b)) for a, b in pairs}
    ru, rv = hg.insertions(64, 6)
    for a, b in zip(ru.tolist(), rv.tolist()):
        if (min(a, b), max(a, b)) not in seen:
            seen.add((min(a, b), max(a, b)))
            pairs.append((a, b))
    u = np.array([p[0] for p in pairs[:48]], dtype=np.int64)
    v = np.array([p[1] for p in pairs[:48]], dtype=np.int64)
    print(f"[giant hub] nv={hg.nv} k={len(roots)} n={len(u)} giant={giant}", flush=True)
    ctx, g, f = _engine(hg, roots)
    cfg = sbc.BatchConfig(window=16, block_threads=1024, park_min=4, recompute_bc=True)
    rec = f.insert_batched(g, ctx, u, v, cfg, records=True)
    assert ctx.last_parked() > 0
    for i in range(len(u)):
        rec_o = o.insert(int(u[i]), int(v[i]), O_FINE, items=True)
        for key in ("kase", "affected", "scanned", "processed", "moved"):
            assert np.array_equal(rec[i][key], rec_o[key]), f"insertion {i} {key}"
    _compare_trees(f, o, "giant hub")
    f.recompute_bc(ctx)
    assert_bc_close(f.bc(), o.bc())


def test_delete_giant_hub_delta_pass_matches_oracle(monkeypatch):
    """Deletions around a 12,000-neighbour hub with the movement case's delta
    pass over the whole GPU: the hub's sum (a fresh sum over its children in
    adjacency order) crosses four LDS slabs of one workgroup, and its parents
    are marked by every wavefront.  Case counts, trees and BC stay the
    oracle's."""
    monkeypatch.setenv("SBC_DELETE_MOVE", "in_place_wide")
    monkeypatch.delenv("SBC_WIDE_GIANT", raising=False)
    rng = np.random.default_rng(41)
    n = 20000
    hub_nb = rng.choice(np.arange(2, n + 1), size=12000, replace=False)
    eu = rng.integers(2, n + 1, size=3 * n).tolist() + [1] * len(hub_nb)
    ev = rng.integers(2, n + 1, size=3 * n).tolist() + hub_nb.tolist()
    hg = sbc.graph_from_edges(n, eu, ev)
    assert hg.degree()[1] == 12000
    roots = hg.pick_roots(32, 7)
    ctx, g, f = _engine(hg, roots)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    # edges of the hub and of its neighbours: deleting them moves vertices
    # below the hub and changes its children
    ops = [(0, 1, int(w)) for w in hub_nb[:6]]
    for w in hub_nb[6:12]:
        nb = hg.col[hg.rowptr[w]:hg.rowptr[w + 1]]
        nb = nb[nb != 1]
        if nb.size:
            ops.append((0, int(w), int(nb[0])))
    print(f"[delete giant] nv={hg.nv} k={len(roots)} ops={len(ops)}", flush=True)
    c = _apply_ops(f, g, ctx, o, ops)
    assert c[:, 2].sum() > 0, "no movement-case root"
    _compare_trees(f, o, "deletions around a giant hub")
    assert_bc_close(f.bc(), o.bc())
