"""CPU tests that pin the oracle (no GPU needed).

Pins, per oracle function (see oracle/sbc_oracle.h):
  static push  -> golden vectors of the reference (bit-exact), networkx x2,
                  closed forms (path, star, cycle, complete), brute force
  static pull  -> push (dist/sigma equal, delta/BC within 1e-12), closed forms
  insert COARSE / FINE -> golden vectors of the reference's COARSE / FINE
                  update at NT=1 (bit-exact, incl. BC), static recompute of the
                  post-insertion graph, the reference's traversal counters
"""
import hashlib
import itertools
import os

import numpy as np
import pytest

import streamingbc_amd as sbc
from oracle_py import COARSE, FINE, MOVE_IN_PLACE, MOVE_REBUILD, PULL, PUSH, Oracle
from sbc_testutil import GOLDEN, LONG_MAX, assert_bc_close

GRAPH_DIR = os.path.join(GOLDEN, "graphs")


def load_graph(name):
    return sbc.read_dimacs(os.path.join(GRAPH_DIR, f"{name}.graph"))


def from_edges(n, edges):
    eu = np.array([a for a, b in edges], dtype=np.int64)
    ev = np.array([b for a, b in edges], dtype=np.int64)
    return sbc.graph_from_edges(n, eu, ev)


# ------------------------------------------------------------ golden vectors

@pytest.mark.parametrize("name", ["karate", "dolphins", "jazz"])
@pytest.mark.parametrize("mode,key", [(COARSE, "coarse"), (FINE, "fine")])
def test_oracle_matches_reference_golden(name, mode, key):
    g = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    hg = load_graph(name)
    roots = np.arange(hg.nv)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PUSH)
    assert np.array_equal(o.bc(), g[f"{key}_bc"][0]), "static BC differs from the reference"
    for i, (a, b) in enumerate(zip(g["u"], g["v"])):
        items = o.insert(int(a), int(b), mode, items=True)
        assert np.array_equal(o.bc(), g[f"{key}_bc"][i + 1]), f"BC after insertion {i}"
        # the reference's own traversal counters (vertex visits, edges scanned,
        # adjacent-case roots, movement-case roots) for this insertion
        ref_c = g[f"{key}_counters"][i]
        assert items["processed"].sum() == ref_c[0]
        assert items["scanned"].sum() == ref_c[1]
        assert (items["kase"] == 1).sum() == ref_c[2]
        assert (items["kase"] == 2).sum() == ref_c[3]
    lv, sg, dl = o.trees()
    assert np.array_equal(lv, g[f"{key}_level"])
    assert np.array_equal(sg, g[f"{key}_sigma"])
    assert np.array_equal(dl, g[f"{key}_delta"])


def test_oracle_matches_reference_golden_rmat10():
    g = np.load(os.path.join(GOLDEN, "ref_rmat10.npz"))
    hg = sbc.gen_rmat(10, 8192, seed=1, keep_lcc=True)
    assert hg.nv == int(g["nv"]) and hg.ne == int(g["ne"]), "generator changed: fixtures are stale"
    u, v = hg.insertions(256, 1)
    assert np.array_equal(u, g["u"]) and np.array_equal(v, g["v"])
    roots = np.arange(hg.nv)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PUSH)
    assert np.array_equal(o.bc(), g["coarse_bc"][0])
    for i, (a, b) in enumerate(zip(u, v)):
        o.insert(int(a), int(b), COARSE)
        if i % 32 == 0 or i == len(u) - 1:
            lv, sg, _ = o.trees()
            h = hashlib.sha256(lv.tobytes() + sg.tobytes()).digest()
            assert h == g["coarse_tree_sha256"][i].tobytes(), f"trees differ after insertion {i}"
        if (i + 1) % 16 == 0:
            assert np.array_equal(o.bc(), g["coarse_bc"][(i + 1) // 16])
    lv, sg, dl = o.trees()
    assert np.array_equal(lv, g["coarse_level"])
    assert np.array_equal(sg, g["coarse_sigma"])
    assert np.array_equal(dl[:64], g["coarse_delta"])


@pytest.mark.parametrize("name", ["karate", "dolphins", "jazz"])
def test_static_matches_networkx(name):
    g = np.load(os.path.join(GOLDEN, f"nx_{name}.npz"))
    hg = load_graph(name)
    for order in (PUSH, PULL):
        o = Oracle(hg.rowptr, hg.col, np.arange(hg.nv))
        o.static(order)
        assert_bc_close(o.bc(), g["bc2"], rel=1e-12)


# ------------------------------------------------------------ closed forms

def _exact_bc(hg, order=PUSH):
    o = Oracle(hg.rowptr, hg.col, np.arange(hg.nv))
    o.static(order)
    return o.bc()


@pytest.mark.parametrize("order", [PUSH, PULL])
def test_closed_forms(order):
    n = 9
    path = from_edges(n, [(i, i + 1) for i in range(1, n)])
    bc = _exact_bc(path, order)
    for k in range(1, n + 1):  # ordered pairs: 2 (k-1)(n-k)
        assert bc[k] == 2 * (k - 1) * (n - k)
    assert bc[0] == 0
    star = from_edges(n, [(1, i) for i in range(2, n + 1)])
    bc = _exact_bc(star, order)
    assert bc[1] == (n - 1) * (n - 2)
    assert np.all(bc[2:] == 0)
    comp = from_edges(6, list(itertools.combinations(range(1, 7), 2)))
    assert np.all(_exact_bc(comp, order) == 0)
    cyc = from_edges(8, [(i, i % 8 + 1) for i in range(1, 9)])
    bc = _exact_bc(cyc, order)
    assert np.allclose(bc[1:], bc[1]) and bc[1] > 0


def _brute_bc(n, edges):
    """BC by enumerating all shortest paths on a tiny graph (ordered pairs)."""
    adj = {x: set() for x in range(1, n + 1)}
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)

    def paths(s, t):
        # all shortest s-t paths by BFS layering
        dist = {s: 0}
        layer = [s]
        while layer and t not in dist:
            nxt = []
            for x in layer:
                for y in adj[x]:
                    if y not in dist:
                        dist[y] = dist[x] + 1
                        nxt.append(y)
            layer = nxt
        if t not in dist:
            return []
        out = []

        def walk(path):
            x = path[-1]
            if x == t:
                out.append(path)
                return
            for y in adj[x]:
                if dist.get(y) == dist[x] + 1 and (y == t or dist[y] < dist[t]):
                    walk(path + [y])
        walk([s])
        return out

    bc = np.zeros(n + 1)
    for s in range(1, n + 1):
        for t in range(1, n + 1):
            if s == t:
                continue
            ps = paths(s, t)
            for p in ps:
                for x in p[1:-1]:
                    bc[x] += 1.0 / len(ps)
    return bc


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_small(seed):
    rng = np.random.default_rng(seed)
    n = 10
    edges = set()
    while len(edges) < 16:
        a, b = rng.integers(1, n + 1, 2)
        if a != b:
            edges.add((min(a, b), max(a, b)))
    edges = sorted(edges)
    hg = from_edges(n, edges)
    want = _brute_bc(n, edges)
    for order in (PUSH, PULL):
        assert_bc_close(_exact_bc(hg, order), want, rel=1e-12)


# ------------------------------------------------------------ invariants

@pytest.mark.parametrize("mode", [COARSE, FINE])
def test_incremental_equals_static_recompute(mode):
    hg = sbc.gen_chung_lu(1500, 2.3, 8.0, seed=2, keep_lcc=True)
    roots = hg.pick_roots(40, 3)
    u, v = hg.insertions(60, 4)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PUSH)
    for a, b in zip(u, v):
        o.insert(int(a), int(b), mode)
    s = Oracle(hg.rowptr, hg.col, roots)
    for a, b in zip(u, v):
        s.add_edge(int(a), int(b))
    s.static(PUSH)
    lv, sg, dl = o.trees()
    ls, ss, ds = s.trees()
    assert np.array_equal(lv, ls)
    assert np.array_equal(sg, ss)
    assert_bc_close(dl, ds, what="delta")
    assert_bc_close(o.bc(), s.bc())


def test_connecting_components_matches_recompute():
    # two components joined by the insertion: one endpoint unreachable from
    # some roots (the reference's LONG_MAX arithmetic path)
    a = [(i, i + 1) for i in range(1, 6)]
    b = [(i, i + 1) for i in range(7, 11)] + [(7, 9)]
    hg = from_edges(11, a + b)
    roots = np.arange(hg.nv)
    for mode in (COARSE, FINE):
        o = Oracle(hg.rowptr, hg.col, roots)
        o.static(PUSH)
        o.insert(3, 8, mode)
        s = Oracle(hg.rowptr, hg.col, roots)
        s.add_edge(3, 8)
        s.static(PUSH)
        lv, sg, dl = o.trees()
        ls, ss, ds = s.trees()
        assert np.array_equal(lv, ls) and np.array_equal(sg, ss)
        assert_bc_close(dl, ds)
        assert_bc_close(o.bc(), s.bc())


def test_fine_and_coarse_agree_and_pull_static_close():
    hg = load_graph("jazz")
    roots = np.arange(hg.nv)
    u, v = hg.insertions(30, 8)
    oc = Oracle(hg.rowptr, hg.col, roots)
    of = Oracle(hg.rowptr, hg.col, roots)
    op = Oracle(hg.rowptr, hg.col, roots)
    oc.static(PUSH)
    of.static(PUSH)
    op.static(PULL)
    lc, sc, dc = oc.trees()
    lp, sp_, dp = op.trees()
    assert np.array_equal(lc, lp) and np.array_equal(sc, sp_)
    assert_bc_close(dp, dc, rel=1e-12, what="pull vs push delta")
    for a, b in zip(u, v):
        oc.insert(int(a), int(b), COARSE)
        of.insert(int(a), int(b), FINE)
    l1, s1, d1 = oc.trees()
    l2, s2, d2 = of.trees()
    assert np.array_equal(l1, l2) and np.array_equal(s1, s2)
    assert_bc_close(d2, d1, rel=1e-10)  # summation order differs (push vs pull)


def test_planted_error_is_caught():
    """The tolerance check rejects a delta perturbed well above rounding."""
    hg = load_graph("karate")
    o = Oracle(hg.rowptr, hg.col, np.arange(hg.nv))
    o.static(PUSH)
    lv, sg, dl = o.trees()
    bad = dl.copy()
    bad[5, 3] *= (1 + 1e-6)
    from sbc_testutil import bc_close
    assert not bc_close(bad, dl)[0]
    lv2 = lv.copy()
    lv2[1, 2] += 1
    assert not np.array_equal(lv2, lv)


def test_sigma_sentinels():
    hg = load_graph("karate")
    o = Oracle(hg.rowptr, hg.col, np.arange(hg.nv))
    o.static(PUSH)
    lv, sg, dl = o.tree(0)  # phantom root: everything else unreachable
    assert lv[0] == 0 and sg[0] == 1
    assert np.all(lv[1:] == LONG_MAX) and np.all(sg[1:] == LONG_MAX) and np.all(dl == 0)
    assert not o.sigma_overflowed()


# ------------------------------------------------------------ edge deletion
# orc_delete is pinned by the reference's deleteEdgeStreamingBC (fixtures
# refdel_*.npz written by the compiled reference: case counts exact, level and
# sigma exact, delta and BC within 1e-9), by a static recompute of the
# post-deletion graph, and by closed forms on paths.

def _replay(o, ops, mode):
    counts = []
    for op, a, b in ops:
        if op == 1:
            o.insert(int(a), int(b), mode)
            counts.append(np.full(3, -1, dtype=np.int64))
        else:
            counts.append(o.delete(int(a), int(b)))
        yield np.array(counts[-1])


@pytest.mark.parametrize("name", ["karate", "dolphins"])
@pytest.mark.parametrize("mode,key", [(COARSE, "coarse"), (FINE, "fine")])
def test_oracle_delete_matches_reference_golden(name, mode, key):
    g = np.load(os.path.join(GOLDEN, f"refdel_{name}.npz"))
    hg = load_graph(name)
    o = Oracle(hg.rowptr, hg.col, np.arange(hg.nv))
    o.static(PUSH)
    assert np.array_equal(o.bc(), g[f"{key}_bc"][0])
    assert (g["ops"][:, 0] == 0).sum() >= 10 and (g["ops"][:, 0] == 1).sum() >= 3
    for i, c in enumerate(_replay(o, g["ops"], mode)):
        assert np.array_equal(c, g[f"{key}_counts"][i]), f"case counts of operation {i}"
        assert_bc_close(o.bc(), g[f"{key}_bc"][i + 1], what=f"BC after operation {i}")
    lv, sg, dl = o.trees()
    assert np.array_equal(lv, g[f"{key}_level"])
    assert np.array_equal(sg, g[f"{key}_sigma"])
    assert_bc_close(dl, g[f"{key}_delta"], what="delta")


def test_oracle_delete_matches_reference_golden_rmat10():
    g = np.load(os.path.join(GOLDEN, "refdel_rmat10.npz"))
    hg = sbc.gen_rmat(10, 8192, seed=1, keep_lcc=True)
    assert hg.nv == int(g["nv"]) and hg.ne == int(g["ne"]), "generator changed: fixtures are stale"
    o = Oracle(hg.rowptr, hg.col, np.arange(hg.nv))
    o.static(PUSH)
    assert np.array_equal(o.bc(), g["coarse_bc"][0])
    kinds = np.zeros(3, dtype=np.int64)
    for i, c in enumerate(_replay(o, g["ops"], COARSE)):
        assert np.array_equal(c, g["coarse_counts"][i]), f"case counts of deletion {i}"
        kinds += c > 0
        if i % 8 == 0 or i == len(g["ops"]) - 1:
            lv, sg, _ = o.trees()
            h = hashlib.sha256(lv.tobytes() + sg.tobytes()).digest()
            assert h == g["coarse_tree_sha256"][i].tobytes(), f"trees differ after deletion {i}"
        if (i + 1) % 8 == 0:
            assert_bc_close(o.bc(), g["coarse_bc"][(i + 1) // 8], what=f"BC after deletion {i}")
    assert np.all(kinds > 0), "the stream must hold same-level, adjacent and movement roots"
    lv, sg, dl = o.trees()
    assert np.array_equal(lv, g["coarse_level"])
    assert np.array_equal(sg, g["coarse_sigma"])
    assert_bc_close(dl[:64], g["coarse_delta"], what="delta")


def test_delete_equals_static_recompute():
    """After deletions alone, every tree equals a fresh static pass over the
    final graph: level and sigma exactly, delta and BC within rounding."""
    hg = sbc.gen_chung_lu(1500, 2.3, 8.0, seed=2, keep_lcc=True)
    roots = hg.pick_roots(40, 3)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    rng = np.random.default_rng(5)
    src = np.repeat(np.arange(hg.nv), np.diff(hg.rowptr))
    edges = [(int(a), int(b)) for a, b in zip(src, hg.col) if a < b]
    gone = [edges[i] for i in rng.choice(len(edges), 80, replace=False)]
    tot = np.zeros(3, dtype=np.int64)
    for a, b in gone:
        tot += o.delete(a, b)
    assert tot.sum() == 80 * len(roots) and tot[1] > 0 and tot[2] > 0
    keep = sorted(set(edges) - set(gone))
    # same adjacency order as the oracle's graph: the original lists minus the deleted entries
    dead = set(gone) | {(b, a) for a, b in gone}
    col = [int(x) for a, x in zip(src, hg.col) if (int(a), int(x)) not in dead]
    rowptr = np.zeros(hg.nv + 1, dtype=np.int64)
    np.add.at(rowptr, np.array([a for a, x in zip(src, hg.col) if (int(a), int(x)) not in dead]) + 1, 1)
    rowptr = np.cumsum(rowptr)
    assert rowptr[-1] == 2 * len(keep)
    s = Oracle(rowptr, np.array(col, dtype=np.int64), roots)
    s.static(PULL)
    lv, sg, dl = o.trees()
    ls, ss, ds = s.trees()
    assert np.array_equal(lv, ls) and np.array_equal(sg, ss)
    assert_bc_close(dl, ds, what="delta")
    assert_bc_close(o.bc(), s.bc())


def _edges_of(hg):
    src = np.repeat(np.arange(hg.nv), np.diff(hg.rowptr))
    return [(int(a), int(b)) for a, b in zip(src, hg.col) if a < b]


@pytest.mark.parametrize("name", ["chunglu_sparse", "rmat11"])
def test_delete_in_place_equals_rebuild_from_static_trees(name):
    """The in-place movement case gives the rebuild's trees bit for bit when
    the trees before the deletion are the static pass's: for each deleted
    edge, two fresh forests, one per mode."""
    if name == "chunglu_sparse":
        hg = sbc.gen_chung_lu(1500, 2.3, 3.0, seed=4, keep_lcc=True)
    else:
        hg = sbc.gen_rmat(11, 8 << 11, seed=3, keep_lcc=True)
    roots = hg.pick_roots(40, 5)
    edges = _edges_of(hg)
    rng = np.random.default_rng(8)
    moved = 0
    for i in rng.choice(len(edges), 25, replace=False):
        a, b = edges[i]
        ob = Oracle(hg.rowptr, hg.col, roots)
        ob.static(PULL)
        oi = Oracle(hg.rowptr, hg.col, roots)
        oi.static(PULL)
        cb = ob.delete(a, b, MOVE_REBUILD)
        ci = oi.delete(a, b, MOVE_IN_PLACE)
        assert np.array_equal(cb, ci)
        moved += int(ci[2])
        for x, y in zip(ob.trees(), oi.trees()):
            assert np.array_equal(x, y), f"deleting ({a},{b})"
        assert_bc_close(oi.bc(), ob.bc(), what=f"BC after deleting ({a},{b})")
    assert moved >= 10, f"only {moved} movement-case roots"


def test_delete_in_place_after_updates_is_close_to_rebuild():
    """After a mix of insertions and deletions, in place and rebuilt agree on
    levels and path counts exactly and on deltas and BC within rounding."""
    hg = sbc.gen_chung_lu(1500, 2.3, 3.0, seed=7, keep_lcc=True)
    roots = hg.pick_roots(40, 6)
    edges = _edges_of(hg)
    rng = np.random.default_rng(9)
    dels = [edges[i] for i in rng.choice(len(edges), 60, replace=False)]
    u, v = hg.insertions(30, 4)
    ob = Oracle(hg.rowptr, hg.col, roots)
    oi = Oracle(hg.rowptr, hg.col, roots)
    ob.static(PULL)
    oi.static(PULL)
    moved = 0
    for i, (a, b) in enumerate(dels):
        if i < len(u):
            ob.insert(int(u[i]), int(v[i]), FINE)
            oi.insert(int(u[i]), int(v[i]), FINE)
        cb = ob.delete(a, b, MOVE_REBUILD)
        ci = oi.delete(a, b, MOVE_IN_PLACE)
        assert np.array_equal(cb, ci)
        moved += int(ci[2])
    assert moved >= 20
    (lb, sb_, db), (li, si, di) = ob.trees(), oi.trees()
    assert np.array_equal(lb, li) and np.array_equal(sb_, si)
    assert_bc_close(di, db, what="delta")
    assert_bc_close(oi.bc(), ob.bc())


@pytest.mark.parametrize("name", ["karate", "dolphins"])
def test_oracle_delete_in_place_matches_reference_golden(name):
    """The in-place movement case against the reference's own deletions
    (FINE): case counts exact, levels and path counts exact, BC and delta
    within 1e-9."""
    g = np.load(os.path.join(GOLDEN, f"refdel_{name}.npz"))
    hg = load_graph(name)
    o = Oracle(hg.rowptr, hg.col, np.arange(hg.nv))
    o.static(PUSH)
    for i, (op, a, b) in enumerate(g["ops"]):
        if op == 1:
            o.insert(int(a), int(b), FINE)
        else:
            c = o.delete(int(a), int(b), MOVE_IN_PLACE)
            assert np.array_equal(c, g["fine_counts"][i]), f"case counts of operation {i}"
        assert_bc_close(o.bc(), g["fine_bc"][i + 1], what=f"BC after operation {i}")
    lv, sg, dl = o.trees()
    assert np.array_equal(lv, g["fine_level"]) and np.array_equal(sg, g["fine_sigma"])
    assert_bc_close(dl, g["fine_delta"], what="delta")


def test_delete_closed_forms():
    """Cutting path 1..9 at (4,5) leaves paths of 4 and 5 vertices, whose BC is
    2(k-1)(m-k) for the k-th vertex of an m-vertex path; cutting a leaf of a
    star lowers the centre's BC from (n-1)(n-2) to (n-2)(n-3)."""
    n = 9
    path = from_edges(n, [(i, i + 1) for i in range(1, n)])
    for mode in (COARSE, FINE):
        o = Oracle(path.rowptr, path.col, np.arange(path.nv))
        o.static(PUSH)
        c = o.delete(4, 5)
        bc = o.bc()
        for k in range(1, 5):
            assert bc[k] == 2 * (k - 1) * (4 - k)
        for k in range(5, 10):
            assert bc[k] == 2 * (k - 5) * (9 - k)
        # roots 1..9: all see (4,5) on different levels; no other parent: movement
        assert c[0] == 1 and c[1] == 0 and c[2] == 9
        star = from_edges(n, [(1, i) for i in range(2, n + 1)])
        o = Oracle(star.rowptr, star.col, np.arange(star.nv))
        o.static(PUSH)
        o.delete(1, 9)
        bc = o.bc()
        assert bc[1] == (n - 2) * (n - 3) and np.all(bc[2:] == 0)
        lv, sg, _ = o.tree(9)
        assert lv[9] == 0 and np.all(lv[1:9] == LONG_MAX) and np.all(sg[1:9] == LONG_MAX)


def test_planted_delete_error_is_caught():
    """Skipping the recompute of one affected root (a plausible dispatch
    mistake) makes the trees differ from the static recompute."""
    hg = load_graph("karate")
    roots = np.arange(hg.nv)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    saved = o.trees()
    c = o.delete(1, 2)
    assert c[1] + c[2] > 0
    lv, sg, dl = o.trees()
    changed = [ri for ri in range(hg.nv) if not np.array_equal(dl[ri], saved[2][ri])]
    assert changed
    bad = [x.copy() for x in (lv, sg, dl)]
    ri = changed[0]
    for x, y in zip(bad, saved):
        x[ri] = y[ri]
    assert not (np.array_equal(bad[0], lv) and np.array_equal(bad[1], sg) and np.array_equal(bad[2], dl))


@pytest.mark.parametrize("name", ["karate", "dolphins"])
def test_edge_counts_match_reference_golden(name):
    """orc_edge_counts (the inputs of the BALANCE estimate used as a scheduling
    comparator) equals the reference's edgesAbove / edgesBelow of every root's
    tree after its static pass (fixtures written by the reference)."""
    from oracle_py import edge_counts
    fx = np.load(os.path.join(GOLDEN, f"refedges_{name}.npz"))
    hg = sbc.read_dimacs(os.path.join(GOLDEN, "graphs", f"{name}.graph"))
    for ri in range(hg.nv):
        a, b = edge_counts(hg.rowptr, hg.col, fx["level"][ri])
        assert np.array_equal(a, fx["above"][ri]) and np.array_equal(b, fx["below"][ri]), ri


def test_balance_work_closed_form():
    """On the path 1-2-3-4 (phantom 0) rooted at 1, edgesAbove[k] = k - 1 and
    edgesBelow[k] = 4 - k; the window (1,3),(2,4) has the deeper endpoints 3
    and 4, so the estimate is (2*1 + 2) + (2*0 + 3) = 7; same-level
    endpoints add nothing."""
    from oracle_py import balance_work
    hg = sbc.graph_from_edges(5, np.array([1, 2, 3]), np.array([2, 3, 4]))
    lv = np.array([[np.iinfo(np.int64).max, 0, 1, 2, 3]], dtype=np.int64)
    w = balance_work(hg.rowptr, hg.col, lv, [1, 2, 3], [3, 4, 3])
    assert w.tolist() == [7]
