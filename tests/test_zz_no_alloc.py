"""The insertion entry point that allocates nothing (sbcInsertEdgeNoAlloc),
called through the C ABI, against the allocating plain form it varies
(sbcInsertEdge, COARSE) on twin forests: item records and trees equal bit
for bit, BC equal after the deterministic recompute.  Two calls in a row
share the fixed scratch: the second runs on a smaller forest (fewer roots,
fewer records) and must not see the first call's contents."""
import numpy as np
import pytest

from oracle_py import COARSE as O_COARSE, PULL, Oracle

pytestmark = pytest.mark.gpu

sbc = pytest.importorskip("streamingbc_amd")


def _twins(hg, roots):
    out = []
    for _ in range(2):
        ctx = sbc.Context(hg.nv)
        g = sbc.Graph(hg)
        f = sbc.Forest(hg.nv, roots)
        f.init_static(g, ctx)
        out.append((ctx, g, f))
    return out


def _check(hg, roots, u, v, what):
    (c1, g1, f1), (c2, g2, f2) = _twins(hg, roots)
    for i, (a, b) in enumerate(zip(u, v)):
        rec = f1.insert_edge_noalloc(g1, c1, int(a), int(b), records=True)
        f2.insert_edge(g2, c2, int(a), int(b), sbc.COARSE)
        want = c2.last_records(f2.k)
        for key in ("root_index", "kase", "gap", "affected", "moved", "processed", "scanned"):
            assert np.array_equal(rec[key], want[key]), f"{what}: insertion {i} {key}"
    for x, y in zip(f1.trees(), f2.trees()):
        assert np.array_equal(x, y), what
    f1.recompute_bc(c1)
    f2.recompute_bc(c2)
    assert np.array_equal(f1.bc(), f2.bc()), what
    return f1


def test_no_alloc_insertion_matches_plain_twice():
    hg = sbc.gen_rmat(12, 16 << 12, seed=9, keep_lcc=True)
    roots = hg.pick_roots(96, 4)
    u, v = hg.insertions(24, 31)
    print(f"[no-alloc] first call: rmat12 nv={hg.nv} K={len(roots)} n={len(u)}", flush=True)
    f = _check(hg, roots, u, v, "first")
    # and the oracle agrees on the trees (COARSE order: delta within rounding)
    o = Oracle(hg.rowptr, hg.col, roots)
    o.static(PULL)
    for a, b in zip(u, v):
        o.insert(int(a), int(b), O_COARSE)
    lv, sg, _ = f.trees()
    lo, so, _ = o.trees()
    assert np.array_equal(lv, lo) and np.array_equal(sg, so)
    # second call: a smaller forest on another graph
    hg2 = sbc.gen_chung_lu(2048, 2.3, 8.0, seed=5, keep_lcc=True)
    roots2 = hg2.pick_roots(16, 3)
    u2, v2 = hg2.insertions(12, 7)
    print(f"[no-alloc] second call: chunglu nv={hg2.nv} K={len(roots2)} n={len(u2)}", flush=True)
    _check(hg2, roots2, u2, v2, "second (smaller)")


def test_no_alloc_refusals_change_nothing():
    hg = sbc.gen_chung_lu(2048, 2.3, 8.0, seed=5, keep_lcc=True)
    roots = hg.pick_roots(8, 3)
    (c1, g1, f1), _ = _twins(hg, roots)
    before = f1.trees()
    a = int(hg.col[hg.rowptr[5]])
    with pytest.raises(sbc.SbcError) as e:
        f1.insert_edge_noalloc(g1, c1, 5, a)  # already an edge
    assert e.value.status == sbc.SBC_ERR_EXISTS
    # fill one vertex's append slack with allocating insertions, then the
    # no-allocation form refuses the relayout the next one would need
    x = int(np.argmax(hg.degree()))
    cands = [y for y in range(1, hg.nv) if y != x and y not in set(hg.col[hg.rowptr[x]:hg.rowptr[x + 1]].tolist())]
    room = 4 + int(hg.degree()[x]) // 16
    for y in cands[:room]:
        f1.insert_edge(g1, c1, x, int(y), sbc.COARSE)
    before = f1.trees()
    deg = g1.degree(x)
    with pytest.raises(sbc.SbcError) as e:
        f1.insert_edge_noalloc(g1, c1, x, int(cands[room]))
    assert e.value.status == sbc.SBC_ERR_STATE
    assert g1.degree(x) == deg
    for p, q in zip(before, f1.trees()):
        assert np.array_equal(p, q)
