"""Writes the golden fixtures in tests/golden/ from the reference StreamingBC
compiled from its own sources (oracle/build_ref.sh -> oracle/_ref/libsbcref.so).

Run in a container where /root/reference is mounted:
    python tests/golden/make_golden.py

Fixtures (inputs and the reference's outputs, nothing else):
  ref_<graph>.npz   karate / dolphins / jazz, exact BC (roots 0..NV-1), an
                    insertion stream from sbcGenInsertions; for COARSE and FINE
                    at NT=1, NO_BALANCE: BC after the static pass and after every
                    insertion, final level/sigma/delta of every root, and the
                    reference's traversal counters per insertion
  ref_rmat10.npz    R-MAT scale 10, 8192 draws, largest component, exact, 256
                    insertions (seed 1): BC after every 16th insertion, SHA-256
                    of (level, sigma) of all roots after every insertion, final
                    level/sigma of all roots and delta of the first 64 roots
  nx_<graph>.npz    2 x networkx.betweenness_centrality(normalized=False) of the
                    bundled graph (phantom vertex 0 has BC 0)
  refdel_<graph>.npz  karate / dolphins, exact: a stream of operations (op 1:
                    insert, 0: delete an edge of the current graph, seed 9), for
                    COARSE and FINE at NT=1, NO_BALANCE: BC after the static pass
                    and after every operation, the (sameLevel, adjacent,
                    movement) counts each deletion returns, final level/sigma/delta
  refedges_<graph>.npz  karate / dolphins, exact: the reference's edgesAbove and
                    edgesBelow of every root's tree after its static pass (the
                    inputs of its BALANCE estimate), with the levels they hang on
  refdel_rmat10.npz R-MAT scale 10 as above, exact, 64 deletions (seed 9), COARSE:
                    BC after every 8th deletion, the counts, SHA-256 of (level,
                    sigma) of all roots after every deletion, final level/sigma
                    and the delta of the first 64 roots
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import streamingbc_amd as sbc  # noqa: E402  (host graph generators)
from oracle_py import COARSE, FINE, Reference  # noqa: E402

GRAPHS = {"karate": 20, "dolphins": 40, "jazz": 100}


def tree_hash(lv, sg):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(lv, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(sg, dtype=np.int64).tobytes())
    return np.frombuffer(h.digest(), dtype=np.uint8)


def run_ref(hg, roots, u, v, gran, bc_every=1, keep_delta_roots=None, hashes=False):
    ref = Reference(hg.rowptr, hg.col, roots, nt=1)
    ref.static()
    bcs = [ref.bc()]
    counters = []
    hs = []
    prev = ref.counters()
    for i, (a, b) in enumerate(zip(u, v)):
        ref.insert(int(a), int(b), 0, gran)
        c = ref.counters()
        counters.append(c - prev)
        prev = c
        if (i + 1) % bc_every == 0:
            bcs.append(ref.bc())
        if hashes:
            lv, sg, _ = ref.trees()
            hs.append(tree_hash(lv, sg))
    lv, sg, dl = ref.trees()
    if keep_delta_roots is not None:
        dl = dl[:keep_delta_roots]
    out = dict(bc=np.array(bcs), level=lv, sigma=sg, delta=dl, counters=np.array(counters))
    if hashes:
        out["tree_sha256"] = np.array(hs)
    ref.close()
    return out


def delete_stream(hg, n_ops, seed, insert_every=0):
    """Operations (op, u, v): deletions of edges of the current graph, drawn
    with numpy's default_rng(seed); with insert_every = m, every m-th
    operation instead inserts the next edge of hg.insertions(n_ops, seed)."""
    rng = np.random.default_rng(seed)
    adj = [set(int(x) for x in hg.col[hg.rowptr[a]:hg.rowptr[a + 1]]) for a in range(hg.nv)]
    iu, iv = hg.insertions(n_ops, seed)
    ops, k = [], 0
    for i in range(n_ops):
        if insert_every and (i + 1) % insert_every == 0 and k < len(iu):
            a, b = int(iu[k]), int(iv[k])
            k += 1
            if b in adj[a]:
                continue
            adj[a].add(b)
            adj[b].add(a)
            ops.append((1, a, b))
            continue
        edges = sorted((a, b) for a in range(hg.nv) for b in adj[a] if a < b)
        a, b = edges[int(rng.integers(len(edges)))]
        adj[a].discard(b)
        adj[b].discard(a)
        ops.append((0, a, b))
    return np.array(ops, dtype=np.int64)


def run_ref_ops(hg, roots, ops, gran, bc_every=1, keep_delta_roots=None, hashes=False):
    ref = Reference(hg.rowptr, hg.col, roots, nt=1)
    ref.static()
    bcs = [ref.bc()]
    counts, hs = [], []
    for i, (op, a, b) in enumerate(ops):
        if op == 1:
            ref.insert(int(a), int(b), 0, gran)
            counts.append(np.full(3, -1, dtype=np.int64))
        else:
            counts.append(ref.delete(int(a), int(b), 0, gran))
        if (i + 1) % bc_every == 0:
            bcs.append(ref.bc())
        if hashes:
            lv, sg, _ = ref.trees()
            hs.append(tree_hash(lv, sg))
    lv, sg, dl = ref.trees()
    if keep_delta_roots is not None:
        dl = dl[:keep_delta_roots]
    out = dict(bc=np.array(bcs), counts=np.array(counts), level=lv, sigma=sg, delta=dl)
    if hashes:
        out["tree_sha256"] = np.array(hs)
    ref.close()
    return out


def main_delete():
    for name, n in (("karate", 24), ("dolphins", 40)):
        hg = sbc.read_dimacs(os.path.join(HERE, "graphs", f"{name}.graph"))
        roots = np.arange(hg.nv)
        ops = delete_stream(hg, n, 9, insert_every=4)
        res = {}
        for gname, gran in (("coarse", COARSE), ("fine", FINE)):
            for k, val in run_ref_ops(hg, roots, ops, gran).items():
                res[f"{gname}_{k}"] = val
        np.savez_compressed(os.path.join(HERE, f"refdel_{name}.npz"), ops=ops, **res)
        print(f"refdel_{name}.npz: nv={hg.nv} operations={len(ops)}")
    hg = sbc.gen_rmat(10, 8192, seed=1, keep_lcc=True)
    roots = np.arange(hg.nv)
    ops = delete_stream(hg, 64, 9)
    r = run_ref_ops(hg, roots, ops, COARSE, bc_every=8, keep_delta_roots=64, hashes=True)
    np.savez_compressed(os.path.join(HERE, "refdel_rmat10.npz"), ops=ops, nv=hg.nv, ne=hg.ne,
                        **{f"coarse_{k}": val for k, val in r.items()})
    print(f"refdel_rmat10.npz: nv={hg.nv} deletions={len(ops)}")


def main():
    for name, n in GRAPHS.items():
        path = os.path.join(HERE, "graphs", f"{name}.graph")
        hg = sbc.read_dimacs(path)
        roots = np.arange(hg.nv)
        u, v = hg.insertions(n, 5)
        res = {}
        for gname, gran in (("coarse", COARSE), ("fine", FINE)):
            r = run_ref(hg, roots, u, v, gran)
            for k, val in r.items():
                res[f"{gname}_{k}"] = val
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), u=u, v=v, **res)
        print(f"ref_{name}.npz: nv={hg.nv} insertions={len(u)}")

        try:
            import networkx as nx
        except ImportError:
            continue
        G = nx.Graph()
        G.add_nodes_from(range(hg.nv))
        for x in range(hg.nv):
            for e in range(hg.rowptr[x], hg.rowptr[x + 1]):
                G.add_edge(x, int(hg.col[e]))
        bc = nx.betweenness_centrality(G, normalized=False)
        arr = np.array([2.0 * bc[x] for x in range(hg.nv)])
        np.savez_compressed(os.path.join(HERE, f"nx_{name}.npz"), bc2=arr,
                            networkx_version=np.array(nx.__version__))

    hg = sbc.gen_rmat(10, 8192, seed=1, keep_lcc=True)
    roots = np.arange(hg.nv)
    u, v = hg.insertions(256, 1)
    for gname, gran in (("coarse", COARSE),):
        r = run_ref(hg, roots, u, v, gran, bc_every=16, keep_delta_roots=64, hashes=True)
        np.savez_compressed(os.path.join(HERE, "ref_rmat10.npz"), u=u, v=v, nv=hg.nv, ne=hg.ne,
                            **{f"{gname}_{k}": val for k, val in r.items()})
    print(f"ref_rmat10.npz: nv={hg.nv} insertions={len(u)}")


def main_edges():
    for name in ("karate", "dolphins"):
        hg = sbc.read_dimacs(os.path.join(HERE, "graphs", f"{name}.graph"))
        roots = np.arange(hg.nv)
        ref = Reference(hg.rowptr, hg.col, roots, nt=1)
        ref.static()
        lv, _, _ = ref.trees()
        ab = [ref.edge_counts(ri) for ri in range(hg.nv)]
        ref.close()
        np.savez_compressed(os.path.join(HERE, f"refedges_{name}.npz"), level=lv,
                            above=np.array([a for a, _ in ab]), below=np.array([b for _, b in ab]))
        print(f"refedges_{name}.npz: nv={hg.nv}")


if __name__ == "__main__":
    if sys.argv[1:] == ["delete"]:
        main_delete()
    elif sys.argv[1:] == ["edges"]:
        main_edges()
    else:
        main()
        main_delete()
