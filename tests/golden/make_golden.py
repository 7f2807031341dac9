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


if __name__ == "__main__":
    main()
