"""CPU tests of the host side: generators, DIMACS IO, streams, and the C ABI
of libstreamingbc_hip.so (loads and exports every declared symbol; no device
calls are made)."""
import ctypes
import os
import re

import numpy as np
import pytest

import streamingbc_amd as sbc
from sbc_testutil import GOLDEN, ROOT


def _check_simple(hg):
    rp, col = hg.rowptr, hg.col
    assert rp[0] == 0 and rp[1] == 0, "phantom vertex 0 must be isolated"
    assert np.all(np.diff(rp) >= 0)
    seen = set()
    for x in range(hg.nv):
        nb = col[rp[x]:rp[x + 1]]
        assert len(set(nb.tolist())) == len(nb), "duplicate neighbour"
        assert x not in set(nb.tolist()), "self loop"
        for y in nb:
            seen.add((x, int(y)))
    for x, y in seen:
        assert (y, x) in seen, "graph not symmetric"


def test_rmat_deterministic_and_simple():
    a = sbc.gen_rmat(10, 8192, seed=1, keep_lcc=True)
    b = sbc.gen_rmat(10, 8192, seed=1, keep_lcc=True)
    c = sbc.gen_rmat(10, 8192, seed=2, keep_lcc=True)
    assert np.array_equal(a.rowptr, b.rowptr) and np.array_equal(a.col, b.col)
    assert not (a.ne == c.ne and np.array_equal(a.col, c.col))
    _check_simple(a)
    lcc = a.largest_component()
    assert lcc[1:].all() and not lcc[0], "keep_lcc leaves one component 1..n'"


def test_rmat_full_keeps_all_ids():
    g = sbc.gen_rmat(12, 16 << 12, seed=1, keep_lcc=False)
    assert g.nv == (1 << 12) + 1
    _check_simple(g)


def test_chung_lu_degrees():
    g = sbc.gen_chung_lu(1 << 12, 2.3, 8.0, seed=1, keep_lcc=True)
    _check_simple(g)
    d = g.degree()[1:]
    assert 5.0 < d.mean() < 9.0
    assert d.max() > 20 * np.median(d), "power-law tail expected"


def test_dimacs_roundtrip(tmp_path):
    g = sbc.read_dimacs(os.path.join(GOLDEN, "graphs", "karate.graph"))
    assert g.nv == 35 and g.ne == 156
    p = str(tmp_path / "k.graph")
    g.write_dimacs(p)
    h = sbc.read_dimacs(p)
    assert np.array_equal(g.rowptr, h.rowptr) and np.array_equal(g.col, h.col)


def test_dimacs_rejects_bad_ids(tmp_path):
    p = tmp_path / "bad.graph"
    p.write_text("3 2 0\n2\n1 7\n\n")
    with pytest.raises(ValueError):
        sbc.read_dimacs(str(p))


def test_insertion_stream_properties():
    g = sbc.gen_chung_lu(2000, 2.3, 8.0, seed=3, keep_lcc=True)
    u, v = g.insertions(500, 9)
    assert len(u) == 500
    lcc = g.largest_component()
    pairs = set()
    for a, b in zip(u, v):
        assert a != b and lcc[a] and lcc[b]
        nb = set(g.col[g.rowptr[a]:g.rowptr[a + 1]].tolist())
        assert int(b) not in nb, "insertion of an existing edge"
        key = (min(a, b), max(a, b))
        assert key not in pairs
        pairs.add(key)
    u2, v2 = g.insertions(500, 9)
    assert np.array_equal(u, u2) and np.array_equal(v, v2)


def test_roots():
    g = sbc.gen_chung_lu(2000, 2.3, 8.0, seed=3, keep_lcc=True)
    r = g.pick_roots(64, 7)
    assert len(set(r.tolist())) == 64
    assert g.largest_component()[r].all()
    assert np.array_equal(g.pick_roots(g.nv, 1), np.arange(g.nv))


def _declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text)
    return [n for n in names if n not in ("if", "sizeof")]


@pytest.mark.parametrize("header", ["streamingbc_hip.h", "sbc_graphgen.h", "sbc_costmodel.h"])
def test_library_exports_declared_symbols(header):
    lib = ctypes.CDLL(sbc.LIB_PATH)
    names = _declared_functions(os.path.join(ROOT, "include", header))
    assert len(names) > 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_footprint_formula():
    # 41 B per (root, vertex) + level offsets + BC
    fp = sbc.forest_footprint(1 << 20, 256)
    assert 256 * (1 << 20) * 41 <= fp < 256 * (1 << 20) * 41 + 64 << 20


def test_remove_edge_return_codes():
    """stinger_remove_edge: 1 removed, -1 absent / self loop / bad id; an
    insertion after a removal is accepted again."""
    hg = sbc.read_dimacs(os.path.join(ROOT, "tests", "golden", "graphs", "karate.graph"))
    g = sbc.Graph(hg)
    d1, d2 = g.degree(1), g.degree(2)
    assert g.remove_edge(1, 2) == 1 and g.remove_edge(2, 1) == 1
    assert g.degree(1) == d1 - 1 and g.degree(2) == d2 - 1
    assert g.remove_edge(1, 2) == -1
    assert g.remove_edge(3, 3) == -1
    assert g.remove_edge(1, hg.nv) == -1
    assert g.insert_edge(1, 2) == 1 and g.degree(1) == d1
    g.close()


def _ckpt_hash(data: bytes) -> int:
    """The checkpoint's section checksum: 8-byte little-endian words (tail
    zero-padded) xored in, times the FNV-64 prime, folded; length last."""
    M = (1 << 64) - 1
    h = 0x9E3779B97F4A7C15

    def word(h, w):
        h = ((h ^ w) * 0x100000001B3) & M
        return h ^ (h >> 32)

    pad = data + bytes((-len(data)) % 8)
    for w in np.frombuffer(pad, dtype="<u8").tolist():
        h = word(h, w)
    return word(h, len(data))


def _ckpt_file(dist_override=None, flip=None):
    """A complete checkpoint of the path 1-2-3 (phantom vertex 0), one root
    (vertex 1), written by hand in the documented layout."""
    nv, K = 4, 1
    deg = np.array([0, 1, 2, 1], np.int32)
    cap = deg + 4
    start = np.concatenate([[0], np.cumsum(cap)[:-1]]).astype(np.int32)
    slots = int(cap.sum())
    col = np.zeros(slots, np.int32)
    col[start[1]] = 2
    col[start[2]:start[2] + 2] = [1, 3]
    col[start[3]] = 2
    seq = np.zeros(slots, np.int32)
    hub = np.zeros((nv, 4), np.int32)
    roots = np.array([1], np.int64)
    dist = np.array([0x7FFFFFFF, 0, 1, 2], np.int32) if dist_override is None else dist_override
    sb = np.array([0x80, 0, 1 << 2, 2 << 2], np.uint8)
    rec = np.zeros((nv, 4), np.float64)
    recb = bytearray(rec.tobytes())
    sig = np.array([np.iinfo(np.int64).max, 1, 1, 1], np.int64)
    dlt = np.array([0.0, 2.0, 1.0, 0.0])
    for x in range(nv):
        recb[32 * x:32 * x + 8] = sig[x].tobytes()
        recb[32 * x + 8:32 * x + 16] = dlt[x].tobytes()
    bc = np.array([0.0, 0.0, 1.0, 0.0])
    hdr = np.array([3, nv, K, 0, 1, slots, 0, 1], np.int64)
    secs = [roots.tobytes(),
            start.tobytes() + deg.tobytes() + deg.tobytes() + cap.tobytes() + col.tobytes() + seq.tobytes()
            + hub.tobytes(),
            dist.tobytes() + sb.tobytes() + bytes(recb),
            bc.tobytes()]
    out = bytearray(b"SBCCKPT3" + hdr.tobytes())
    for sec in secs:
        out += sec + np.array([_ckpt_hash(sec)], "<u8").tobytes()
    out += b"SBCEND03"
    if flip is not None:
        out[flip] ^= 0x10
    return bytes(out)


def test_checkpoint_load_rejects_bad_files(tmp_path):
    """sbcCheckpointLoad checks the header, the exact file size, every
    section's checksum and the ranges of the graph and trees before any
    device allocation: a missing file, a wrong magic, a truncated body, a
    flipped bit anywhere after the header, or a level beyond the level
    capacity is SBC_ERR_ARG.  A sound file passes those checks and only then
    reaches the device (here without a GPU: any status but SBC_ERR_ARG)."""
    p = tmp_path / "state.ckpt"
    with pytest.raises(sbc.SbcError) as e:
        sbc.checkpoint_load(str(p))
    assert e.value.status == sbc.SBC_ERR_ARG
    p.write_bytes(b"NOTACKPT" + bytes(64))
    with pytest.raises(sbc.SbcError) as e:
        sbc.checkpoint_load(str(p))
    assert e.value.status == sbc.SBC_ERR_ARG
    hdr = np.array([3, 35, 35, 1, 1, 200, 0, 1], dtype=np.int64)
    p.write_bytes(b"SBCCKPT3" + hdr.tobytes() + bytes(1000))
    with pytest.raises(sbc.SbcError) as e:
        sbc.checkpoint_load(str(p))
    assert e.value.status == sbc.SBC_ERR_ARG
    good = _ckpt_file()
    # a flipped bit in each section (roots, graph, tree, BC) fails a checksum
    for at in (72 + 3, 72 + 8 + 8 + 20, len(good) - 8 - 8 - 32 - 100, len(good) - 8 - 8 - 3):
        p.write_bytes(_ckpt_file(flip=at))
        with pytest.raises(sbc.SbcError) as e:
            sbc.checkpoint_load(str(p))
        assert e.value.status == sbc.SBC_ERR_ARG and "checksum" in str(e.value), at
    # a level beyond the level capacity, with valid checksums
    p.write_bytes(_ckpt_file(dist_override=np.array([0x7FFFFFFF, 0, 1, 9999], np.int32)))
    with pytest.raises(sbc.SbcError) as e:
        sbc.checkpoint_load(str(p))
    assert e.value.status == sbc.SBC_ERR_ARG and "range" in str(e.value)
    p.write_bytes(good)
    try:
        f, g = sbc.checkpoint_load(str(p))
    except sbc.SbcError as e:
        assert e.status != sbc.SBC_ERR_ARG, str(e)
    else:  # a GPU is present: the file loads
        f.close()
        g.close()
