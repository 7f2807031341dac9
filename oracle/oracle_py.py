"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the reference
built from its sources (oracle/_ref/libsbcref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libsbcref.so")

i64 = C.c_int64
P64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
VP = C.c_void_p

PUSH, PULL = 0, 1
COARSE, FINE = 0, 1
MOVE_REBUILD, MOVE_IN_PLACE = 0, 1

ITEM_DTYPE = np.dtype([("kase", np.int32), ("gap", np.int32), ("processed", np.int64), ("scanned", np.int64),
                       ("affected", np.int64), ("moved", np.int64), ("descent", np.int64)])

_olib = None
_rlib = None


def _p64(a):
    return a.ctypes.data_as(P64)


def _pd(a):
    return a.ctypes.data_as(PD)


def oracle_lib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_LIB)
        for name, res, args in [
            ("orc_graph_new", VP, [i64, P64, P64]),
            ("orc_graph_free", None, [VP]),
            ("orc_graph_insert", C.c_int, [VP, i64, i64]),
            ("orc_graph_remove", C.c_int, [VP, i64, i64]),
            ("orc_delete", None, [VP, VP, i64, i64, P64]),
            ("orc_delete_mode", None, [VP, VP, i64, i64, P64, C.c_int]),
            ("orc_graph_degree", i64, [VP, i64]),
            ("orc_forest_new", VP, [i64, i64, P64]),
            ("orc_forest_free", None, [VP]),
            ("orc_static", None, [VP, VP, C.c_int]),
            ("orc_insert", None, [VP, VP, i64, i64, C.c_int, C.c_void_p]),
            ("orc_get_tree", None, [VP, i64, P64, P64, PD]),
            ("orc_set_tree", None, [VP, i64, P64, P64, PD]),
            ("orc_get_bc", None, [VP, PD]),
            ("orc_set_bc", None, [VP, PD]),
            ("orc_sigma_overflowed", C.c_int, [VP]),
            ("orc_balance_work", None, [i64, P64, P64, i64, P64, i64, P64, P64, P64]),
            ("orc_edge_counts", None, [i64, P64, P64, P64, P64, P64]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _olib = L
    return _olib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    global _rlib
    if _rlib is None:
        if not ref_available():
            raise ImportError(f"{REF_LIB} missing: run oracle/build_ref.sh where the reference is mounted")
        L = C.CDLL(REF_LIB)
        for name, res, args in [
            ("ref_new", VP, [i64, i64, P64, P64, i64, P64, i64]),
            ("ref_static", C.c_double, [VP]),
            ("ref_insert", C.c_double, [VP, i64, i64, i64, i64]),
            ("ref_delete", C.c_double, [VP, i64, i64, i64, i64, P64]),
            ("ref_get_tree", None, [VP, i64, P64, P64, PD]),
            ("ref_get_bc", None, [VP, PD]),
            ("ref_get_edge_counts", None, [VP, i64, P64, P64]),
            ("ref_counters", None, [VP, P64]),
            ("ref_free", None, [VP]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _rlib = L
    return _rlib


class Oracle:
    """Sequential CPU restatement (see oracle/sbc_oracle.h)."""

    def __init__(self, rowptr, col, roots):
        L = oracle_lib()
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int64)
        self.roots = np.ascontiguousarray(roots, dtype=np.int64)
        self.nv = self.rowptr.shape[0] - 1
        self.k = self.roots.shape[0]
        self.g = L.orc_graph_new(self.nv, _p64(self.rowptr), _p64(self.col))
        self.f = L.orc_forest_new(self.nv, self.k, _p64(self.roots))

    def close(self):
        L = oracle_lib()
        if self.f:
            L.orc_forest_free(self.f)
            self.f = None
        if self.g:
            L.orc_graph_free(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def static(self, order=PUSH):
        oracle_lib().orc_static(self.f, self.g, order)

    def add_edge(self, u, v) -> int:
        return oracle_lib().orc_graph_insert(self.g, u, v)

    def update(self, u, v, mode=COARSE, items=False):
        rec = np.zeros(self.k, dtype=ITEM_DTYPE) if items else None
        oracle_lib().orc_insert(self.f, self.g, u, v, mode, rec.ctypes.data if items else None)
        return rec

    def insert(self, u, v, mode=COARSE, items=False):
        r = self.add_edge(u, v)
        if r != 1:
            raise ValueError(f"edge ({u},{v}) not insertable: {r}")
        return self.update(u, v, mode, items)

    def delete(self, u, v, move=MOVE_REBUILD):
        """Removes (u,v) and runs the deletion update; returns the root counts
        (sameLevel, adjacent, movement).  move: movement-case roots rebuilt by
        the static pull pass (MOVE_REBUILD) or updated in place
        (MOVE_IN_PLACE)."""
        L = oracle_lib()
        if L.orc_graph_remove(self.g, u, v) != 1:
            raise ValueError(f"edge ({u},{v}) not present")
        out = np.zeros(3, dtype=np.int64)
        L.orc_delete_mode(self.f, self.g, u, v, _p64(out), move)
        return out

    def tree(self, ri):
        lv = np.zeros(self.nv, dtype=np.int64)
        sg = np.zeros(self.nv, dtype=np.int64)
        dl = np.zeros(self.nv, dtype=np.float64)
        oracle_lib().orc_get_tree(self.f, ri, _p64(lv), _p64(sg), _pd(dl))
        return lv, sg, dl

    def trees(self):
        lv = np.zeros((self.k, self.nv), dtype=np.int64)
        sg = np.zeros((self.k, self.nv), dtype=np.int64)
        dl = np.zeros((self.k, self.nv), dtype=np.float64)
        L = oracle_lib()
        for ri in range(self.k):
            L.orc_get_tree(self.f, ri, _p64(lv[ri]), _p64(sg[ri]), _pd(dl[ri]))
        return lv, sg, dl

    def set_trees(self, lv, sg, dl):
        L = oracle_lib()
        lv = np.ascontiguousarray(lv, dtype=np.int64)
        sg = np.ascontiguousarray(sg, dtype=np.int64)
        dl = np.ascontiguousarray(dl, dtype=np.float64)
        for ri in range(self.k):
            L.orc_set_tree(self.f, ri, _p64(lv[ri]), _p64(sg[ri]), _pd(dl[ri]))

    def bc(self):
        out = np.zeros(self.nv, dtype=np.float64)
        oracle_lib().orc_get_bc(self.f, _pd(out))
        return out

    def set_bc(self, bc):
        bc = np.ascontiguousarray(bc, dtype=np.float64)
        oracle_lib().orc_set_bc(self.f, _pd(bc))

    def sigma_overflowed(self) -> bool:
        return bool(oracle_lib().orc_sigma_overflowed(self.f))


def edge_counts(rowptr, col, levels):
    """(edgesAbove, edgesBelow) of one tree from its levels (orc_edge_counts)."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int64)
    levels = np.ascontiguousarray(levels, dtype=np.int64)
    nv = levels.shape[0]
    a = np.zeros(nv, dtype=np.int64)
    b = np.zeros(nv, dtype=np.int64)
    oracle_lib().orc_edge_counts(nv, _p64(rowptr), _p64(col), _p64(levels), _p64(a), _p64(b))
    return a, b


def balance_work(rowptr, col, levels, u, v):
    """The reference's BALANCE estimate per root summed over a window
    (orc_balance_work): levels is (K, nv) int64 with LONG_MAX unreachable."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int64)
    levels = np.ascontiguousarray(levels, dtype=np.int64)
    u = np.ascontiguousarray(u, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.int64)
    K, nv = levels.shape
    out = np.zeros(K, dtype=np.int64)
    oracle_lib().orc_balance_work(nv, _p64(rowptr), _p64(col), K, _p64(levels), u.shape[0], _p64(u), _p64(v),
                                  _p64(out))
    return out


class OracleGroup:
    """The oracle over a root list split into `parts` forests that share one
    graph; each insertion's per-root updates run on host threads (ctypes
    releases the GIL, and roots never touch each other's state).  Every
    root's tree is the one the single-forest oracle computes, bit for bit;
    BC is the sum of the parts' BC (another summation order, so it agrees
    with the single forest within rounding, not bit for bit)."""

    def __init__(self, rowptr, col, roots, parts=None):
        from concurrent.futures import ThreadPoolExecutor
        L = oracle_lib()
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int64)
        self.roots = np.ascontiguousarray(roots, dtype=np.int64)
        self.nv = self.rowptr.shape[0] - 1
        self.k = self.roots.shape[0]
        if parts is None:
            parts = min(self.k, max(1, min(len(os.sched_getaffinity(0)), 16)))
        bounds = np.linspace(0, self.k, parts + 1).astype(np.int64)
        self.g = L.orc_graph_new(self.nv, _p64(self.rowptr), _p64(self.col))
        self.chunks = []  # (first root index, root array, forest)
        for a, b in zip(bounds[:-1], bounds[1:]):
            if b > a:
                r = np.ascontiguousarray(self.roots[a:b])
                self.chunks.append((int(a), r, L.orc_forest_new(self.nv, r.shape[0], _p64(r))))
        self.pool = ThreadPoolExecutor(len(self.chunks))

    def close(self):
        L = oracle_lib()
        if self.pool is not None:
            self.pool.shutdown()
            self.pool = None
        for _, _, f in self.chunks:
            L.orc_forest_free(f)
        self.chunks = []
        if self.g:
            L.orc_graph_free(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _each(self, fn):
        return list(self.pool.map(fn, self.chunks))

    def static(self, order=PULL):
        L = oracle_lib()
        self._each(lambda c: L.orc_static(c[2], self.g, order))

    def insert(self, u, v, mode=FINE, items=False):
        L = oracle_lib()
        r = L.orc_graph_insert(self.g, u, v)
        if r != 1:
            raise ValueError(f"edge ({u},{v}) not insertable: {r}")
        rec = np.zeros(self.k, dtype=ITEM_DTYPE) if items else None

        def one(c):
            a, rr, f = c
            p = rec[a:].ctypes.data if items else None
            L.orc_insert(f, self.g, u, v, mode, p)

        self._each(one)
        return rec

    def tree(self, ri):
        for a, r, f in self.chunks:
            if a <= ri < a + r.shape[0]:
                lv = np.zeros(self.nv, dtype=np.int64)
                sg = np.zeros(self.nv, dtype=np.int64)
                dl = np.zeros(self.nv, dtype=np.float64)
                oracle_lib().orc_get_tree(f, ri - a, _p64(lv), _p64(sg), _pd(dl))
                return lv, sg, dl
        raise IndexError(ri)

    def bc(self):
        acc = np.zeros(self.nv, dtype=np.float64)
        for _, _, f in self.chunks:
            out = np.zeros(self.nv, dtype=np.float64)
            oracle_lib().orc_get_bc(f, _pd(out))
            acc += out
        return acc

    def sigma_overflowed(self) -> bool:
        return any(bool(oracle_lib().orc_sigma_overflowed(f)) for _, _, f in self.chunks)


class Reference:
    """The reference library compiled from its own sources (oracle/_ref)."""

    def __init__(self, rowptr, col, roots, nt=1):
        L = ref_lib()
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int64)
        self.roots = np.ascontiguousarray(roots, dtype=np.int64)
        self.nv = self.rowptr.shape[0] - 1
        self.k = self.roots.shape[0]
        self.h = L.ref_new(self.nv, self.col.shape[0], _p64(self.rowptr), _p64(self.col), self.k,
                           _p64(self.roots), nt)

    def close(self):
        if self.h:
            ref_lib().ref_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def static(self) -> float:
        return ref_lib().ref_static(self.h)

    def insert(self, u, v, lb=0, gran=COARSE) -> float:
        t = ref_lib().ref_insert(self.h, u, v, lb, gran)
        if t < 0:
            raise ValueError(f"edge ({u},{v}) already present")
        return t

    def delete(self, u, v, lb=0, gran=COARSE):
        """Returns the reference's (sameLevel, adjacent, movement) counts."""
        out = np.zeros(3, dtype=np.int64)
        t = ref_lib().ref_delete(self.h, u, v, lb, gran, _p64(out))
        if t < 0:
            raise ValueError(f"edge ({u},{v}) not present")
        return out

    def tree(self, ri):
        lv = np.zeros(self.nv, dtype=np.int64)
        sg = np.zeros(self.nv, dtype=np.int64)
        dl = np.zeros(self.nv, dtype=np.float64)
        ref_lib().ref_get_tree(self.h, ri, _p64(lv), _p64(sg), _pd(dl))
        return lv, sg, dl

    def trees(self):
        lv = np.zeros((self.k, self.nv), dtype=np.int64)
        sg = np.zeros((self.k, self.nv), dtype=np.int64)
        dl = np.zeros((self.k, self.nv), dtype=np.float64)
        L = ref_lib()
        for ri in range(self.k):
            L.ref_get_tree(self.h, ri, _p64(lv[ri]), _p64(sg[ri]), _pd(dl[ri]))
        return lv, sg, dl

    def bc(self):
        out = np.zeros(self.nv, dtype=np.float64)
        ref_lib().ref_get_bc(self.h, _pd(out))
        return out

    def edge_counts(self, ri):
        a = np.zeros(self.nv, dtype=np.int64)
        b = np.zeros(self.nv, dtype=np.int64)
        ref_lib().ref_get_edge_counts(self.h, ri, _p64(a), _p64(b))
        return a, b

    def counters(self):
        out = np.zeros(4, dtype=np.int64)
        ref_lib().ref_counters(self.h, _p64(out))
        return out
