"""MI355X streaming betweenness centrality under edge insertions.

Python bindings (ctypes) over the C ABI of ``lib/libstreamingbc_hip.so``
(declared in ``include/streamingbc_hip.h`` and ``include/sbc_graphgen.h``).
The library is the only compute path: importing this package fails loudly if
it has not been built (``python streamingbc_amd/build.py``), and every device
call raises :class:`SbcError` if the HIP runtime reports an error.

Names follow the reference's C interface (src/streamingbc.h:40-64):
``Forest`` wraps ``bcForest`` (streamingBCCreateForest*), ``Context`` wraps
the ``extraArraysPerThread`` array (streamingBCCreateAuxilary), ``Graph``
wraps the device ``struct stinger`` (CreateStingerFromCSR).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libstreamingbc_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python streamingbc_amd/build.py` "
        "(the HIP library is the only compute path)")

lib = C.CDLL(LIB_PATH)

i64 = C.c_int64
i32 = C.c_int32
u32 = C.c_uint32
P64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)


class SbcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


(SBC_OK, SBC_ERR_ARG, SBC_ERR_HIP, SBC_ERR_NOMEM, SBC_ERR_STATE, SBC_ERR_EXISTS, SBC_ERR_LEVELS, SBC_ERR_BUDGET,
 SBC_ERR_ABSENT) = range(9)
COARSE, FINE = 0, 1
NO_BALANCE, BALANCE = 0, 1


class HostGraphC(C.Structure):
    _fields_ = [("nv", i64), ("ne", i64), ("rowptr", P64), ("col", P64)]


class BcForestC(C.Structure):
    _fields_ = [("dev", C.c_void_p), ("totalBC", PD), ("NV", i64)]


class ItemRecord(C.Structure):
    _fields_ = [("root_index", i32), ("insertion", i32), ("kase", i32), ("gap", i32),
                ("processed", i64), ("scanned", i64), ("affected", i64), ("moved", i64)]


class ExtraInfoC(C.Structure):
    _fields_ = [("movement", i64), ("adjacent", i64), ("sameLevel", i64), ("connectedComponents", i64)]


class BatchConfigC(C.Structure):
    _fields_ = [("window", i32), ("block_threads", i32), ("grid_blocks", i32), ("use_graph", i32),
                ("recompute_bc", i32), ("heavy_roots", i32), ("root_order", C.POINTER(C.c_int32)),
                ("park_min", i32), ("use_context_order", i32), ("park_scan", i64)]


RECORD_DTYPE = np.dtype([("root_index", np.int32), ("insertion", np.int32), ("kase", np.int32),
                         ("gap", np.int32), ("processed", np.int64), ("scanned", np.int64),
                         ("affected", np.int64), ("moved", np.int64)])
assert RECORD_DTYPE.itemsize == C.sizeof(ItemRecord)

HGP = C.POINTER(HostGraphC)
FP = C.POINTER(BcForestC)
VP = C.c_void_p
VPP = C.POINTER(C.c_void_p)


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


# generators (include/sbc_graphgen.h)
_sig("sbcGenRmat", HGP, [C.c_int, i64, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int])
_sig("sbcGenChungLu", HGP, [i64, C.c_double, C.c_double, C.c_uint64, C.c_int])
_sig("sbcGraphFromEdges", HGP, [i64, i64, P64, P64])
_sig("sbcReadDimacs", HGP, [C.c_char_p])
_sig("sbcWriteDimacs", C.c_int, [HGP, C.c_char_p])
_sig("sbcHostGraphFree", None, [HGP])
_sig("sbcLargestComponent", i64, [HGP, C.POINTER(C.c_uint8)])
_sig("sbcPickRoots", C.c_int, [HGP, i64, C.c_uint64, C.c_int, P64])
_sig("sbcGenInsertions", i64, [HGP, i64, C.c_uint64, P64, P64])
# engine (include/streamingbc_hip.h)
_sig("sbcLastError", C.c_int, [])
_sig("sbcLastErrorString", C.c_char_p, [])
_sig("sbcGraphCreate", C.c_int, [i64, P64, P64, VPP])
_sig("sbcForestCreate", C.c_int, [i64, i64, P64, C.POINTER(FP)])
_sig("sbcInitStatic", C.c_int, [FP, VP, VPP])
_sig("sbcInsertEdge", C.c_int, [FP, VP, VPP, i64, i64, u32])
_sig("sbcInsertEdgeNoAlloc", C.c_int, [FP, VP, VPP, i64, i64, C.POINTER(ItemRecord)])
_sig("sbcDeleteEdge", C.c_int, [FP, VP, VPP, i64, i64, C.POINTER(ExtraInfoC)])
_sig("sbcInsertVertex", C.c_int, [FP, VP, VPP, i64, P64, i64])
_sig("sbcCheckpointSave", C.c_int, [FP, VP, C.c_char_p])
_sig("sbcCheckpointLoad", C.c_int, [C.c_char_p, C.POINTER(FP), C.POINTER(C.c_void_p)])
_sig("sbcDeleteVertex", C.c_int, [FP, VP, VPP, i64, P64, i64, P64, C.POINTER(ExtraInfoC)])
_sig("streamingBCCreateAuxilary", VPP, [i64, i64])
_sig("streamingBCDeleteAuxilary", None, [VPP, i64, i64])
_sig("streamingBCDeleteForestApproximate", None, [C.POINTER(FP), i64, P64])
_sig("stinger_free_all", VP, [VP])
_sig("stinger_insert_edge", C.c_int, [VP, i64, i64, i64, i64, i64])
_sig("stinger_remove_edge", C.c_int, [VP, i64, i64, i64])
_sig("stinger_outdegree", i64, [VP, i64])
_sig("streamingBCGetBC", C.c_int, [FP, PD])
_sig("streamingBCGetTree", C.c_int, [FP, i64, P64, P64, PD])
_sig("streamingBCRecomputeBC", C.c_int, [FP, VPP])
_sig("streamingBCInsertEdgesBatched", C.c_int, [FP, VP, VPP, P64, P64, i64, C.POINTER(BatchConfigC),
                                                 C.POINTER(ItemRecord)])
_sig("insertEdgeStreamingBC", ExtraInfoC, [FP, VP, i64, i64, P64, i64, i64, i64, VPP, u32, u32])
_sig("deleteEdgeStreamingBC", ExtraInfoC, [FP, VP, i64, i64, P64, i64, i64, i64, VPP, u32, u32])
_sig("streamingBCLastRecords", C.c_int, [VPP, C.POINTER(ItemRecord), i64])
_sig("streamingBCKernelStats", C.c_int, [VPP, i32, P64, PD])
_sig("streamingBCResetStats", C.c_int, [VPP, i32])
_sig("streamingBCForestFootprint", i64, [i64, i64])
_sig("streamingBCResidentFootprint", i64, [P64, i64, i64, i32])
_sig("streamingBCSynchronize", C.c_int, [VPP])
_sig("streamingBCLastParked", i64, [VPP])
_sig("streamingBCComputeUnits", i32, [VPP])
_sig("streamingBCWorkTotals", C.c_int, [VPP, i64, P64])
_sig("streamingBCGraphCounts", C.c_int, [VPP, P64, P64])
# cost model (include/sbc_costmodel.h)
PF = C.POINTER(C.c_float)
_sig("sbcCostModelNumParams", i32, [])
_sig("sbcCostModelCreate", C.c_int, [C.c_uint64, VPP])
_sig("sbcCostModelFree", None, [VP])
_sig("sbcCostModelGetParams", C.c_int, [VP, PF])
_sig("sbcCostModelSetParams", C.c_int, [VP, PF])
_sig("sbcCostModelTrain", C.c_int, [VP, PF, PF, i64, i32, i32, C.c_float, C.c_uint64, PF])
_sig("sbcCostModelGradients", C.c_int, [VP, PF, PF, i64, PF, PF])
_sig("sbcCostModelPredict", C.c_int, [VP, PF, i64, PF])
_sig("sbcCostModelPredictFP32", C.c_int, [VP, PF, i64, PF])
_sig("streamingBCWindowFeatures", C.c_int, [FP, VP, VPP, P64, P64, i64, PF, PF])
_sig("streamingBCScheduleWindow", C.c_int, [FP, VP, VPP, VP, P64, P64, i64, PF, C.c_float, C.c_float,
                                            C.POINTER(C.c_int32), PF])
_sig("streamingBCSetRootHistory", C.c_int, [VPP, i64, PF])
_sig("streamingBCUpdateRootHistory", C.c_int, [VPP, i64, C.c_float])


def _check(st):
    if st != SBC_OK:
        raise SbcError(st, lib.sbcLastErrorString().decode(errors="replace"))


def _p64(a):
    return a.ctypes.data_as(P64)


def _pd(a):
    return a.ctypes.data_as(PD)


# ----------------------------------------------------------------- host graphs

@dataclass
class HostGraph:
    """Undirected simple graph in CSR form with the reference's DIMACS
    convention: ids 1..n, isolated phantom vertex 0, nv = n + 1."""
    rowptr: np.ndarray  # int64[nv + 1]
    col: np.ndarray     # int64[ne]

    @property
    def nv(self) -> int:
        return int(self.rowptr.shape[0] - 1)

    @property
    def ne(self) -> int:
        return int(self.col.shape[0])

    def degree(self) -> np.ndarray:
        return np.diff(self.rowptr)

    def _as_c(self):
        hg = HostGraphC(self.nv, self.ne, _p64(self.rowptr), _p64(self.col))
        return hg

    def largest_component(self) -> np.ndarray:
        hg = self._as_c()
        mem = np.zeros(self.nv, dtype=np.uint8)
        lib.sbcLargestComponent(C.byref(hg), mem.ctypes.data_as(C.POINTER(C.c_uint8)))
        return mem.astype(bool)

    def pick_roots(self, k: int, seed: int, from_lcc: bool = True) -> np.ndarray:
        hg = self._as_c()
        out = np.zeros(k, dtype=np.int64)
        if lib.sbcPickRoots(C.byref(hg), k, seed, 1 if from_lcc else 0, _p64(out)) != 0:
            raise ValueError("cannot pick that many roots")
        return out

    def insertions(self, count: int, seed: int):
        hg = self._as_c()
        u = np.zeros(count, dtype=np.int64)
        v = np.zeros(count, dtype=np.int64)
        n = lib.sbcGenInsertions(C.byref(hg), count, seed, _p64(u), _p64(v))
        return u[:n], v[:n]

    def write_dimacs(self, path: str) -> None:
        hg = self._as_c()
        if lib.sbcWriteDimacs(C.byref(hg), path.encode()) != 0:
            raise OSError(f"cannot write {path}")


def _take(hgp) -> HostGraph:
    if not hgp:
        raise ValueError("graph generation failed")
    g = hgp.contents
    rowptr = np.ctypeslib.as_array(g.rowptr, shape=(g.nv + 1,)).copy()
    col = np.ctypeslib.as_array(g.col, shape=(g.ne,)).copy() if g.ne > 0 else np.zeros(0, np.int64)
    lib.sbcHostGraphFree(hgp)
    return HostGraph(rowptr, col)


def gen_rmat(scale: int, draws: int, a=0.57, b=0.19, c=0.19, seed: int = 1, keep_lcc: bool = True) -> HostGraph:
    return _take(lib.sbcGenRmat(scale, draws, a, b, c, seed, 1 if keep_lcc else 0))


def gen_chung_lu(n: int, gamma: float = 2.3, mean_deg: float = 8.0, seed: int = 1, keep_lcc: bool = True) -> HostGraph:
    return _take(lib.sbcGenChungLu(n, gamma, mean_deg, seed, 1 if keep_lcc else 0))


def graph_from_edges(n: int, eu, ev) -> HostGraph:
    eu = np.ascontiguousarray(eu, dtype=np.int64)
    ev = np.ascontiguousarray(ev, dtype=np.int64)
    return _take(lib.sbcGraphFromEdges(n, eu.shape[0], _p64(eu), _p64(ev)))


def read_dimacs(path: str) -> HostGraph:
    return _take(lib.sbcReadDimacs(path.encode()))


# ----------------------------------------------------------------- device engine

class Context:
    """Device work context (HIP stream, timing, staging): the reference's
    extraArraysPerThread array, src/streamingbc.h:47-48."""

    def __init__(self, nv: int = 0):
        self._p = lib.streamingBCCreateAuxilary(1, nv)
        if not self._p:
            raise SbcError(SBC_ERR_HIP, "context creation failed")

    def close(self):
        if self._p:
            lib.streamingBCDeleteAuxilary(self._p, 1, 0)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset_stats(self, timing: bool = True):
        _check(lib.streamingBCResetStats(self._p, 1 if timing else 0))

    def kernel_stats(self, kernel: int):
        n = i64(0)
        ms = C.c_double(0)
        _check(lib.streamingBCKernelStats(self._p, kernel, C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    @property
    def num_cu(self) -> int:
        """Compute units of the context's device."""
        return int(lib.streamingBCComputeUnits(self._p))

    def work_totals(self, k: int) -> dict:
        """Per-root work totals of the batched calls since reset_stats: arrays
        of K for nonzero items, processed, scanned, affected, moved and the
        largest affected count of one item."""
        out = np.zeros(8 * k, dtype=np.int64)
        _check(lib.streamingBCWorkTotals(self._p, k, _p64(out)))
        t = out.reshape(k, 8)
        return {"nonzero": t[:, 0], "processed": t[:, 1], "scanned": t[:, 2], "affected": t[:, 3],
                "moved": t[:, 4], "max_affected": t[:, 5]}

    def graph_counts(self):
        """(window graphs built, windows launched as graph replays) of the
        batched calls of this context."""
        b = np.zeros(1, dtype=np.int64)
        n = np.zeros(1, dtype=np.int64)
        _check(lib.streamingBCGraphCounts(self._p, _p64(b), _p64(n)))
        return int(b[0]), int(n[0])

    def last_parked(self) -> int:
        """Items of the last batched call finished by the whole-GPU kernels."""
        return int(lib.streamingBCLastParked(self._p))

    def synchronize(self):
        _check(lib.streamingBCSynchronize(self._p))

    def last_records(self, k: int) -> np.ndarray:
        out = np.zeros(k, dtype=RECORD_DTYPE)
        _check(lib.streamingBCLastRecords(self._p, out.ctypes.data_as(C.POINTER(ItemRecord)), k))
        return out


KSTAT_UPDATE, KSTAT_STATIC, KSTAT_BC, KSTAT_APPEND, KSTAT_WIDE = 0, 1, 2, 3, 4


class Graph:
    """Device-resident CSR with append slack (replaces STINGER)."""

    def __init__(self, hg: HostGraph):
        self.nv = hg.nv
        p = C.c_void_p()
        _check(lib.sbcGraphCreate(hg.nv, _p64(hg.rowptr), _p64(hg.col), C.byref(p)))
        self._p = p

    def close(self):
        if self._p:
            lib.stinger_free_all(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def insert_edge(self, u: int, v: int) -> int:
        """stinger_insert_edge(u -> v); 1 inserted, 0 exists, -1 invalid."""
        return lib.stinger_insert_edge(self._p, 0, u, v, 0, 0)

    def remove_edge(self, u: int, v: int) -> int:
        """stinger_remove_edge(u -> v); 1 removed, -1 absent or invalid."""
        return lib.stinger_remove_edge(self._p, 0, u, v)

    def degree(self, v: int) -> int:
        return lib.stinger_outdegree(self._p, v)


def checkpoint_save(forest: "Forest", g: Graph, path: str) -> None:
    """Writes the graph and every root's tree and BC to path (sbcCheckpointSave)."""
    _check(lib.sbcCheckpointSave(forest._p, g._p, os.fsencode(path)))


def checkpoint_load(path: str):
    """Rebuilds (Forest, Graph) from a checkpoint (sbcCheckpointLoad)."""
    fp = FP()
    gp = C.c_void_p()
    _check(lib.sbcCheckpointLoad(os.fsencode(path), C.byref(fp), C.byref(gp)))
    hdr = np.fromfile(path, dtype=np.int64, count=8, offset=8)
    nv, k = int(hdr[1]), int(hdr[2])
    g = Graph.__new__(Graph)
    g.nv = nv
    g._p = gp
    f = Forest.__new__(Forest)
    f.nv = nv
    f.roots = np.fromfile(path, dtype=np.int64, count=k, offset=8 + 64)
    f.k = k
    f._p = fp
    return f, g


def forest_footprint(nv: int, k: int) -> int:
    return int(lib.streamingBCForestFootprint(nv, k))


def resident_footprint(hg: HostGraph, k: int, num_cu: int = 0) -> int:
    """Device bytes for a resident graph: forest + device CSR + scratch."""
    return int(lib.streamingBCResidentFootprint(_p64(hg.rowptr), hg.nv, k, num_cu))


@dataclass
class BatchConfig:
    window: int = 64
    block_threads: int = 256
    grid_blocks: int = 0
    use_graph: bool = True
    recompute_bc: bool = False
    heavy_roots: int = -1
    root_order: np.ndarray | None = None
    park_min: int = 0  # 0: library default, < 0: never hand an item to the whole-GPU kernels
    use_context_order: bool = False  # root order left by schedule_window (device)
    park_scan: int = 0  # 0: library default (512K adjacency entries)


class Forest:
    """Per-root BFS trees in HBM plus global BC (bcForest)."""

    def __init__(self, nv: int, roots=None):
        self.nv = nv
        fp = FP()
        if roots is None:
            self.roots = np.arange(nv, dtype=np.int64)
            _check(lib.sbcForestCreate(nv, nv, None, C.byref(fp)))
        else:
            self.roots = np.ascontiguousarray(roots, dtype=np.int64)
            _check(lib.sbcForestCreate(nv, self.roots.shape[0], _p64(self.roots), C.byref(fp)))
        self._p = fp
        self.k = int(self.roots.shape[0])

    def close(self):
        if self._p:
            lib.streamingBCDeleteForestApproximate(C.byref(self._p), self.k, _p64(self.roots))
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_static(self, g: Graph, ctx: Context):
        _check(lib.sbcInitStatic(self._p, g._p, ctx._p))

    def insert_edge(self, g: Graph, ctx: Context, u: int, v: int, granularity: int = FINE):
        """Adds (u,v) to the graph and updates every root (the plain form)."""
        _check(lib.sbcInsertEdge(self._p, g._p, ctx._p, u, v, granularity))

    def insert_edge_noalloc(self, g: Graph, ctx: Context, u: int, v: int, records: bool = False):
        """sbcInsertEdgeNoAlloc: the plain COARSE insertion with its scratch in
        the library's fixed per-GPU array (no device allocation); returns the
        K item records when asked."""
        rec = np.zeros(self.k, dtype=RECORD_DTYPE) if records else None
        rp = rec.ctypes.data_as(C.POINTER(ItemRecord)) if records else None
        _check(lib.sbcInsertEdgeNoAlloc(self._p, g._p, ctx._p, u, v, rp))
        return rec

    def update_inserted(self, g: Graph, ctx: Context, u: int, v: int, granularity: int = FINE):
        """insertEdgeStreamingBC for an edge already added with Graph.insert_edge."""
        lib.insertEdgeStreamingBC(self._p, g._p, u, v, _p64(self.roots), self.k, self.nv, 1, ctx._p, 0,
                                  granularity)

    def delete_edge(self, g: Graph, ctx: Context, u: int, v: int) -> dict:
        """Removes (u,v) from the graph and updates every root; returns the
        root counts of the reference's three deletion cases."""
        info = ExtraInfoC()
        _check(lib.sbcDeleteEdge(self._p, g._p, ctx._p, u, v, C.byref(info)))
        return {"sameLevel": info.sameLevel, "adjacent": info.adjacent, "movement": info.movement}

    def insert_vertex(self, g: Graph, ctx: Context, src: int, adj) -> None:
        """Connects src to every listed vertex (edges already there skipped)."""
        adj = np.ascontiguousarray(adj, dtype=np.int64)
        _check(lib.sbcInsertVertex(self._p, g._p, ctx._p, src, _p64(adj), adj.shape[0]))

    def delete_vertex(self, g: Graph, ctx: Context, src: int):
        """Removes every edge of src; returns (removed neighbours in list
        order, summed case counts)."""
        out = np.zeros(max(g.degree(src), 1), dtype=np.int64)
        n = np.zeros(1, dtype=np.int64)
        info = ExtraInfoC()
        _check(lib.sbcDeleteVertex(self._p, g._p, ctx._p, src, _p64(out), out.shape[0], _p64(n), C.byref(info)))
        return out[:int(n[0])], {"sameLevel": info.sameLevel, "adjacent": info.adjacent, "movement": info.movement}

    def update_deleted(self, g: Graph, ctx: Context, u: int, v: int) -> dict:
        """deleteEdgeStreamingBC for an edge already removed with Graph.remove_edge."""
        info = lib.deleteEdgeStreamingBC(self._p, g._p, u, v, _p64(self.roots), self.k, self.nv, 1, ctx._p, 0, FINE)
        return {"sameLevel": info.sameLevel, "adjacent": info.adjacent, "movement": info.movement}

    def insert_batched(self, g: Graph, ctx: Context, u, v, cfg: BatchConfig, records: bool = False):
        u = np.ascontiguousarray(u, dtype=np.int64)
        v = np.ascontiguousarray(v, dtype=np.int64)
        n = u.shape[0]
        order = None
        c = BatchConfigC(cfg.window, cfg.block_threads, cfg.grid_blocks, 1 if cfg.use_graph else 0,
                         1 if cfg.recompute_bc else 0, cfg.heavy_roots, None, cfg.park_min,
                         1 if cfg.use_context_order else 0, cfg.park_scan)
        if cfg.root_order is not None:
            order = np.ascontiguousarray(cfg.root_order, dtype=np.int32)
            c.root_order = order.ctypes.data_as(C.POINTER(C.c_int32))
        rec = np.zeros(n * self.k, dtype=RECORD_DTYPE) if records else None
        rp = rec.ctypes.data_as(C.POINTER(ItemRecord)) if records else None
        _check(lib.streamingBCInsertEdgesBatched(self._p, g._p, ctx._p, _p64(u), _p64(v), n, C.byref(c), rp))
        return rec.reshape(n, self.k) if records else None

    def bc(self) -> np.ndarray:
        out = np.zeros(self.nv, dtype=np.float64)
        _check(lib.streamingBCGetBC(self._p, _pd(out)))
        return out

    def recompute_bc(self, ctx: Context):
        _check(lib.streamingBCRecomputeBC(self._p, ctx._p))

    def tree(self, ri: int):
        lv = np.zeros(self.nv, dtype=np.int64)
        sg = np.zeros(self.nv, dtype=np.int64)
        dl = np.zeros(self.nv, dtype=np.float64)
        _check(lib.streamingBCGetTree(self._p, ri, _p64(lv), _p64(sg), _pd(dl)))
        return lv, sg, dl

    def levels(self) -> np.ndarray:
        """(K, nv) levels of every root's tree (LONG_MAX: unreachable)."""
        lv = np.zeros((self.k, self.nv), dtype=np.int64)
        for ri in range(self.k):
            _check(lib.streamingBCGetTree(self._p, ri, _p64(lv[ri]), None, None))
        return lv

    def trees(self):
        lv = np.zeros((self.k, self.nv), dtype=np.int64)
        sg = np.zeros((self.k, self.nv), dtype=np.int64)
        dl = np.zeros((self.k, self.nv), dtype=np.float64)
        for ri in range(self.k):
            _check(lib.streamingBCGetTree(self._p, ri, _p64(lv[ri]), _p64(sg[ri]), _pd(dl[ri])))
        return lv, sg, dl


# ----------------------------------------------------------------- cost model

COST_FEATURES = 12


def _pf(a):
    return a.ctypes.data_as(PF)


def _f32(a, shape_tail=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape_tail is not None and a.shape[-1] != shape_tail:
        raise ValueError(f"expected trailing dimension {shape_tail}, got {a.shape}")
    return a


class CostModel:
    """The 12-64-64-1 MLP cost model (sbcCostModel): FP32 training kernels,
    BF16-MFMA inference.  Predicts log1p(affected vertices) of an item."""

    def __init__(self, seed: int = 1):
        p = C.c_void_p()
        _check(lib.sbcCostModelCreate(seed, C.byref(p)))
        self._p = p
        self.n_params = int(lib.sbcCostModelNumParams())

    def close(self):
        if self._p:
            lib.sbcCostModelFree(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(lib.sbcCostModelGetParams(self._p, _pf(out)))
        return out

    @params.setter
    def params(self, p):
        p = _f32(p)
        if p.shape != (self.n_params,):
            raise ValueError(f"expected {self.n_params} parameters")
        _check(lib.sbcCostModelSetParams(self._p, _pf(p)))

    def train(self, X, y, epochs: int = 10, batch: int = 512, lr: float = 1e-3, seed: int = 1) -> float:
        X = _f32(X, COST_FEATURES).reshape(-1, COST_FEATURES)
        y = _f32(y).reshape(-1)
        if X.shape[0] != y.shape[0]:
            raise ValueError("X and y disagree in length")
        loss = C.c_float(0.0)
        _check(lib.sbcCostModelTrain(self._p, _pf(X), _pf(y), X.shape[0], epochs, batch, lr, seed, C.byref(loss)))
        return float(loss.value)

    def gradients(self, X, y):
        X = _f32(X, COST_FEATURES).reshape(-1, COST_FEATURES)
        y = _f32(y).reshape(-1)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float(0.0)
        _check(lib.sbcCostModelGradients(self._p, _pf(X), _pf(y), X.shape[0], _pf(g), C.byref(loss)))
        return g, float(loss.value)

    def predict(self, X, mfma: bool = True) -> np.ndarray:
        X = _f32(X, COST_FEATURES)
        shape = X.shape[:-1]
        X = X.reshape(-1, COST_FEATURES)
        out = np.zeros(X.shape[0], dtype=np.float32)
        f = lib.sbcCostModelPredict if mfma else lib.sbcCostModelPredictFP32
        _check(f(self._p, _pf(X), X.shape[0], _pf(out)))
        return out.reshape(shape)


def window_features(forest: Forest, g: Graph, ctx: Context, u, v, root_hist=None) -> np.ndarray:
    """Cost-model features of every (insertion, root) item of a window, from
    the forest's current state: array (n, K, 12)."""
    u = np.ascontiguousarray(u, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.int64)
    n = u.shape[0]
    X = np.zeros((n, forest.k, COST_FEATURES), dtype=np.float32)
    hp = None
    if root_hist is not None:
        h = _f32(root_hist)
        if h.shape != (forest.k,):
            raise ValueError("root_hist needs one value per root")
        hp = _pf(h)
    _check(lib.streamingBCWindowFeatures(forest._p, g._p, ctx._p, _p64(u), _p64(v), n, hp, _pf(X)))
    return X


def set_root_history(ctx: Context, hist) -> None:
    """Uploads the per-root history feature to the context's device copy."""
    h = _f32(hist)
    _check(lib.streamingBCSetRootHistory(ctx._p, h.shape[0], _pf(h)))


def update_root_history(ctx: Context, k: int, decay: float) -> None:
    """Folds the last batched window's item records into the device history
    (no records copied)."""
    _check(lib.streamingBCUpdateRootHistory(ctx._p, k, decay))


def schedule_window(forest: Forest, g: Graph, ctx: Context, model: CostModel, u, v, root_hist=None,
                    light_max: float = 0.0, item_overhead: float = 4.0, predictions: bool = False):
    """Device scheduling of the next window (streamingBCScheduleWindow):
    leaves the root order in the context for a batched call with
    use_context_order=True.  root_hist None: the context's device history.
    Returns (heavy_roots, predictions (n, K) or None)."""
    u = np.ascontiguousarray(u, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.int64)
    n = u.shape[0]
    hp = None
    if root_hist is not None:
        h = _f32(root_hist)
        if h.shape != (forest.k,):
            raise ValueError("root_hist needs one value per root")
        hp = _pf(h)
    heavy = C.c_int32(-1)
    pred = np.zeros((n, forest.k), dtype=np.float32) if predictions else None
    _check(lib.streamingBCScheduleWindow(forest._p, g._p, ctx._p, model._p, _p64(u), _p64(v), n, hp, light_max,
                                         item_overhead, C.byref(heavy), _pf(pred) if predictions else None))
    return int(heavy.value), pred
