"""xgminer_amd - Python host binding of libxgm (gfx950 graph pattern counting).

The product is the C ABI in ``include/xgm.h`` / ``include/xgm_graph.h``,
implemented by ``libxgm.so`` (built in-tree by ``make -C x-gminer_amd``).
This module binds it with ctypes and mirrors the reference's solver entry
points so callers read like GraphMiner's drivers:

* ``TCSolver(g)``             <- codes/GraphMiner/src/triangle/main.cc:5
* ``CliqueSolver(g, k)``      <- codes/GraphMiner/src/clique/main.cc:6
* ``SglSolver(g, pattern)``   <- codes/GraphMiner/src/sgl/main.cc:7
  (pattern names "diamond" / "rectangle" as in codes/GraphMiner/include/pattern.hh:62-63)

Errors raise ``XGMError`` carrying the C status code; there is no CPU
fallback: if ``libxgm.so`` is missing the import-time load raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxgm.so")

# status codes (include/xgm_graph.h)
XGM_OK, XGM_EINVAL, XGM_ENOMEM, XGM_EUNSUPPORTED, XGM_EHIP, XGM_ENODEV, XGM_EIO = range(7)
# csr flags
CSR_SORTED, CSR_SYMMETRIC, CSR_DAG = 0x1, 0x2, 0x4
# patterns (include/xgm.h)
TRIANGLE, CLIQUE4, CLIQUE5, DIAMOND_EDGE, CYCLE4_EDGE = 1, 2, 3, 4, 5
# vertex-induced 4-motifs (GLUMIN P1 / P2 / P3), each embedding once
STAR3_V, DIAMOND_V, TAILEDTRI_V = 6, 7, 8
PATTERN_NAMES = {TRIANGLE: "triangle", CLIQUE4: "clique4", CLIQUE5: "clique5",
                 DIAMOND_EDGE: "diamond", CYCLE4_EDGE: "rectangle", STAR3_V: "3-star",
                 DIAMOND_V: "diamond-induced", TAILEDTRI_V: "tailed-triangle"}
# strategies
(STRAT_AUTO, STRAT_HASH, STRAT_BSEARCH, STRAT_LUT, STRAT_LIST, STRAT_ELUT, STRAT_LUT_RANK, STRAT_BUCKET,
 STRAT_MERGE) = range(9)
SETOP_INTERSECT, SETOP_DIFF, SETOP_DIFF_EXCL = 1, 2, 3
# strategy flags OR-ed into a set operation (include/xgm.h)
SETOP_MERGE, SETOP_HYBRID = 0x100, 0x200
SETOP_MAX_ANCESTORS = 4
NOALLOC_MAX_PAIRS = 1 << 20
# where task lists are built (Context.set_prep)
PREP_AUTO, PREP_HOST, PREP_DEVICE = 0, 1, 2


class XGMError(RuntimeError):
    def __init__(self, status: int, msg: str = ""):
        self.status = status
        super().__init__(f"xgm status {status}: {msg}")


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64),
                ("rowptr", C.POINTER(C.c_int64)), ("colidx", C.POINTER(C.c_int32)),
                ("flags", C.c_uint32)]


class _Stats(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("max_degree", C.c_int64),
                ("sum_deg_sq", C.c_double), ("sum_in_out", C.c_double)]


class _Opts(C.Structure):
    _fields_ = [("strategy", C.c_int), ("lut_max_degree", C.c_int),
                ("flags", C.c_uint32), ("hash_shift", C.c_int),
                ("wave_max_small", C.c_int), ("wave_max_large", C.c_int), ("hub_rows", C.c_int),
                ("wedge_range", C.c_int)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "flags"}


def _opts(strategy=0, lut_max_degree=0, hash_shift=0, wave_max_small=0, wave_max_large=0, hub_rows=0,
          wedge_range=0):
    return _Opts(strategy, lut_max_degree, 0, hash_shift, wave_max_small, wave_max_large, hub_rows, wedge_range)


class _KTime(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("ms", C.c_float), ("tasks", C.c_int64)]


class _Profile(C.Structure):
    _fields_ = [("total_ms", C.c_float), ("n_kernels", C.c_int), ("k", _KTime * 16)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run `make -C x-gminer_amd` "
                          "(or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P, I64, I32, U64, VP = C.POINTER, C.c_int64, C.c_int32, C.c_uint64, C.c_void_p
    sigs = {
        "xgm_strerror": (C.c_char_p, [C.c_int]),
        "xgm_hgraph_view": (C.c_int, [VP, P(_Csr)]),
        "xgm_hgraph_free": (None, [VP]),
        "xgm_gen_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, U64, P(VP)]),
        "xgm_gen_chung_lu": (C.c_int, [I64, C.c_double, C.c_double, C.c_double, U64, P(VP)]),
        "xgm_gen_erdos_renyi": (C.c_int, [I64, C.c_double, U64, P(VP)]),
        "xgm_build_undirected": (C.c_int, [I64, I64, P(I32), P(I32), P(VP)]),
        "xgm_orient_dag": (C.c_int, [P(_Csr), P(VP)]),
        "xgm_load_csr": (C.c_int, [C.c_char_p, C.c_uint32, P(VP)]),
        "xgm_save_csr": (C.c_int, [C.c_char_p, P(_Csr)]),
        "xgm_load_edgelist": (C.c_int, [C.c_char_p, P(VP)]),
        "xgm_csr_validate": (C.c_int, [P(_Csr)]),
        "xgm_csr_hash": (U64, [P(_Csr)]),
        "xgm_csr_get_stats": (C.c_int, [P(_Csr), P(_Stats)]),
        "xgm_ctx_create": (C.c_int, [C.c_int, P(VP)]),
        "xgm_ctx_destroy": (None, [VP]),
        "xgm_ctx_last_error": (C.c_char_p, [VP]),
        "xgm_graph_upload": (C.c_int, [VP, P(_Csr), P(VP)]),
        "xgm_graph_free": (None, [VP]),
        "xgm_graph_orient": (C.c_int, [VP, VP, P(VP)]),
        "xgm_graph_download": (C.c_int, [VP, VP, P(VP)]),
        "xgm_graph_task_order": (C.c_int, [VP, P(I32), I64, P(I64)]),
        "xgm_graph_tc_items": (C.c_int, [VP, VP, P(_Opts), P(I32), I64, P(I64), P(I64)]),
        "xgm_graph_task_list_names": (C.c_int, [VP, C.c_char_p, I64, P(I64)]),
        "xgm_graph_task_list": (C.c_int, [VP, VP, C.c_char_p, P(I32), I64, P(I64), P(I64), C.c_int,
                                          P(C.c_int)]),
        "xgm_ctx_set_prep": (C.c_int, [VP, C.c_int]),
        "xgm_count": (C.c_int, [VP, VP, C.c_int, P(_Opts), P(U64)]),
        "xgm_count_csr": (C.c_int, [VP, P(_Csr), C.c_int, P(_Opts), P(U64)]),
        "xgm_last_profile": (C.c_int, [VP, P(_Profile)]),
        "xgm_graph_upload_batch": (C.c_int, [VP, P(_Csr), C.c_int, P(VP)]),
        "xgm_graph_instances": (C.c_int, [VP]),
        "xgm_count_instances": (C.c_int, [VP, VP, C.c_int, P(_Opts), P(U64)]),
        "xgm_plan_create": (C.c_int, [VP, VP, VP, P(C.c_int), C.c_int, P(_Opts), P(VP)]),
        "xgm_plan_create_sched": (C.c_int, [VP, VP, VP, P(C.c_int), C.c_int, P(_Opts), P(VP), P(VP)]),
        "xgm_plan_run": (C.c_int, [VP, VP, P(U64), P(C.c_float)]),
        "xgm_plan_free": (None, [VP]),
        "xgm_autotune": (C.c_int, [VP, VP, C.c_int, C.c_int, P(_Opts), P(C.c_float)]),
        "xgm_count_batch": (C.c_int, [VP, P(_Csr), P(C.c_int), C.c_int, P(U64)]),
        "xgm_setop_pairs_noalloc": (C.c_int, [VP, VP, C.c_int, P(I32), P(I32), P(I32), I64, P(C.c_uint32)]),
        "xgm_setop_pairs_anc": (C.c_int, [VP, VP, C.c_int, P(I32), P(I32), P(I32), P(I32), C.c_int, I64,
                                          P(C.c_uint32)]),
        "xgm_copy_bandwidth": (C.c_int, [VP, C.c_size_t, C.c_int, P(C.c_float)]),
        "xgm_motifs": (C.c_int, [VP, VP, C.c_int, P(U64)]),
        "xgm_setop_pairs": (C.c_int, [VP, VP, C.c_int, P(I32), P(I32), P(I32), I64, P(C.c_uint32), P(I64),
                                      P(I32), I64]),
        "xgm_mlp_create": (C.c_int, [C.c_int, U64, P(VP)]),
        "xgm_mlp_free": (None, [VP]),
        "xgm_mlp_last_error": (C.c_char_p, [VP]),
        "xgm_mlp_set_params": (C.c_int, [VP, P(C.c_float)]),
        "xgm_mlp_get_params": (C.c_int, [VP, P(C.c_float)]),
        "xgm_mlp_predict": (C.c_int, [VP, P(C.c_float), I64, P(C.c_float)]),
        "xgm_mlp_train": (C.c_int, [VP, P(C.c_float), P(C.c_float), P(C.c_float), I64, C.c_int, C.c_int,
                                    C.c_float, U64, P(C.c_float)]),
        "xgm_mlp_grads": (C.c_int, [VP, P(C.c_float), P(C.c_float), P(C.c_float), I64, P(C.c_float),
                                    P(C.c_float)]),
        "xgm_mfma_gemm_check": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_float),
                                          P(C.c_float)]),
        "xgm_sched_heuristic": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, P(C.c_int8)]),
        "xgm_sched_features": (C.c_int, [VP, VP, C.c_int, P(C.c_float)]),
        "xgm_sched_learned": (C.c_int, [VP, VP, C.c_int, VP, P(C.c_int8)]),
        "xgm_sched_create": (C.c_int, [VP, VP, C.c_int, P(C.c_int8), P(VP)]),
        "xgm_sched_free": (None, [VP]),
        "xgm_count_sched": (C.c_int, [VP, VP, VP, P(C.c_float), P(U64)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(rc: int, ctx=None):
    if rc != XGM_OK:
        msg = lib.xgm_strerror(rc).decode()
        if ctx is not None:
            msg += " (" + lib.xgm_ctx_last_error(ctx).decode() + ")"
        raise XGMError(rc, msg)


# --------------------------------------------------------------------------
# host graphs

class Graph:
    """A CSR graph: int64 rowptr[n+1], int32 colidx[nnz], flags.

    Either owned by the library (generators, loader, orientation) or a view
    of caller-owned numpy arrays (``Graph.from_arrays``)."""

    def __init__(self, handle=None, rowptr=None, colidx=None, flags=0):
        self._h = handle
        if handle is not None:
            v = _Csr()
            _check(lib.xgm_hgraph_view(handle, C.byref(v)))
            self.n, self.nnz, self.flags = v.n, v.nnz, v.flags
            self.rowptr = np.ctypeslib.as_array(v.rowptr, shape=(v.n + 1,))
            self.colidx = (np.ctypeslib.as_array(v.colidx, shape=(v.nnz,)) if v.nnz
                           else np.zeros(0, np.int32))
        else:
            self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
            self.colidx = np.ascontiguousarray(colidx, dtype=np.int32)
            self.n = len(self.rowptr) - 1
            self.nnz = len(self.colidx)
            self.flags = flags
        self._csr = _Csr(self.n, self.nnz,
                         self.rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                         self.colidx.ctypes.data_as(C.POINTER(C.c_int32)), self.flags)

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            lib.xgm_hgraph_free(self._h)
            self._h = None

    @classmethod
    def from_arrays(cls, rowptr, colidx, flags):
        return cls(rowptr=rowptr, colidx=colidx, flags=flags)

    @property
    def csr(self):
        return C.byref(self._csr)

    def degree(self):
        return np.diff(self.rowptr)

    def orient(self) -> "Graph":
        """Degree-oriented DAG (GraphMiner Graph::orientation semantics)."""
        h = C.c_void_p()
        _check(lib.xgm_orient_dag(self.csr, C.byref(h)))
        return Graph(h)

    def validate(self) -> bool:
        return lib.xgm_csr_validate(self.csr) == XGM_OK

    def hash(self) -> int:
        return int(lib.xgm_csr_hash(self.csr))

    def stats(self) -> dict:
        s = _Stats()
        _check(lib.xgm_csr_get_stats(self.csr, C.byref(s)))
        return dict(n=s.n, nnz=s.nnz, max_degree=s.max_degree,
                    sum_deg_sq=s.sum_deg_sq, sum_in_out=s.sum_in_out)

    def save(self, prefix: str):
        _check(lib.xgm_save_csr(prefix.encode(), self.csr))

    @property
    def num_edges(self) -> int:
        """Undirected edges (= DAG edges)."""
        return self.nnz if self.flags & CSR_DAG else self.nnz // 2


def _gen(fn, *args) -> Graph:
    h = C.c_void_p()
    _check(fn(*args, C.byref(h)))
    return Graph(h)


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1) -> Graph:
    return _gen(lib.xgm_gen_rmat, scale, edge_factor, a, b, c, seed)


def gen_chung_lu(n, avg_degree, beta, w_max, seed) -> Graph:
    return _gen(lib.xgm_gen_chung_lu, n, avg_degree, beta, w_max, seed)


def gen_erdos_renyi(n, avg_degree, seed) -> Graph:
    return _gen(lib.xgm_gen_erdos_renyi, n, avg_degree, seed)


def from_edges(n, src, dst) -> Graph:
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    return _gen(lib.xgm_build_undirected, n, len(src),
                src.ctypes.data_as(C.POINTER(C.c_int32)), dst.ctypes.data_as(C.POINTER(C.c_int32)))


def load_csr(prefix: str, flags=CSR_SORTED | CSR_SYMMETRIC) -> Graph:
    return _gen(lib.xgm_load_csr, prefix.encode(), flags)


def load_edgelist(path: str) -> Graph:
    """Edge-list text ("n m" header, then "u v" lines), ids renumbered in
    order of first appearance; the undirected CSR."""
    return _gen(lib.xgm_load_edgelist, path.encode())


CLI_PATH = os.path.join(_HERE, "xgminer")


# --------------------------------------------------------------------------
# device

@dataclass
class KernelTime:
    name: str
    ms: float
    tasks: int


class Context:
    """One HIP device context (stream, scratch, events).  Not thread-safe."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib.xgm_ctx_create(device, C.byref(h)))
        self._h = h
        self._graphs = weakref.WeakSet()
        self._plans = weakref.WeakSet()

    def close(self):
        if getattr(self, "_h", None) is not None:
            for p in list(self._plans):
                p.close()
            for g in list(self._graphs):
                g.close()
            lib.xgm_ctx_destroy(self._h)
            self._h = None

    __del__ = close

    def upload(self, g: Graph) -> "DeviceGraph":
        return DeviceGraph(self, g)

    def set_prep(self, route: int):
        """Build uncached task lists on the host (PREP_HOST), in HBM
        (PREP_DEVICE), or in HBM unless a host copy of the CSR is at hand
        (PREP_AUTO, the default)."""
        _check(lib.xgm_ctx_set_prep(self._h, route), self._h)

    def orient(self, dg: "DeviceGraph") -> "DeviceGraph":
        """Degree-oriented DAG of an uploaded undirected graph, built in HBM
        (equal to ``upload(g.orient())``)."""
        h = C.c_void_p()
        _check(lib.xgm_graph_orient(self._h, dg._h, C.byref(h)), self._h)
        d = DeviceGraph(self, None, handle=h)
        d.n = dg.n
        return d

    def upload_batch(self, graphs) -> "DeviceGraph":
        """Disjoint union of graphs sharing one kind (DAG or undirected)."""
        return DeviceGraph(self, list(graphs))

    def count_instances(self, dg: "DeviceGraph", pattern: int, **opts) -> np.ndarray:
        """One count per instance of a batch (one entry for a single graph)."""
        out = np.zeros(dg.n_inst, np.uint64)
        o = _opts(**opts)
        _check(lib.xgm_count_instances(self._h, dg._h, pattern, C.byref(o),
                                       out.ctypes.data_as(C.POINTER(C.c_uint64))), self._h)
        return out

    def autotune(self, dg: "DeviceGraph", pattern: int, reps: int = 3):
        """Fastest option set for pattern on dg: (opts dict, ms)."""
        o, ms = _Opts(), C.c_float()
        _check(lib.xgm_autotune(self._h, dg._h, pattern, reps, C.byref(o), C.byref(ms)), self._h)
        return o.as_dict(), float(ms.value)

    def copy_bandwidth(self, nbytes: int = 1 << 30, reps: int = 5) -> float:
        """Measured HBM GB/s of a plain copy kernel (read + write bytes)."""
        g = C.c_float()
        _check(lib.xgm_copy_bandwidth(self._h, nbytes, reps, C.byref(g)), self._h)
        return float(g.value)

    def motifs(self, dg: "DeviceGraph", k: int) -> np.ndarray:
        """Vertex-induced k-motif census (k = 3: wedges, triangles; k = 4:
        3-stars, 4-paths, tailed triangles, 4-cycles, diamonds, 4-cliques),
        one row per instance."""
        per = 2 if k == 3 else 6
        out = np.zeros((dg.n_inst, per), np.uint64)
        _check(lib.xgm_motifs(self._h, dg._h, k, out.ctypes.data_as(C.POINTER(C.c_uint64))), self._h)
        return out

    def count_batch(self, graphs, patterns) -> np.ndarray:
        """One count per (undirected graph, pattern) pair (xgm_count_batch)."""
        arr = (_Csr * len(graphs))(*[g._csr for g in graphs])
        pats = (C.c_int * len(patterns))(*patterns)
        out = np.zeros(len(graphs), np.uint64)
        _check(lib.xgm_count_batch(self._h, arr, pats, len(graphs), out.ctypes.data_as(C.POINTER(C.c_uint64))),
               self._h)
        return out

    def setop_pairs(self, dg: "DeviceGraph", op: int, a, b, bounds=None, materialise=False):
        """Set operation op (SETOP_*) on N(a[i]), N(b[i]) below bounds[i]:
        counts, or (counts, offsets, values) with materialise=True."""
        a = np.ascontiguousarray(a, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        n = len(a)
        pi32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
        bd = None if bounds is None else np.ascontiguousarray(bounds, np.int32)
        pbd = pi32(bd) if bd is not None else None
        counts = np.zeros(n, np.uint32)
        pc = counts.ctypes.data_as(C.POINTER(C.c_uint32))
        if not materialise:
            _check(lib.xgm_setop_pairs(self._h, dg._h, op, pi32(a), pi32(b), pbd, n, pc, None, None, 0), self._h)
            return counts
        offsets = np.zeros(n + 1, np.int64)
        po = offsets.ctypes.data_as(C.POINTER(C.c_int64))
        probe = np.zeros(1, np.int32)
        rc = lib.xgm_setop_pairs(self._h, dg._h, op, pi32(a), pi32(b), pbd, n, pc, po, pi32(probe), 0)
        if rc != XGM_OK and offsets[-1] == 0:   # a failure other than "values_cap too small"
            _check(rc, self._h)
        vals = np.zeros(max(1, int(offsets[-1])), np.int32)
        _check(lib.xgm_setop_pairs(self._h, dg._h, op, pi32(a), pi32(b), pbd, n, pc, po, pi32(vals), len(vals)),
               self._h)
        return counts, offsets, vals[:int(offsets[-1])]

    def setop_pairs_anc(self, dg: "DeviceGraph", op: int, a, b, ancestors, bounds=None):
        """Counts of op on N(a[i]), N(b[i]) below bounds[i], excluding the
        vertices ancestors[i, :] (an n x k int array, k <= 4, -1 unused)."""
        a = np.ascontiguousarray(a, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        anc = np.ascontiguousarray(ancestors, np.int32).reshape(len(a), -1)
        pi32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
        bd = None if bounds is None else np.ascontiguousarray(bounds, np.int32)
        counts = np.zeros(len(a), np.uint32)
        _check(lib.xgm_setop_pairs_anc(self._h, dg._h, op, pi32(a), pi32(b), pi32(bd) if bd is not None else None,
                                       pi32(anc), anc.shape[1], len(a), counts.ctypes.data_as(C.POINTER(C.c_uint32))),
               self._h)
        return counts

    def setop_pairs_noalloc(self, dg: "DeviceGraph", op: int, a, b, bounds=None):
        """Counts of setop_pairs through the allocation-free entry point
        (the library's static staging array; at most NOALLOC_MAX_PAIRS)."""
        a = np.ascontiguousarray(a, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        pi32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
        bd = None if bounds is None else np.ascontiguousarray(bounds, np.int32)
        counts = np.zeros(len(a), np.uint32)
        _check(lib.xgm_setop_pairs_noalloc(self._h, dg._h, op, pi32(a), pi32(b), pi32(bd) if bd is not None else None,
                                           len(a), counts.ctypes.data_as(C.POINTER(C.c_uint32))), self._h)
        return counts

    def plan(self, patterns, dag=None, und=None, opts=None, scheds=None) -> "Plan":
        """scheds: None, or one Schedule (or None) per pattern: clique
        patterns then replay that schedule (e.g. the learned one)."""
        return Plan(self, patterns, dag, und, opts, scheds)

    def count(self, g, pattern: int, strategy: int = STRAT_AUTO, lut_max_degree: int = 0, **kw) -> int:
        opts = _opts(strategy, lut_max_degree, **kw)
        out = C.c_uint64()
        if isinstance(g, DeviceGraph):
            _check(lib.xgm_count(self._h, g._h, pattern, C.byref(opts), C.byref(out)), self._h)
        else:
            _check(lib.xgm_count_csr(self._h, g.csr, pattern, C.byref(opts), C.byref(out)), self._h)
        return int(out.value)

    def last_profile(self):
        p = _Profile()
        _check(lib.xgm_last_profile(self._h, C.byref(p)))
        ks = [KernelTime(p.k[i].name.decode(), p.k[i].ms, p.k[i].tasks) for i in range(p.n_kernels)]
        return p.total_ms, ks


class Schedule:
    """Per-vertex clique strategy codes (0 LUT, 1 rank LUT, 2 edge LUT, 3 edge
    lists, 4 rank LUT one bin up, -1 no task) uploaded as task lists."""

    CODES = ("lut", "lut_rank", "edge_lut", "edge_list", "lut_rank_up")

    def __init__(self, dg: "DeviceGraph", pattern: int, codes):
        self.codes = np.ascontiguousarray(codes, dtype=np.int8)
        h = C.c_void_p()
        _check(lib.xgm_sched_create(dg.ctx._h, dg._h, pattern, self.codes.ctypes.data_as(C.POINTER(C.c_int8)),
                                    C.byref(h)), dg.ctx._h)
        self._h, self.dg, self.pattern = h, dg, pattern

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib.xgm_sched_free(self._h)
            self._h = None

    __del__ = close

    def count(self, log=False):
        """Returns count, or (count, per-vertex microseconds) with log=True."""
        out = C.c_uint64()
        us = np.zeros(self.dg.n, np.float32) if log else None
        _check(lib.xgm_count_sched(self.dg.ctx._h, self.dg._h, self._h,
                                   us.ctypes.data_as(C.POINTER(C.c_float)) if log else None, C.byref(out)),
               self.dg.ctx._h)
        return (int(out.value), us) if log else int(out.value)


class Plan:
    """Counts of several patterns over resident batches captured as one
    hipGraph (xgm_plan_create); run() replays it."""

    def __init__(self, ctx: Context, patterns, dag=None, und=None, opts=None, scheds=None):
        pats = (C.c_int * len(patterns))(*patterns)
        oa = None
        if opts is not None:
            oa = (_Opts * len(patterns))(*[_opts(**o) for o in opts])
        h = C.c_void_p()
        if scheds is None:
            _check(lib.xgm_plan_create(ctx._h, dag._h if dag else None, und._h if und else None, pats,
                                       len(patterns), oa, C.byref(h)), ctx._h)
        else:
            if len(scheds) != len(patterns):
                raise ValueError("one schedule (or None) per pattern")
            sa = (C.c_void_p * len(patterns))(*[s._h if s is not None else None for s in scheds])
            _check(lib.xgm_plan_create_sched(ctx._h, dag._h if dag else None, und._h if und else None, pats,
                                             len(patterns), oa, sa, C.byref(h)), ctx._h)
        self._h, self.ctx, self.patterns = h, ctx, list(patterns)
        # the captured kernels read these graphs' buffers: keep them alive
        # as long as the plan (the library also defers their release)
        self._graphs = (dag, und)
        self.n_inst = (dag or und).n_inst
        ctx._plans.add(self)

    def run(self):
        """(counts[n_patterns, n_instances], replay ms)."""
        out = np.zeros((len(self.patterns), self.n_inst), np.uint64)
        ms = C.c_float()
        _check(lib.xgm_plan_run(self.ctx._h, self._h, out.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(ms)),
               self.ctx._h)
        return out, float(ms.value)

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib.xgm_plan_free(self._h)
            self._h = None
            self._graphs = None

    __del__ = close


class DeviceGraph:
    def sched_heuristic(self, pattern, strategy=STRAT_AUTO, lut_max_degree=0):
        codes = np.zeros(self.n, np.int8)
        _check(lib.xgm_sched_heuristic(self._h, pattern, strategy, lut_max_degree,
                                       codes.ctypes.data_as(C.POINTER(C.c_int8))))
        return codes

    def sched_features(self, pattern):
        F = np.zeros((self.n, MLP_IN), np.float32)
        _check(lib.xgm_sched_features(self.ctx._h, self._h, pattern, F.ctypes.data_as(C.POINTER(C.c_float))),
               self.ctx._h)
        return F

    def sched_learned(self, pattern, model: "MLP"):
        codes = np.zeros(self.n, np.int8)
        _check(lib.xgm_sched_learned(self.ctx._h, self._h, pattern, model._h,
                                     codes.ctypes.data_as(C.POINTER(C.c_int8))), self.ctx._h)
        return codes

    def __init__(self, ctx: Context, g, handle=None):
        h = C.c_void_p()
        if handle is not None:  # built on the device (Context.orient)
            h = handle
            self.instances = None
            self._host = None
            self.n = 0
        elif isinstance(g, list):  # batch of instances (disjoint union)
            arr = (_Csr * len(g))(*[x._csr for x in g])
            _check(lib.xgm_graph_upload_batch(ctx._h, arr, len(g), C.byref(h)), ctx._h)
            self.instances = g
            self._host = None
            self.n = sum(x.n for x in g)
        else:
            _check(lib.xgm_graph_upload(ctx._h, g.csr, C.byref(h)), ctx._h)
            self.instances = None
            self._host = g
            self.n = g.n
        self._h, self.ctx = h, ctx
        self.n_inst = int(lib.xgm_graph_instances(h))
        ctx._graphs.add(self)

    @property
    def host(self):
        """Host CSR of a single graph (downloaded once for a device-built one)."""
        if self._host is None and self.instances is None:
            self._host = self.download()
        return self._host

    def download(self) -> Graph:
        """The CSR as stored in HBM, copied to a new host graph."""
        h = C.c_void_p()
        _check(lib.xgm_graph_download(self.ctx._h, self._h, C.byref(h)), self.ctx._h)
        return Graph(h)

    def task_order(self) -> np.ndarray:
        """DAG vertex tasks (d+ >= 2) by decreasing work estimate."""
        n = C.c_int64()
        _check(lib.xgm_graph_task_order(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.int32)
        _check(lib.xgm_graph_task_order(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), n.value, C.byref(n)))
        return out

    def tc_items(self, **opts):
        """Triangle task items (n x 4: t, k0, k1, 0; tiers concatenated) and
        the items per tier."""
        o = _opts(**opts)
        n, tc = C.c_int64(), (C.c_int64 * 4)()
        _check(lib.xgm_graph_tc_items(self.ctx._h, self._h, C.byref(o), None, 0, C.byref(n), tc), self.ctx._h)
        out = np.zeros(max(1, n.value), np.int32)
        _check(lib.xgm_graph_tc_items(self.ctx._h, self._h, C.byref(o), out.ctypes.data_as(C.POINTER(C.c_int32)),
                                      n.value, C.byref(n), tc), self.ctx._h)
        return out[:n.value].reshape(-1, 4), list(tc)

    def task_list_names(self):
        """Names of the task lists cached on this graph."""
        n = C.c_int64()
        _check(lib.xgm_graph_task_list_names(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(max(1, n.value))
        _check(lib.xgm_graph_task_list_names(self._h, buf, n.value, C.byref(n)))
        return [x for x in buf.raw[:n.value].decode().split("\n") if x]

    def task_list(self, name: str):
        """A cached task list: (ints, tiers concatenated; tasks per tier)."""
        n, nc = C.c_int64(), C.c_int()
        cnt = (C.c_int64 * 16)()
        key = name.encode()
        _check(lib.xgm_graph_task_list(self.ctx._h, self._h, key, None, 0, C.byref(n), cnt, 16, C.byref(nc)),
               self.ctx._h)
        out = np.zeros(max(1, n.value), np.int32)
        _check(lib.xgm_graph_task_list(self.ctx._h, self._h, key, out.ctypes.data_as(C.POINTER(C.c_int32)), n.value,
                                       C.byref(n), cnt, 16, C.byref(nc)), self.ctx._h)
        return out[:n.value], list(cnt)[:nc.value]

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib.xgm_graph_free(self._h)
            self._h = None

    __del__ = close


# --------------------------------------------------------------------------
# reference-shaped solver entry points (GraphMiner main.cc contracts)

_default_ctx = None


def _ctx():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def TCSolver(g: Graph) -> int:
    """Triangles of a DAG-oriented graph (GraphMiner: Graph g(path, USE_DAG))."""
    return _ctx().count(g, TRIANGLE)


def CliqueSolver(g: Graph, k: int) -> int:
    if k == 3:
        return _ctx().count(g, TRIANGLE)
    if k not in (4, 5):
        raise XGMError(XGM_EUNSUPPORTED, f"k={k} not supported")
    return _ctx().count(g, CLIQUE4 if k == 4 else CLIQUE5)


def SglSolver(g: Graph, pattern: str) -> int:
    pid = {"diamond": DIAMOND_EDGE, "rectangle": CYCLE4_EDGE}.get(pattern)
    if pid is None:
        raise XGMError(XGM_EUNSUPPORTED, f"pattern {pattern!r} not supported")
    return _ctx().count(g, pid)


def PatternSolver(g: Graph, k: int) -> int:
    """GLUMIN's pattern entry (GLUMIN::run with --patternID 16 + k,
    xgminer/include/glumin.h:36-86): k = 1 3-star, 2 diamond, 3 tailed
    triangle (vertex-induced, on the undirected graph), 4 / 5 = 4- / 5-clique
    (on its degree-oriented DAG).  Counts are unique embeddings; GLUMIN's own
    P2 kernel reports twice the diamonds."""
    pid = {1: STAR3_V, 2: DIAMOND_V, 3: TAILEDTRI_V, 4: CLIQUE4, 5: CLIQUE5}.get(k)
    if pid is None:
        raise XGMError(XGM_EUNSUPPORTED, f"pattern {k} not supported")
    return _ctx().count(g.orient() if k >= 4 else g, pid)


# --------------------------------------------------------------------------
# learned scheduler cost model (16-64-64-5 MLP on the matrix cores)

MLP_IN, MLP_OUT, MLP_PARAMS = 16, 5, 5573


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    def __init__(self, device: int = 0, seed: int = 1):
        h = C.c_void_p()
        _check(lib.xgm_mlp_create(device, seed, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib.xgm_mlp_free(self._h)
            self._h = None

    __del__ = close

    def _chk(self, rc):
        if rc != XGM_OK:
            raise XGMError(rc, lib.xgm_strerror(rc).decode() + " (" + lib.xgm_mlp_last_error(self._h).decode() + ")")

    @property
    def params(self):
        out = np.zeros(MLP_PARAMS, np.float32)
        self._chk(lib.xgm_mlp_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    @params.setter
    def params(self, p):
        p, pp = _f32(p, (MLP_PARAMS,))
        self._chk(lib.xgm_mlp_set_params(self._h, pp))

    def predict(self, X):
        X, px = _f32(X, (-1, MLP_IN))
        Y = np.zeros((len(X), MLP_OUT), np.float32)
        self._chk(lib.xgm_mlp_predict(self._h, px, len(X), Y.ctypes.data_as(C.POINTER(C.c_float))))
        return Y

    def train(self, X, T, M, epochs=20, batch=4096, lr=1e-3, seed=7):
        X, px = _f32(X, (-1, MLP_IN))
        T, pt = _f32(T, (-1, MLP_OUT))
        M, pm = _f32(M, (-1, MLP_OUT))
        loss = C.c_float()
        self._chk(lib.xgm_mlp_train(self._h, px, pt, pm, len(X), epochs, batch, lr, seed, C.byref(loss)))
        return float(loss.value)

    def grads(self, X, T, M):
        X, px = _f32(X, (-1, MLP_IN))
        T, pt = _f32(T, (-1, MLP_OUT))
        M, pm = _f32(M, (-1, MLP_OUT))
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._chk(lib.xgm_mlp_grads(self._h, px, pt, pm, len(X), g.ctypes.data_as(C.POINTER(C.c_float)),
                                    C.byref(loss)))
        return float(loss.value), g


def mfma_gemm(A, B, device=0):
    A, pa = _f32(A)
    B, pb = _f32(B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    Cm = np.zeros((M, N), np.float32)
    _check(lib.xgm_mfma_gemm_check(device, M, N, K, pa, pb, Cm.ctypes.data_as(C.POINTER(C.c_float))))
    return Cm
