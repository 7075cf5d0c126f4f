"""ORACLE - test infrastructure only (the checker, never the product).

ctypes binding of ``oracle/liboracle.so`` (a C++/OpenMP restatement of the
reference GraphMiner CPU path, see oracle.cpp) and a runner for the
reference's own GraphMiner binaries built from source into ``oracle/_ref/``
by ``oracle/Makefile``.  Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this package.

Pinning status of each function is recorded in oracle.cpp's header and the
tests that pin it (tests/test_oracle.py).
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P, I64, I32, U64 = C.POINTER, C.c_int64, C.c_int32, C.c_uint64
        for name in ("orc_tc", "orc_clique4", "orc_clique5", "orc_diamond", "orc_cycle4"):
            f = getattr(L, name)
            f.restype = U64
            f.argtypes = [I64, P(I64), P(I32), I64, I64, C.c_int, P(U64)]
        for name in ("orc_intersect_num",):
            f = getattr(L, name)
            f.restype = U64
            f.argtypes = [P(I32), I64, P(I32), I64]
        L.orc_intersect_num_bound.restype = U64
        L.orc_intersect_num_bound.argtypes = [P(I32), I64, P(I32), I64, I32]
        L.orc_intersect_set.restype = I64
        L.orc_intersect_set.argtypes = [P(I32), I64, P(I32), I64, P(I32)]
        L.orc_difference_num.restype = U64
        L.orc_difference_num.argtypes = [P(I32), I64, P(I32), I64, I32, I32]
        L.orc_difference_num_noexcl.restype = U64
        L.orc_difference_num_noexcl.argtypes = [P(I32), I64, P(I32), I64, I32]
        L.orc_orient.restype = I64
        L.orc_orient.argtypes = [I64, P(I64), P(I32), P(I64), P(I32)]
        _lib = L
    return _lib


def _p32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


def _p64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(C.POINTER(C.c_int64))


# ---- set primitives ----
def intersect_num(a, b):
    a, pa = _p32(a); b, pb = _p32(b)
    return int(lib().orc_intersect_num(pa, len(a), pb, len(b)))


def intersect_num_bound(a, b, up):
    a, pa = _p32(a); b, pb = _p32(b)
    return int(lib().orc_intersect_num_bound(pa, len(a), pb, len(b), up))


def intersect_set(a, b):
    a, pa = _p32(a); b, pb = _p32(b)
    out = np.zeros(max(1, min(len(a), len(b))), np.int32)
    k = lib().orc_intersect_set(pa, len(a), pb, len(b), out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out[:k].copy()


def difference_num(a, b, up, vid_b):
    a, pa = _p32(a); b, pb = _p32(b)
    return int(lib().orc_difference_num(pa, len(a), pb, len(b), up, vid_b))


def difference_num_noexcl(a, b, up):
    a, pa = _p32(a); b, pb = _p32(b)
    return int(lib().orc_difference_num_noexcl(pa, len(a), pb, len(b), up))


def orient(rowptr, colidx):
    rp, prp = _p64(rowptr); ci, pci = _p32(colidx)
    n = len(rp) - 1
    orp = np.zeros(n + 1, np.int64)
    oci = np.zeros(max(1, len(ci) // 2), np.int32)
    m = lib().orc_orient(n, prp, pci, orp.ctypes.data_as(C.POINTER(C.c_int64)),
                         oci.ctypes.data_as(C.POINTER(C.c_int32)))
    return orp, oci[:m].copy()


_PATTERN_FN = {"triangle": "orc_tc", "clique4": "orc_clique4", "clique5": "orc_clique5",
               "diamond": "orc_diamond", "rectangle": "orc_cycle4"}


def count(pattern: str, rowptr, colidx, v_begin=0, v_end=-1, threads=0):
    """Count `pattern` with the restatement; returns (count, algorithmic_bytes).

    triangle / clique4 / clique5 expect a DAG CSR; diamond / rectangle the
    undirected CSR."""
    rp, prp = _p64(rowptr); ci, pci = _p32(colidx)
    b = C.c_uint64()
    fn = getattr(lib(), _PATTERN_FN[pattern])
    c = fn(len(rp) - 1, prp, pci, v_begin, v_end, threads, C.byref(b))
    return int(c), int(b.value)


def count_tasks(pattern: str, rowptr, colidx, tasks, threads=0):
    """count() over an explicit list of vertex tasks (the vertex loop visits
    exactly these vertices, e.g. a sample of them); returns (count,
    algorithmic_bytes) of those tasks."""
    t = np.ascontiguousarray(tasks, dtype=np.int32)
    L = lib()
    L.orc_set_tasks.restype = None
    L.orc_set_tasks.argtypes = [C.POINTER(C.c_int32), C.c_int64]
    if len(t) == 0:
        return 0, 0
    L.orc_set_tasks(t.ctypes.data_as(C.POINTER(C.c_int32)), len(t))
    try:
        return count(pattern, rowptr, colidx, 0, -1, threads)
    finally:
        L.orc_set_tasks(None, 0)


def cycle4_wedge(rowptr, colidx, threads=0):
    """Edge-induced 4-cycles of an undirected CSR by wedge aggregation (a
    second CPU algorithm for large golden counts; pinned against
    count("rectangle") in tests/test_oracle.py)."""
    rp, prp = _p64(rowptr); ci, pci = _p32(colidx)
    fn = lib().orc_cycle4_wedge
    fn.restype = C.c_uint64
    fn.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int]
    return int(fn(len(rp) - 1, prp, pci, threads))


def clique_bitset(rowptr, colidx, k, v_begin=0, v_end=-1, threads=0):
    """k-cliques (4 or 5) of a DAG CSR by bitset sub-DAGs (a second CPU
    algorithm for large golden counts; pinned against count("clique4/5"))."""
    rp, prp = _p64(rowptr); ci, pci = _p32(colidx)
    fn = lib().orc_clique_bitset
    fn.restype = C.c_uint64
    fn.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int, C.c_int64, C.c_int64, C.c_int]
    return int(fn(len(rp) - 1, prp, pci, k, v_begin, v_end, threads))


MOTIF_NAMES = {3: ("wedges", "triangles"),
               4: ("3-stars", "4-paths", "tailed-triangles", "4-cycles", "diamonds", "4-cliques")}


def motif(rowptr, colidx, k, threads=0):
    """Vertex-induced k-motif census (k = 3 or 4) of an undirected CSR by the
    reference's formula path; returns the counts in the reference's order."""
    rp, prp = _p64(rowptr); ci, pci = _p32(colidx)
    out = (C.c_uint64 * 6)()
    fn = lib().orc_motif
    fn.restype = C.c_int
    fn.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int, C.c_int, C.POINTER(C.c_uint64)]
    if fn(len(rp) - 1, prp, pci, k, threads, out):
        raise ValueError("k must be 3 or 4")
    return [int(out[i]) for i in range(len(MOTIF_NAMES[k]))]


def ref_motif_timed(g, k, threads=0, timeout=3600):
    """The reference's motif_omp_formula binary on an undirected graph:
    (census, seconds of its counting loop)."""
    env = dict(os.environ)
    if threads > 0:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory(prefix="xgm_ref_") as d:
        prefix = os.path.join(d, "graph")
        g.save(prefix)
        out = subprocess.run([os.path.join(REF_DIR, "motif_omp_formula"), prefix, str(k)], capture_output=True,
                             text=True, timeout=timeout, env=env, check=True).stdout
    got = dict((int(a), int(b)) for a, b in re.findall(r"pattern (\d+): (\d+)", out))
    t = re.search(r"runtime[^=]*= ([0-9.eE+-]+)", out)
    return [got[i] for i in range(len(MOTIF_NAMES[k]))], (float(t.group(1)) if t else float("nan"))


def ref_motif_graph(g, k, threads=0, timeout=3600):
    """The reference's motif_omp_formula binary on an undirected graph."""
    return ref_motif_timed(g, k, threads, timeout)[0]


# ---- the reference's own binaries (oracle/_ref) ----
_REF_CMD = {
    "triangle": ("tc_omp_base", [], r"total_num_triangles = (\d+)"),
    "clique4": ("clique_omp_base", ["4"], r"num_4-cliques = (\d+)"),
    "clique5": ("clique_omp_base", ["5"], r"num_5-cliques = (\d+)"),
    "diamond": ("sgl_omp_base", ["diamond"], r"total_num = (\d+)"),
    "rectangle": ("sgl_omp_base", ["rectangle"], r"total_num = (\d+)"),
}


def ref_available(pattern="triangle") -> bool:
    return os.access(os.path.join(REF_DIR, _REF_CMD[pattern][0]), os.X_OK)


def ref_count(pattern: str, prefix: str, threads: int = 0, timeout: float = 3600):
    """Run the reference GraphMiner binary on an undirected CSR saved at
    `prefix` (meta/vertex/edge files).  The binary orients internally for
    triangle/clique.  Returns (count, seconds of its counting loop)."""
    exe, extra, pat = _REF_CMD[pattern]
    env = dict(os.environ)
    if threads > 0:
        env["OMP_NUM_THREADS"] = str(threads)
    out = subprocess.run([os.path.join(REF_DIR, exe), prefix] + extra, capture_output=True,
                         text=True, timeout=timeout, env=env, check=True).stdout
    m = re.search(pat, out)
    if not m:
        raise RuntimeError(f"reference output not understood:\n{out[-2000:]}")
    t = re.search(r"runtime[^=]*= ([0-9.eE+-]+)", out)
    return int(m.group(1)), (float(t.group(1)) if t else float("nan"))


def ref_count_graph(pattern: str, g, threads: int = 0, timeout: float = 3600):
    """Save an undirected xgminer_amd.Graph to a temp dir and run ref_count."""
    with tempfile.TemporaryDirectory(prefix="xgm_ref_") as d:
        prefix = os.path.join(d, "graph")
        g.save(prefix)
        return ref_count(pattern, prefix, threads, timeout)
