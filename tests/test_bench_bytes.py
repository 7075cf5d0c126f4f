"""bench.py's algorithmic-byte accounting, on the CPU: the per-kernel split
of triangle-counting bytes follows the library's default tier bounds
(x-gminer_amd/csrc/kernels.h), and its total equals both the oracle's
instrumented op stream and the closed form 4 * sum_v d+(v) * d(v) + 40 * m
(every DAG edge u->v reads both out-lists once, 4 * (d+(u) + d+(v)) bytes,
plus 40 bytes of task data; GraphMiner src/triangle/omp_base.cc:6-26)."""
import os
import re
import sys

import numpy as np

from conftest import CITESEER, random_graph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _kernels_h_int(name):
    src = open(os.path.join(ROOT, "x-gminer_amd", "csrc", "kernels.h")).read()
    m = re.search(r"constexpr int " + re.escape(name) + r"(\[\w*\])? = ([^;]+);", src)
    assert m, name
    def value(x):  # an integer literal or "a << b"
        parts = [int(p) for p in x.split("<<")]
        return parts[0] << parts[1] if len(parts) == 2 else parts[0]
    body = m.group(2).strip()
    if body.startswith("{"):
        return [value(x) for x in body.strip("{}").split(",")]
    return value(body)


def test_tc_tiers_follow_library_defaults():
    import bench
    caps = _kernels_h_int("kTcWaveCap")
    large = _kernels_h_int("kTcWaveLargeDefault")
    maxdeg = _kernels_h_int("kTierMaxDeg")
    tiers = bench.TC_TIERS
    assert tiers[0][1] == 1
    for (_, _, hi), (_, lo, _) in zip(tiers, tiers[1:]):
        assert lo == hi + 1  # contiguous, no gaps or overlaps
    bounds = {name: (lo, hi) for name, lo, hi in tiers}
    assert bounds["tc_wave_64"] == (1, caps[0])
    assert bounds["tc_block_2048"] == (max(caps[0], large) + 1, maxdeg[2])
    assert bounds["tc_block_big"][0] == maxdeg[2] + 1


def test_tc_bytes_equal_oracle_and_closed_form(xgm, orc):
    import bench
    for g in (xgm.load_csr(CITESEER), random_graph(xgm, 3000, 40000, 7), xgm.gen_rmat(12, 16, 0.57, 0.19, 0.19, 3)):
        d = g.orient()
        per_tier, total = bench.tc_alg_bytes_per_tier(d)
        assert sum(per_tier.values()) == total
        dplus = np.diff(np.asarray(d.rowptr)).astype(np.int64)
        deg = np.diff(np.asarray(g.rowptr)).astype(np.int64)  # undirected degree = d+ + d-
        m = int(dplus.sum())
        assert total == 4 * int((dplus * deg).sum()) + 40 * m
        assert total == orc.count("triangle", d.rowptr, d.colidx)[1]


def _brute_tc_kernel_bytes(d, tiers):
    rp, ci = np.asarray(d.rowptr), np.asarray(d.colidx)
    dout = np.diff(rp)
    out = {name: 0 for name, _, _ in tiers}
    tier = lambda x: next(name for name, lo, hi in tiers if lo <= x <= hi)
    staged = set()
    for u in range(d.n):
        for v in ci[rp[u]:rp[u + 1]]:
            t, p = (u, v) if dout[u] > dout[v] else (v, u)
            out[tier(dout[t])] += 20 + 4 * int(dout[p])
            staged.add(int(t))
    for t in staged:
        out[tier(dout[t])] += 32 + 4 * int(dout[t])
    return out


def test_kernel_byte_models_match_loops(xgm):
    """The per-kernel byte models (read-once bytes of each kernel's own
    algorithm) equal straightforward loops over the graph."""
    import bench
    g = random_graph(xgm, 400, 3000, 11)
    d = g.orient()
    assert bench.tc_kernel_bytes(d) == _brute_tc_kernel_bytes(d, bench.TC_TIERS)
    # cliques: rank-LUT codes everywhere; vertex v0 reads N+(v0) and every
    # member's row pointers, in-list bounds and out-list
    rp, ci = np.asarray(d.rowptr), np.asarray(d.colidx)
    dout = np.diff(rp)
    codes = np.where(dout >= 3, 1, -1).astype(np.int8)
    got = bench.clique_kernel_bytes(d, codes, 4)
    want = 0
    for v in range(d.n):
        if dout[v] < 3:
            continue
        want += 16 + 4 * int(dout[v]) + sum(32 + 4 * int(dout[u]) for u in ci[rp[v]:rp[v + 1]])
    assert sum(got.values()) == want
    assert set(got) <= {"k4_rlut_64", "k4_rlut_256"}
    # diamonds: owner = (degree, id)-larger endpoint of degree >= 3
    grp, gci = np.asarray(g.rowptr), np.asarray(g.colidx)
    deg = np.diff(grp)
    want, owners = 0, set()
    for u in range(g.n):
        for v in gci[grp[u]:grp[u + 1]]:
            if (deg[u], u) > (deg[v], v) and deg[u] >= 3:
                want += 20 + 4 * int(deg[v])
                owners.add(u)
    want += sum(16 + 4 * int(deg[t]) for t in owners)
    assert sum(bench.diamond_kernel_bytes(g).values()) == want
    # 4-cycles: relabel by (degree, id); task x reads its lower list and the
    # part of each lower neighbour's list below x
    key = sorted(range(g.n), key=lambda v: (deg[v], v))
    rank = np.empty(g.n, np.int64)
    rank[key] = np.arange(g.n)
    nl = [sorted(rank[gci[grp[v]:grp[v + 1]]]) for v in key]  # lists in new ids
    want = 0
    for x in range(g.n):
        low = [v for v in nl[x] if v < x]
        if low:
            want += 16 + 4 * len(low)
        for v in low:
            want += 16 + 4 * sum(1 for w in nl[v] if w < x)
    assert bench.cycle4_bytes(g) == want


def test_oracle_task_sample_partitions_counts(xgm, orc):
    """The oracle's explicit task lists (the CPU baseline's samples) visit
    exactly the given vertices: two halves add up to the full count and op
    stream."""
    g = xgm.gen_rmat(11, 16, 0.57, 0.19, 0.19, 5)
    d = g.orient()
    for p, gr in (("triangle", d), ("clique4", d), ("clique5", d), ("diamond", g), ("rectangle", g)):
        full = orc.count(p, gr.rowptr, gr.colidx)
        a = orc.count_tasks(p, gr.rowptr, gr.colidx, np.arange(0, gr.n, 2))
        b = orc.count_tasks(p, gr.rowptr, gr.colidx, np.arange(1, gr.n, 2))
        assert (a[0] + b[0], a[1] + b[1]) == full, p
        assert orc.count(p, gr.rowptr, gr.colidx) == full  # the list is cleared afterwards
