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
