"""GPU parity: every count from the HIP path, called through the C ABI, is
bit-exact against the CPU oracle and the reference's golden counts."""
import json
import math
import os
import sys

import pytest

from conftest import CITESEER, GOLDEN, complete_graph, random_graph

pytestmark = pytest.mark.gpu

DAG_PATTERNS = {"triangle": 1}
STRATS = {"hash": 1, "bsearch": 2, "bucket": 7, "merge": 8}


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(xgm):
    c = xgm.Context(0)
    yield c
    c.close()


def _gold():
    return json.load(open(os.path.join(GOLDEN, "counts.json")))


@pytest.mark.parametrize("strat", list(STRATS))
def test_tc_citeseer(xgm, ctx, strat):
    _say("tc citeseer", strat)
    d = xgm.load_csr(CITESEER).orient()
    assert ctx.count(d, xgm.TRIANGLE, STRATS[strat]) == 1166


@pytest.mark.parametrize("n", [3, 5, 40, 70, 300, 2100])
def test_tc_complete(xgm, ctx, n):
    # n-1 out-degree spans every tier: 64-wave, 256-wave, 2048-block, >2048 global
    _say("tc complete", n)
    d = complete_graph(xgm, n).orient()
    for s in STRATS.values():
        assert ctx.count(d, xgm.TRIANGLE, s) == math.comb(n, 3)


def test_tc_empty_and_tiny(xgm, ctx):
    _say("tc empty/tiny")
    import numpy as np
    g = xgm.from_edges(10, np.zeros(0, np.int32), np.zeros(0, np.int32)).orient()
    assert ctx.count(g, xgm.TRIANGLE) == 0
    g = xgm.from_edges(3, [0, 1], [1, 2]).orient()
    assert ctx.count(g, xgm.TRIANGLE) == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tc_random_vs_oracle(xgm, orc, ctx, seed):
    _say("tc random", seed)
    g = random_graph(xgm, 3000, 60000, seed).orient()
    want = orc.count("triangle", g.rowptr, g.colidx)[0]
    for s in STRATS.values():
        assert ctx.count(g, xgm.TRIANGLE, s) == want


@pytest.mark.parametrize("name", ["rmat10_s7", "rmat12_s3", "er2000_d16_s5", "chunglu20k_s9", "rmat16_s1"])
def test_tc_golden(xgm, ctx, name):
    from golden.make_golden import SMALL, make_graph
    _say("tc golden", name)
    gold = _gold()[name]
    g = make_graph(xgm, SMALL[name][0], SMALL[name][1])
    assert f"{g.hash():016x}" == gold["hash"]
    d = g.orient()
    dg = ctx.upload(d)
    for s in STRATS.values():
        assert ctx.count(dg, xgm.TRIANGLE, s) == gold["counts"]["triangle"]
    dg.close()


def test_tc_solver_entry(xgm):
    _say("TCSolver entry")
    assert xgm.TCSolver(xgm.load_csr(CITESEER).orient()) == 1166


def test_profile_reports_kernels(xgm, ctx):
    _say("profile")
    d = complete_graph(xgm, 100).orient()
    ctx.count(d, xgm.TRIANGLE)
    total, ks = ctx.last_profile()
    assert ks and all(k.ms >= 0 for k in ks) and total >= 0


# ---------------------------------------------------------------- k-cliques

CLQ = {"clique4": (2, 4), "clique5": (3, 5)}
CLQ_STRATS = {"auto": 0, "lut": 3, "list": 4, "elut": 5, "lut_rank": 6}


@pytest.mark.parametrize("n", [4, 6, 40, 100, 300, 600])
def test_clique_complete_lut_tiers(xgm, ctx, n):
    # out-degrees n-1 exercise the 64 / 256 / 512 / 1024-row LUT tiers
    _say("clique complete", n)
    d = complete_graph(xgm, n).orient()
    dg = ctx.upload(d)
    assert ctx.count(dg, 2) == math.comb(n, 4)
    assert ctx.count(dg, 3) == math.comb(n, 5)
    dg.close()


def test_clique4_list_above_lut_threshold(xgm, ctx):
    # d+ up to 1039 > 1024: those vertices take the list path under auto
    _say("clique4 complete 1040")
    d = complete_graph(xgm, 1040).orient()
    assert ctx.count(d, 2) == math.comb(1040, 4)


@pytest.mark.parametrize("strat", list(CLQ_STRATS))
@pytest.mark.parametrize("seed", [4, 5])
def test_clique_random_vs_oracle(xgm, orc, ctx, strat, seed):
    _say("clique random", strat, seed)
    g = random_graph(xgm, 2000, 40000, seed).orient()
    dg = ctx.upload(g)
    for name, (pid, _) in CLQ.items():
        want = orc.count(name, g.rowptr, g.colidx)[0]
        assert ctx.count(dg, pid, CLQ_STRATS[strat]) == want, name
    dg.close()


def test_clique_lut_threshold_option(xgm, ctx):
    _say("clique lut_max_degree sweep")
    g = _gold()["rmat12_s3"]
    from golden.make_golden import SMALL, make_graph
    d = make_graph(xgm, SMALL["rmat12_s3"][0], SMALL["rmat12_s3"][1]).orient()
    dg = ctx.upload(d)
    for lm in (16, 64, 200, 1024):
        assert ctx.count(dg, 2, 0, lm) == g["counts"]["clique4"]
        assert ctx.count(dg, 3, 0, lm) == g["counts"]["clique5"]
    dg.close()


@pytest.mark.parametrize("name", ["rmat10_s7", "rmat12_s3", "er2000_d16_s5", "chunglu20k_s9", "rmat16_s1"])
def test_clique_golden(xgm, ctx, name):
    from golden.make_golden import SMALL, make_graph
    _say("clique golden", name)
    gold = _gold()[name]
    d = make_graph(xgm, SMALL[name][0], SMALL[name][1]).orient()
    dg = ctx.upload(d)
    for pname, (pid, _) in CLQ.items():
        assert ctx.count(dg, pid) == gold["counts"][pname], pname
    if name != "rmat16_s1":
        for pname, (pid, _) in CLQ.items():
            assert ctx.count(dg, pid, 4) == gold["counts"][pname], pname + " list"
    dg.close()


def test_clique_citeseer(xgm, ctx):
    _say("clique citeseer")
    d = xgm.load_csr(CITESEER).orient()
    assert xgm.CliqueSolver(d, 4) == 255
    assert xgm.CliqueSolver(d, 5) == 46


# ------------------------------------------------- edge-induced diamond / 4-cycle

SGL = {"diamond": 4, "rectangle": 5}


@pytest.mark.parametrize("n", [4, 5, 30, 70, 1100])
def test_sgl_complete(xgm, ctx, n):
    # degrees n-1 cover the 64 / 1024 / 8192 diamond tiers
    _say("sgl complete", n)
    g = complete_graph(xgm, n)
    dg = ctx.upload(g)
    assert ctx.count(dg, 4) == 6 * math.comb(n, 4)
    assert ctx.count(dg, 5) == 3 * math.comb(n, 4)
    if n <= 70:
        assert ctx.count(dg, 5, 4) == 3 * math.comb(n, 4)  # reference loop order (PAIR)
    dg.close()


@pytest.mark.parametrize("a,b", [(2, 2), (3, 7), (60, 200)])  # 60x200: dense-counter tier
def test_sgl_bipartite(xgm, ctx, a, b):
    from conftest import complete_bipartite
    _say("sgl bipartite", a, b)
    g = complete_bipartite(xgm, a, b)
    assert ctx.count(g, 4) == 0
    assert ctx.count(g, 5) == math.comb(a, 2) * math.comb(b, 2)
    assert ctx.count(g, 5, 4) == math.comb(a, 2) * math.comb(b, 2)


def test_sgl_hub_star_plus_clique(xgm, ctx, orc):
    # one vertex of degree > 8192 (diamond tier with the set in HBM)
    import numpy as np
    _say("sgl hub")
    n = 9000
    src = np.concatenate([np.zeros(n - 1, np.int64), np.arange(1, 200)])
    dst = np.concatenate([np.arange(1, n), np.arange(2, 201)])
    g = xgm.from_edges(n, src, dst)
    for name, pid in SGL.items():
        want = orc.count(name, g.rowptr, g.colidx)[0]
        assert ctx.count(g, pid) == want, name


def test_sgl_binned_hub_sets(xgm, ctx, orc):
    """Owners above 8192 neighbours search N(t) as a binned LDS set (2^14 bins
    of 2^6 ids for 2^20 vertices): hubs of 50,000 random neighbours, of 20,000
    neighbours packed into 59,000 ids (about 22 per bin), and of 60,000 (above
    the binned capacity: the sample index over HBM), with a community among
    the hub neighbours so partners share neighbours with the hub; diamonds,
    4-cycles and the 4-motif census (the same kernel in motif mode) against
    the oracle."""
    import numpy as np
    _say("sgl binned hubs")
    rng = np.random.default_rng(11)
    n = 1 << 20
    h0 = rng.choice(np.arange(3, n), 50000, replace=False)
    h1 = rng.choice(np.arange(1000, 60000), 20000, replace=False)
    h2 = rng.choice(np.arange(3, n), 60000, replace=False)
    comm = np.concatenate([h0[:2000], h1[:2000], h2[:1000]])
    src = np.concatenate([rng.integers(0, n, 300000), np.zeros(50000, np.int64), np.ones(20000, np.int64),
                          np.full(60000, 2), rng.choice(comm, 40000)])
    dst = np.concatenate([rng.integers(0, n, 300000), h0, h1, h2, rng.choice(comm, 40000)])
    g = xgm.from_edges(n, src, dst)
    dg = ctx.upload(g)
    for name, pid in SGL.items():
        assert ctx.count(dg, pid) == orc.count(name, g.rowptr, g.colidx)[0], name
    assert ctx.motifs(dg, 4)[0].tolist() == list(orc.motif(g.rowptr, g.colidx, 4))
    dg.close()


@pytest.mark.parametrize("seed", [6, 7])
def test_sgl_random_vs_oracle(xgm, orc, ctx, seed):
    _say("sgl random", seed)
    g = random_graph(xgm, 3000, 40000, seed)
    dg = ctx.upload(g)
    for name, pid in SGL.items():
        want = orc.count(name, g.rowptr, g.colidx)[0]
        assert ctx.count(dg, pid) == want, name
        assert ctx.count(dg, pid) == want, name + " (repeat: dense counters left zero)"
    assert ctx.count(dg, 5, 4) == orc.count("rectangle", g.rowptr, g.colidx)[0]
    dg.close()


@pytest.mark.parametrize("name", ["rmat10_s7", "rmat12_s3", "er2000_d16_s5", "chunglu20k_s9", "rmat16_s1"])
def test_sgl_golden(xgm, ctx, name):
    from golden.make_golden import SMALL, make_graph
    _say("sgl golden", name)
    gold = _gold()[name]
    g = make_graph(xgm, SMALL[name][0], SMALL[name][1])
    dg = ctx.upload(g)
    for pname, pid in SGL.items():
        assert ctx.count(dg, pid) == gold["counts"][pname], pname
    if name in ("rmat10_s7", "er2000_d16_s5", "chunglu20k_s9"):
        assert ctx.count(dg, 5, 4) == gold["counts"]["rectangle"], "pair"
    dg.close()


def test_sgl_citeseer(xgm, ctx):
    _say("sgl citeseer")
    g = xgm.load_csr(CITESEER)
    assert xgm.SglSolver(g, "diamond") == 3730
    assert xgm.SglSolver(g, "rectangle") == 6059


# ------------------------------------------- the reference's own drivers on libxgm

_REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref")


@pytest.mark.skipif(not os.path.exists(os.path.join(_REF, "triangle_xgm")),
                    reason="GraphMiner drivers linked with libxgm not built (needs the reference tree)")
@pytest.mark.parametrize("exe,args,pattern,want", [
    ("triangle_xgm", [], r"total_num_triangles = (\d+)", 1166),
    ("clique_xgm", ["4"], r"num_4-cliques = (\d+)", 255),
    ("clique_xgm", ["5"], r"num_5-cliques = (\d+)", 46),
    ("sgl_xgm", ["diamond"], r"total_num = (\d+)", 3730),
    ("sgl_xgm", ["rectangle"], r"total_num = (\d+)", 6059),
])
def test_graphminer_drivers_linked_to_libxgm(exe, args, pattern, want):
    """GraphMiner's main.cc/graph.cc (orientation, file loading) with the
    solver entry points resolved to libxgm (x-gminer_amd/integration)."""
    import re
    import subprocess
    _say("graphminer driver", exe, *args)
    out = subprocess.run([os.path.join(_REF, exe), CITESEER] + args, capture_output=True, text=True,
                         timeout=120, check=True).stdout
    assert int(re.search(pattern, out).group(1)) == want


def test_lut_rank_rejects_foreign_orientation(xgm, ctx):
    """An id-oriented DAG (u -> v iff v > u) is acyclic but not in (degree, id)
    order: the rank LUT must refuse it rather than miscount; LUT accepts it."""
    import numpy as np
    _say("lut_rank foreign orientation")
    g = random_graph(xgm, 300, 4000, 9)
    src = np.repeat(np.arange(g.n), g.degree())
    keep = g.colidx > src
    rp = np.concatenate([[0], np.cumsum(np.bincount(src[keep], minlength=g.n))]).astype(np.int64)
    d = xgm.Graph.from_arrays(rp, g.colidx[keep], xgm.CSR_SORTED | xgm.CSR_DAG)
    import oracle
    want = oracle.count("clique4", d.rowptr, d.colidx)[0]
    assert ctx.count(d, 2, 3) == want
    with pytest.raises(xgm.XGMError):
        ctx.count(d, 2, 6)


@pytest.mark.parametrize("pattern,name", [(2, "clique4"), (3, "clique5")])
def test_schedules_every_code_and_mixtures_exact(xgm, ctx, pattern, name):
    """Every per-vertex strategy code, and random mixtures of codes, count the
    same: the scheduler cannot change a count."""
    import numpy as np
    from golden.make_golden import SMALL, make_graph
    _say("schedules", name)
    gold = _gold()["rmat12_s3"]["counts"][name]
    d = make_graph(xgm, SMALL["rmat12_s3"][0], SMALL["rmat12_s3"][1]).orient()
    dg = ctx.upload(d)
    task = np.diff(d.rowptr) >= (pattern + 1)  # d+ >= k - 1
    for c in range(5):
        s = xgm.Schedule(dg, pattern, np.where(task, c, -1))
        n, us = s.count(log=True)
        assert n == gold, c
        assert np.all(us >= 0) and us[task].sum() > 0
        s.close()
    rng = np.random.default_rng(pattern)
    for _ in range(3):
        s = xgm.Schedule(dg, pattern, np.where(task, rng.integers(0, 5, d.n), -1))
        assert s.count() == gold
        s.close()
    heur = dg.sched_heuristic(pattern)
    assert xgm.Schedule(dg, pattern, heur).count() == gold
    dg.close()


def _hub_graph(xgm, nb=1100, nl=1200, inner=10, seed=3):
    """v (id 0) -> B (nb vertices) <- L (nl vertices), B x L complete, B
    sparse inside: v and every L vertex get out-degree nb > 1024 under the
    degree orientation, with sparse sub-DAGs."""
    import numpy as np
    rng = np.random.default_rng(seed)
    B = 1 + np.arange(nb)
    L = 1 + nb + np.arange(nl)
    src = [np.zeros(nb, np.int64), np.repeat(B, nl), np.repeat(B, inner)]
    dst = [B, np.tile(L, nb), rng.choice(B, nb * inner)]
    return xgm.from_edges(1 + nb + nl, np.concatenate(src), np.concatenate(dst))


@pytest.mark.parametrize("nb,nl", [(1100, 1200), (1500, 1500)])
def test_clique_hubs_hbm_lut(xgm, orc, ctx, nb, nl):
    """Hub tasks above the 1024-row LDS LUT: out-degree 1100 runs the packed
    triangular LDS LUT (<= 1344 rows), 1500 the LUT with rows in HBM."""
    import numpy as np
    _say("clique hubs lut", nb, nl)
    d = _hub_graph(xgm, nb=nb, nl=nl).orient()
    assert int(np.diff(d.rowptr).max()) == nb
    want4 = orc.count("clique4", d.rowptr, d.colidx)[0]
    want5 = orc.count("clique5", d.rowptr, d.colidx)[0]
    dg = ctx.upload(d)
    for strat in (0, 6, 5):
        assert ctx.count(dg, 2, strat) == want4, strat
        assert ctx.count(dg, 3, strat) == want5, strat
    task = np.diff(d.rowptr) >= 4
    for c in (1, 4):
        s = xgm.Schedule(dg, 3, np.where(task, c, -1))
        assert s.count() == want5
        s.close()
    dg.close()


def test_clique_hub_pieces(xgm, orc, ctx):
    """Two hub tasks (out-degree 1100) and nothing else in the triangular LDS
    LUT bin: each is split into 64 row pieces over as many blocks, and the
    pieces' counts must add up to the oracle's."""
    import numpy as np
    _say("clique hub pieces")
    rng = np.random.default_rng(5)
    nb, nleaf = 1100, 1100
    B = 2 + np.arange(nb)                      # two hubs 0, 1 point to every B vertex
    leaves = 2 + nb + np.arange(nb * nleaf)    # each B vertex gets its own leaves, so deg(B) > deg(hub)
    src = np.concatenate([np.zeros(nb, np.int64), np.ones(nb, np.int64), np.repeat(B, nleaf), np.repeat(B, 40)])
    dst = np.concatenate([B, B, leaves, rng.choice(B, nb * 40)])
    d = xgm.from_edges(2 + nb + nb * nleaf, src, dst).orient()
    dout = np.diff(d.rowptr)
    assert int(((dout > 1024) & (dout <= 1344)).sum()) == 2
    want4 = orc.count("clique4", d.rowptr, d.colidx)[0]
    want5 = orc.count("clique5", d.rowptr, d.colidx)[0]
    assert want5 > 0
    dg = ctx.upload(d)
    for strat in (0, 6):
        assert ctx.count(dg, 2, strat) == want4, strat
        assert ctx.count(dg, 3, strat) == want5, strat
    task = dout >= 4
    for c in (1, 4):
        s = xgm.Schedule(dg, 3, np.where(task, c, -1))
        assert s.count() == want5, c
        s.close()
    dg.close()


def test_clique_edge_hub_pairs_merge_path(xgm, orc, ctx):
    """Edge tasks between two vertices of out-degree > 1024 take the merge
    path for W1 = N+(v0) ∩ N+(v1) (the hybrid rule): the L vertices of the
    hub graph, all pointing to the 1100 B vertices, get a sparse ring among
    themselves, so DAG edges L_i -> L_j join two lists of 1100+ ids.  The
    edge-task strategies and schedule codes must match the oracle."""
    import numpy as np
    _say("clique edge hub pairs merge path")
    nb, nl = 1100, 1200
    g = _hub_graph(xgm, nb=nb, nl=nl)
    L = 1 + nb + np.arange(nl)
    src = np.concatenate([np.repeat(np.arange(g.n), np.diff(g.rowptr)), L, L[::5]])
    dst = np.concatenate([g.colidx, np.roll(L, 1), np.roll(L, 7)[::5]])
    d = xgm.from_edges(g.n, src, dst).orient()
    dout = np.diff(d.rowptr)
    both = (dout[np.repeat(np.arange(d.n), dout)] > 1024) & (dout[d.colidx] > 1024)
    assert both.sum() > 1000  # hub x hub DAG edges exist
    want4 = orc.count("clique4", d.rowptr, d.colidx)[0]
    want5 = orc.count("clique5", d.rowptr, d.colidx)[0]
    dg = ctx.upload(d)
    for strat in (xgm.STRAT_LIST, xgm.STRAT_ELUT):
        assert ctx.count(dg, 2, strat) == want4, strat
        assert ctx.count(dg, 3, strat) == want5, strat
    task = dout >= 4
    for c in (2, 3):
        s = xgm.Schedule(dg, 3, np.where(task, c, -1))
        assert s.count() == want5, c
        s.close()
    dg.close()


@pytest.mark.parametrize("n,spread", [(40000, 1), (20000, 1), (20000, 97)])
def test_sgl_sampled_hub_owner(xgm, ctx, orc, n, spread):
    """An owner of degree n - 1: above the 36,352-entry LDS set the probes go
    through the 1024-sample index over HBM, below it to the sorted set in
    LDS (spread 97 puts the ring on ids spaced 97 apart).  Diamonds and the
    motif census."""
    import numpy as np
    _say("sgl hub owner", n, spread)
    ring = np.arange(1, 3001)
    src = np.concatenate([np.zeros(n - 1, np.int64), ring, ring[::3]])
    dst = np.concatenate([np.arange(1, n), ring % 3000 + 1, (ring[::3] + 7) % 3000 + 1])
    if spread > 1:  # ring vertex r -> id r * spread (mod n - 1) + 1, a bijection on 1..n-1
        perm = np.arange(n)
        perm[1:] = (np.arange(n - 1) * spread) % (n - 1) + 1
        src, dst = perm[src], perm[dst]
    g = xgm.from_edges(n, src, dst)
    assert int(np.diff(g.rowptr).max()) == n - 1
    assert ctx.count(g, 4) == orc.count("diamond", g.rowptr, g.colidx)[0]
    dg = ctx.upload(g)
    assert ctx.motifs(dg, 4)[0].tolist() == orc.motif(g.rowptr, g.colidx, 4)
    dg.close()


def test_hash_bucket_overflow(xgm, orc, ctx):
    """Neighbourhoods crowded into one LDS hash bucket: 120 of the 300 graph
    vertices share the home bucket of every bucketed hash here (the same top
    11 bits of id * 0x9E3779B1; the tables use 5 to 11 of them), so staged
    sets overflow their home bucket by many buckets, and probes continue past
    full buckets (the rare branch of the triangle, diamond, rank-LUT and
    packed-LUT probes) for hits and misses alike."""
    import numpy as np
    _say("hash bucket overflow")
    n = 300_000
    ids = np.arange(n, dtype=np.uint64)
    home = ((ids * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)) >> np.uint64(21)
    counts = np.bincount(home.astype(np.int64), minlength=2048)
    crowd = np.nonzero(home == np.uint64(int(np.argmax(counts))))[0][:120]
    assert len(crowd) == 120
    rng = np.random.default_rng(11)
    rest = rng.choice(np.setdiff1d(np.arange(n), crowd), 180, replace=False)
    verts = np.concatenate([crowd, rest])
    iu, ju = np.triu_indices(len(verts), 1)
    keep = rng.random(len(iu)) < 0.35
    g = xgm.from_edges(n, verts[iu[keep]], verts[ju[keep]])
    dg = ctx.upload(g)
    for name, pid in SGL.items():
        assert ctx.count(dg, pid) == orc.count(name, g.rowptr, g.colidx)[0], name
    dg.close()
    d = g.orient()
    assert int(np.diff(d.rowptr).max()) > 64  # block-tier triangle items, 256-row LUT bin
    dd = ctx.upload(d)
    assert ctx.count(dd, xgm.TRIANGLE) == orc.count("triangle", d.rowptr, d.colidx)[0]
    for name, (pid, _) in CLQ.items():
        want = orc.count(name, d.rowptr, d.colidx)[0]
        for strat in (0, xgm.STRAT_LUT, xgm.STRAT_LUT_RANK):
            assert ctx.count(dd, pid, strat) == want, (name, strat)
    dd.close()
