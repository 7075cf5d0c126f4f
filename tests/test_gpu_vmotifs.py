"""Vertex-induced 4-motifs (GLUMIN P1 3-star, P2 diamond, P3 tailed
triangle) through xgm_count, under the LUT (AND-NOT row popcounts) and the
set-difference list strategies: the reference's citeseer counts
(codes/GraphMiner/src/motif/README.md:53-56), the oracle's census (the
reference's formula path restated, pinned in test_oracle.py) on seeded
graphs whose degrees reach every LUT tier and the list path, complete and
complete bipartite closed forms, and per-instance batches."""
import math
import sys

import numpy as np
import pytest

from conftest import CITESEER, complete_bipartite, complete_graph, random_graph

pytestmark = pytest.mark.gpu

# census order of oracle.motif(k=4): 3-stars, 4-paths, tailed, 4-cycles, diamonds, 4-cliques
VM = {"star": (6, 0), "diamond": (7, 4), "tailed": (8, 2)}
# auto: LUT kernels, or the census identities when a task lies above 1024
# neighbours; lut: LUT kernels with the set-difference list path above 1024
STRATS = {"auto": 0, "lut": 3, "list": 4}


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(xgm):
    c = xgm.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("strat", list(STRATS))
def test_vmotif_citeseer_known_answers(xgm, ctx, strat):
    _say("vmotif citeseer", strat)
    dg = ctx.upload(xgm.load_csr(CITESEER))
    want = {"star": 222630, "diamond": 2200, "tailed": 22900}
    for name, (pid, _) in VM.items():
        assert ctx.count(dg, pid, STRATS[strat]) == want[name], name
    dg.close()


def _hub_graph(xgm, seed, hub_deg):
    """Random background plus hubs of the given degrees (LUT tiers 64 / 256 /
    1024 and, above 1024, the list path)."""
    rng = np.random.default_rng(seed)
    n = 3000
    src, dst = [rng.integers(0, n, 12000)], [rng.integers(0, n, 12000)]
    for h, k in enumerate(hub_deg):
        src.append(np.full(k, h))
        dst.append(rng.choice(np.arange(len(hub_deg), n), k, replace=False))
    return xgm.from_edges(n, np.concatenate(src), np.concatenate(dst))


@pytest.mark.parametrize("seed", [1, 2])
def test_vmotif_vs_oracle_census(xgm, orc, ctx, seed):
    _say("vmotif vs oracle", seed)
    g = _hub_graph(xgm, 70 + seed, [40, 200, 900, 1300])
    census = orc.motif(g.rowptr, g.colidx, 4)
    assert np.diff(g.rowptr).max() > 1024
    dg = ctx.upload(g)
    for strat in STRATS.values():
        for name, (pid, ci) in VM.items():
            assert ctx.count(dg, pid, strat) == census[ci], (name, strat)
    dg.close()


def test_vmotif_rmat_golden(xgm, orc, ctx):
    from golden.make_golden import SMALL, make_graph
    _say("vmotif rmat10")
    g = make_graph(xgm, SMALL["rmat10_s7"][0], SMALL["rmat10_s7"][1])
    census = orc.motif(g.rowptr, g.colidx, 4)
    dg = ctx.upload(g)
    for name, (pid, ci) in VM.items():
        assert ctx.count(dg, pid) == census[ci], name
    dg.close()


def test_vmotif_closed_forms(xgm, ctx):
    _say("vmotif closed forms")
    for n in (4, 30, 80):  # K_n: no induced star, diamond or tailed triangle
        dg = ctx.upload(complete_graph(xgm, n))
        assert [ctx.count(dg, pid) for pid, _ in VM.values()] == [0, 0, 0]
        dg.close()
    for a, b in ((3, 5), (40, 70), (2, 300)):  # K_{a,b}: 3-stars a*C(b,3) + b*C(a,3)
        dg = ctx.upload(complete_bipartite(xgm, a, b))
        assert ctx.count(dg, 6) == a * math.comb(b, 3) + b * math.comb(a, 3)
        assert ctx.count(dg, 6, 4) == a * math.comb(b, 3) + b * math.comb(a, 3)
        assert ctx.count(dg, 7) == 0 and ctx.count(dg, 8) == 0
        dg.close()
    # a graph with no vertex of degree 3 or more
    dg = ctx.upload(xgm.from_edges(6, [0, 1, 2, 3], [1, 2, 3, 4]))
    assert [ctx.count(dg, pid) for pid, _ in VM.values()] == [0, 0, 0]
    dg.close()


def test_vmotif_batch_instances(xgm, orc, ctx):
    _say("vmotif batch")
    inst = [random_graph(xgm, 300 + 50 * i, 2500 + 300 * i, 80 + i) for i in range(5)]
    inst.append(_hub_graph(xgm, 90, [70, 1100]))
    dg = ctx.upload_batch(inst)
    for name, (pid, ci) in VM.items():
        got = ctx.count_instances(dg, pid)
        want = [orc.motif(g.rowptr, g.colidx, 4)[ci] for g in inst]
        assert got.tolist() == want, name
    dg.close()


def test_vmotif_rejects_dag_and_solver_entry(xgm, ctx):
    _say("vmotif errors + PatternSolver")
    d = random_graph(xgm, 200, 1000, 3).orient()
    with pytest.raises(xgm.XGMError):
        ctx.count(d, 6)
    g = xgm.load_csr(CITESEER)
    assert [xgm.PatternSolver(g, k) for k in (1, 2, 3, 4)] == [222630, 2200, 22900, 255]


def test_vmotif_auto_route(xgm, orc, ctx):
    """XGM_STRAT_AUTO counts graphs whose tasks fit the LUT tiers with the LUT
    kernels, and graphs with a task above 1024 neighbours by the census
    identities; both equal the kernels' counts and the oracle's census."""
    _say("vmotif auto route")
    small = ctx.upload(xgm.load_csr(CITESEER))  # largest degree 99
    for name, (pid, _) in VM.items():
        ctx.count(small, pid)
        assert any(k.name.startswith("vm_") for k in ctx.last_profile()[1]), name
    small.close()
    g = _hub_graph(xgm, 75, [50, 300, 1500, 2500])
    census = orc.motif(g.rowptr, g.colidx, 4)
    dg = ctx.upload(g)
    for name, (pid, ci) in VM.items():
        auto = ctx.count(dg, pid)
        assert not any(k.name.startswith("vm_") for k in ctx.last_profile()[1]), name
        assert auto == ctx.count(dg, pid, STRATS["lut"]) == census[ci], name
    dg.close()


def test_vmotif_diamond_bucket_prefilter(xgm, orc, ctx):
    """Induced diamonds with X-GMiner's bucket prefilter (XGM_STRAT_BUCKET,
    the reference's P2 kernel): citeseer's published 2,200, hub graphs and
    batches against the oracle's census; other motifs and degrees >= 65536
    are refused."""
    _say("vmotif bucket")
    bucket = 7
    dg = ctx.upload(xgm.load_csr(CITESEER))
    assert ctx.count(dg, 7, bucket) == 2200
    with pytest.raises(xgm.XGMError):
        ctx.count(dg, 6, bucket)
    dg.close()
    for seed in (3, 4):
        g = _hub_graph(xgm, 60 + seed, [40, 200, 900, 1300, 2500])
        dg = ctx.upload(g)
        assert ctx.count(dg, 7, bucket) == orc.motif(g.rowptr, g.colidx, 4)[4], seed
        dg.close()
    inst = [random_graph(xgm, 300 + 50 * i, 2500 + 300 * i, 120 + i) for i in range(4)]
    inst.append(_hub_graph(xgm, 91, [70, 1100]))
    dg = ctx.upload_batch(inst)
    assert ctx.count_instances(dg, 7, strategy=bucket).tolist() == [orc.motif(g.rowptr, g.colidx, 4)[4] for g in inst]
    dg.close()
    leaves = np.arange(1, 70001, dtype=np.int32)
    dg = ctx.upload(xgm.from_edges(70001, np.zeros(70000, np.int32), leaves))
    with pytest.raises(xgm.XGMError):
        ctx.count(dg, 7, bucket)
    dg.close()
