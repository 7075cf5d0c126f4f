"""Batches of instances (disjoint union, one count per instance), the
hipGraph plan and the autotuner, through the C ABI: every per-instance count
is bit-exact against the oracle run on that instance alone, or against a
closed form where the oracle would take minutes (complete bipartite)."""
import json
import math
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, complete_bipartite, complete_graph, random_graph

pytestmark = pytest.mark.gpu

DAG = {"triangle": 1, "clique4": 2, "clique5": 3}
UND = {"diamond": 4, "rectangle": 5}


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(xgm):
    c = xgm.Context(0)
    yield c
    c.close()


def _bip_counts(a, b):
    return dict(triangle=0, clique4=0, clique5=0, diamond=0, rectangle=math.comb(a, 2) * math.comb(b, 2))


@pytest.fixture(scope="module")
def mixed(xgm, orc):
    """Instances of uneven size: every 4-cycle tier (LDS hash, per-block HBM
    rows, hub rows shared by several blocks) appears at a non-zero id offset,
    twice, and empty / edgeless instances sit between them."""
    from golden.make_golden import SMALL, make_graph
    inst, want = [], []

    def add(g, counts=None):
        inst.append(g)
        if counts is None:
            d = g.orient()
            counts = {p: orc.count(p, d.rowptr, d.colidx)[0] for p in DAG}
            counts.update({p: orc.count(p, g.rowptr, g.colidx)[0] for p in UND})
        want.append(counts)

    add(random_graph(xgm, 300, 2000, 31))
    add(complete_bipartite(xgm, 520, 520), _bip_counts(520, 520))
    add(xgm.from_edges(5, np.zeros(0, np.int32), np.zeros(0, np.int32)))
    add(complete_graph(xgm, 30))
    add(random_graph(xgm, 3000, 40000, 32))
    add(make_graph(xgm, *SMALL["chunglu20k_s9"][:2]))
    add(xgm.from_edges(1, np.zeros(0, np.int32), np.zeros(0, np.int32)))
    add(complete_bipartite(xgm, 520, 520), _bip_counts(520, 520))
    add(make_graph(xgm, *SMALL["rmat12_s3"][:2]))
    return inst, want


def _col(want, p):
    return np.array([w[p] for w in want], np.uint64)


def test_batch_per_instance_counts(xgm, ctx, mixed):
    inst, want = mixed
    _say("batch per-instance", len(inst))
    dag = ctx.upload_batch([g.orient() for g in inst])
    und = ctx.upload_batch(inst)
    assert dag.n_inst == und.n_inst == len(inst)
    for p, pid in DAG.items():
        _say("batch", p)
        got = ctx.count_instances(dag, pid)
        assert np.array_equal(got, _col(want, p)), p
        assert ctx.count(dag, pid) == int(_col(want, p).sum()), p + " (union total)"
    for p, pid in UND.items():
        _say("batch", p)
        got = ctx.count_instances(und, pid)
        assert np.array_equal(got, _col(want, p)), p
        # second call: the HBM counter rows were left zero by the first
        assert np.array_equal(ctx.count_instances(und, pid), _col(want, p)), p + " (repeat)"
    _say("batch rectangle hub rounds")
    for rows in (1, 2, 7):
        assert np.array_equal(ctx.count_instances(und, 5, hub_rows=rows), _col(want, "rectangle")), rows
    _say("batch rectangle LDS range counters")
    for rm in (1, 2, 3):
        assert np.array_equal(ctx.count_instances(und, 5, wedge_range=rm), _col(want, "rectangle")), rm
    assert np.array_equal(ctx.count_instances(und, 5, wedge_range=1, hub_rows=3), _col(want, "rectangle"))
    _say("batch rectangle pair strategy")
    assert np.array_equal(ctx.count_instances(und, 5, strategy=4), _col(want, "rectangle"))
    _say("batch triangle bucket strategy")
    assert np.array_equal(ctx.count_instances(dag, 1, strategy=7), _col(want, "triangle"))
    _say("batch clique strategies")
    for strat in (3, 4, 5, 6):
        assert np.array_equal(ctx.count_instances(dag, 3, strategy=strat), _col(want, "clique5")), strat
    dag.close()
    und.close()


def test_batch_plan_learned_schedule(xgm, ctx, mixed):
    """Learned + batched: the cost model's codes over the whole batch DAG
    (one schedule for all instances), captured with the other patterns in one
    plan and replayed; every per-instance count equals the oracle's, and the
    plan keeps replaying after the schedule object is freed."""
    from xgminer_amd import scheduler
    inst, want = mixed
    _say("plan with learned clique schedules")
    dag = ctx.upload_batch([g.orient() for g in inst])
    und = ctx.upload_batch(inst)
    model = scheduler.load()
    s5 = xgm.Schedule(dag, xgm.CLIQUE5, scheduler.schedule(dag, model))
    s4 = xgm.Schedule(dag, xgm.CLIQUE4, dag.sched_heuristic(xgm.CLIQUE4, xgm.STRAT_ELUT))  # edge-LUT items
    pats = [1, 2, 3, 4]
    plan = ctx.plan(pats, dag=dag, und=und, scheds=[None, s4, s5, None])
    s4.close()
    s5.close()
    expect = np.stack([_col(want, p) for p in ("triangle", "clique4", "clique5", "diamond")])
    for rep in range(3):
        _say("learned plan replay", rep)
        got, ms = plan.run()
        assert np.array_equal(got, expect), rep
        assert ms > 0
        assert ctx.count(dag, 3) == int(expect[2].sum())  # eager work between replays
    plan.close()
    dag.close()
    und.close()


def test_batch_plan_replays(xgm, ctx, mixed):
    inst, want = mixed
    _say("plan capture")
    dag = ctx.upload_batch([g.orient() for g in inst])
    und = ctx.upload_batch(inst)
    pats = [1, 2, 4, 5]
    plan = ctx.plan(pats, dag=dag, und=und)
    expect = np.stack([_col(want, p) for p in ("triangle", "clique4", "diamond", "rectangle")])
    for rep in range(4):
        _say("plan replay", rep)
        got, ms = plan.run()
        assert np.array_equal(got, expect), rep
        assert ms > 0
        # eager work on the same stream and buffers between replays
        assert np.array_equal(ctx.count_instances(und, 5), expect[3])
        assert ctx.count(dag, 1) == int(expect[0].sum())
    plan.close()
    _say("plan with options")
    opts = [dict(strategy=2), dict(strategy=6), {}, dict(hub_rows=3)]
    plan = ctx.plan(pats, dag=dag, und=und, opts=opts)
    for rep in range(2):
        got, _ = plan.run()
        assert np.array_equal(got, expect), rep
    plan.close()
    dag.close()
    und.close()
    # eager counts whose task splits need more counter / cursor rows than the
    # plan's replace the graph's scratch buffers; the plan keeps replaying on
    # the ones it captured
    _say("plan replay after scratch growth")
    und = ctx.upload_batch(inst)
    plan = ctx.plan([4, 5], und=und)
    for kw in (dict(wedge_range=1), dict(wedge_range=3), dict(wedge_range=2), dict(wedge_range=2, hub_rows=1)):
        assert np.array_equal(ctx.count_instances(und, 5, **kw), expect[3]), kw
    for rep in range(2):
        got, _ = plan.run()
        assert np.array_equal(got, expect[2:]), rep
    plan.close()
    und.close()


def test_plan_single_graph(xgm, orc, ctx):
    _say("plan single graph")
    g = random_graph(xgm, 2000, 30000, 33)
    d = g.orient()
    dag, und = ctx.upload(d), ctx.upload(g)
    plan = ctx.plan([1, 3, 4], dag=dag, und=und)
    got, _ = plan.run()
    assert got.shape == (3, 1)
    assert int(got[0, 0]) == orc.count("triangle", d.rowptr, d.colidx)[0]
    assert int(got[1, 0]) == orc.count("clique5", d.rowptr, d.colidx)[0]
    assert int(got[2, 0]) == orc.count("diamond", g.rowptr, g.colidx)[0]
    plan.close()


def test_plan_outlives_dropped_graphs(xgm, orc, ctx):
    """A plan keeps the graphs its kernels read: a temporary upload dropped
    before run(), and a graph closed while the plan is alive, still replay
    the exact counts; the memory goes with the plan."""
    import gc
    _say("plan outlives its graphs")
    g = random_graph(xgm, 1500, 20000, 34)
    d = g.orient()
    want_tc = orc.count("triangle", d.rowptr, d.colidx)[0]
    want_dm = orc.count("diamond", g.rowptr, g.colidx)[0]
    plan = ctx.plan([1], dag=ctx.upload(d))
    gc.collect()
    for _ in range(2):
        assert int(plan.run()[0][0, 0]) == want_tc
    und = ctx.upload(g)
    plan2 = ctx.plan([4], und=und)
    und.close()
    ctx.count(ctx.upload(g), 4)  # other device work between replays
    assert int(plan2.run()[0][0, 0]) == want_dm
    plan.close()
    plan2.close()


def test_autotune_picks_an_exact_candidate(xgm, ctx, mixed):
    inst, want = mixed
    dag = ctx.upload_batch([g.orient() for g in inst])
    und = ctx.upload_batch(inst)
    for g, pid, p in ((dag, 1, "triangle"), (dag, 2, "clique4"), (und, 5, "rectangle")):
        _say("autotune", p)
        best, ms = ctx.autotune(g, pid, reps=2)
        assert ms > 0
        assert np.array_equal(ctx.count_instances(g, pid, **best), _col(want, p)), (p, best)
    dag.close()
    und.close()


def test_cycle4_hub_rounds_single_graph(xgm, ctx):
    """K_{520,520}: the top ranks have wedge work 520 * (x - 520) > 2^18,
    so about 15 hub tasks; every row cap gives the closed form."""
    _say("cycle4 hub rounds single graph")
    dg = ctx.upload(complete_bipartite(xgm, 520, 520))
    want = math.comb(520, 2) ** 2
    for rows in (0, 1, 4, 15, 100):
        assert ctx.count(dg, 5, hub_rows=rows) == want, rows
    for rm in (0, 1, 2, 3):
        assert ctx.count(dg, 5, wedge_range=rm) == want, rm
    dg.close()


def test_cycle4_range_counters_golden(xgm, ctx):
    """Forced LDS range items on the golden graphs (Chung-Lu 20k has tasks
    in every wedge tier), and the one-range and many-range cases."""
    import json
    from golden.make_golden import SMALL, make_graph
    gold = json.load(open(os.path.join(GOLDEN, "counts.json")))
    for name in ("chunglu20k_s9", "rmat16_s1", "rmat12_s3"):
        _say("range golden", name)
        g = make_graph(xgm, *SMALL[name][:2])
        dg = ctx.upload(g)
        for rm in (1, 2, 0, 3):
            assert ctx.count(dg, 5, wedge_range=rm) == gold[name]["counts"]["rectangle"], (name, rm)
        dg.close()


def test_cycle4_range_spans(xgm, orc, ctx):
    """Tasks whose ids below x span several ranges (n = 300,000): 16-bit
    counters over 65,536 ids (wedge_range 1) and 32-bit counters over
    32,768 ids (wedge_range 3) give the oracle's wedge count."""
    _say("range spans chung-lu 300k")
    g = xgm.gen_chung_lu(300000, 12.0, 2.1, 3000.0, 11)
    want = orc.cycle4_wedge(g.rowptr, g.colidx)
    dg = ctx.upload(g)
    for rm in (0, 1, 2, 3):
        assert ctx.count(dg, 5, wedge_range=rm) == want, rm
    dg.close()


def test_tc_tile_options_exact(xgm, orc, ctx):
    _say("tc tile options")
    d = xgm.gen_rmat(14, 16, 0.57, 0.19, 0.19, 8).orient()
    want = orc.count("triangle", d.rowptr, d.colidx)[0]
    dg = ctx.upload(d)
    for hs in (1, 2):
        for ws, wl in ((64, 256), (32, 128), (16, 64), (1, 1), (64, 64)):
            assert ctx.count(dg, 1, 1, hash_shift=hs, wave_max_small=ws, wave_max_large=wl) == want, (hs, ws, wl)
    dg.close()


def test_batch_rejects_mixed_kinds(xgm, ctx):
    _say("batch mixed kinds")
    g = random_graph(xgm, 100, 300, 34)
    with pytest.raises(xgm.XGMError):
        ctx.upload_batch([g, g.orient()])
    dag = ctx.upload_batch([g.orient()] * 2)
    und = ctx.upload_batch([g] * 3)
    with pytest.raises(xgm.XGMError):
        ctx.plan([1, 4], dag=dag, und=und)      # instance counts differ
    with pytest.raises(xgm.XGMError):
        ctx.plan([4], dag=dag)                  # no undirected batch
    dag.close()
    und.close()


@pytest.mark.timeout(900)
def test_batch256_golden(xgm, ctx):
    """The 256-instance workload: all 1,024 counts from one plan replay
    against the oracle's per-instance counts."""
    path = os.path.join(GOLDEN, "batch_counts.json")
    gold = json.load(open(path))["instances"]
    from xgminer_amd import workloads
    specs = workloads.batch_specs()
    if len(gold) < len(specs):
        pytest.skip(f"batch_counts.json holds {len(gold)} of {len(specs)} instances")
    _say("batch256 generate")
    inst = workloads.batch_graphs(xgm, specs)
    for g, row in zip(inst, gold):
        assert f"{g.hash():016x}" == row["hash"], row["seed"]
    _say("batch256 upload")
    dag = ctx.upload_batch([g.orient() for g in inst])
    und = ctx.upload_batch(inst)
    _say("batch256 plan")
    plan = ctx.plan([1, 2, 4, 5], dag=dag, und=und)
    got, ms = plan.run()
    _say("batch256 replay ms", ms)
    expect = np.array([[r["counts"][p] for r in gold] for p in ("triangle", "clique4", "diamond", "rectangle")],
                      np.uint64)
    assert np.array_equal(got, expect)
    plan.close()
    dag.close()
    und.close()
