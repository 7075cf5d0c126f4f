"""Parity at the full benchmark sizes: R-MAT-22 (seed 2) triangles and
4-cliques against the reference binaries' counts, Chung-Lu 4M (seed 3)
diamonds and 4-cycles against the oracle's, each graph tied to its count by
its CSR hash.  5-cliques on R-MAT-22 (no golden count yet) are checked for
agreement across every strategy, which share no kernel."""
import json
import os
import sys

import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(xgm):
    c = xgm.Context(0)
    yield c
    c.close()


def _gold(name):
    return json.load(open(os.path.join(GOLDEN, "counts.json")))[name]


@pytest.mark.timeout(600)
def test_rmat22_full(xgm, ctx):
    from xgminer_amd import workloads
    gold = _gold("rmat22_s2")
    _say("rmat22 generate")
    g = workloads.rmat(xgm, 22, 2)
    assert f"{g.hash():016x}" == gold["hash"]
    d = g.orient()
    dg = ctx.upload(d)
    _say("rmat22 device orientation")
    du = ctx.upload(g)
    dd = ctx.orient(du)
    assert dd.download().hash() == d.hash()
    assert (dd.task_order() == dg.task_order()).all()
    assert ctx.count(dd, 1) == gold["counts"]["triangle"]
    dd.close()
    du.close()
    _say("rmat22 triangle")
    assert ctx.count(dg, 1) == gold["counts"]["triangle"]
    _say("rmat22 triangle bucket strategy")
    assert ctx.count(dg, 1, 7) == gold["counts"]["triangle"]
    _say("rmat22 clique4")
    for strat in (0, 6):
        assert ctx.count(dg, 2, strat) == gold["counts"]["clique4"], strat
    _say("rmat22 clique5")
    k5 = {s: ctx.count(dg, 3, s) for s in (0, 6)}
    assert len(set(k5.values())) == 1, k5
    if "clique5" in gold["counts"]:
        assert k5[0] == gold["counts"]["clique5"]
    dg.close()


@pytest.mark.timeout(600)
def test_chunglu4m_full(xgm, ctx):
    from xgminer_amd import workloads
    gold = _gold("chunglu4m_s3")
    _say("chunglu4m generate")
    g = workloads.chung_lu_4m(xgm)
    assert f"{g.hash():016x}" == gold["hash"]
    dg = ctx.upload(g)
    _say("chunglu4m diamond")
    assert ctx.count(dg, 4) == gold["counts"]["diamond"]
    _say("chunglu4m rectangle")
    assert ctx.count(dg, 5) == gold["counts"]["rectangle"]
    assert ctx.count(dg, 5, hub_rows=64) == gold["counts"]["rectangle"]
    dg.close()


@pytest.mark.timeout(600)
def test_rmat23_clique5_full(xgm, ctx):
    """The held-out learned-scheduler graph: 5-cliques on R-MAT-23 (seed 4)
    under the rank LUT and the learned schedule, against the oracle's
    bitset count; the DAG is oriented on the device."""
    from xgminer_amd import scheduler, workloads
    gold = _gold("rmat23_s4")
    _say("rmat23 generate")
    g = workloads.rmat(xgm, 23, 4)
    assert f"{g.hash():016x}" == gold["hash"]
    du = ctx.upload(g)
    dg = ctx.orient(du)
    du.close()
    want = gold["counts"]["clique5"]
    _say("rmat23 clique5 rank LUT")
    assert ctx.count(dg, 3, 6) == want
    _say("rmat23 clique5 learned")
    codes = scheduler.schedule(dg, scheduler.load(), xgm.CLIQUE5)
    s = xgm.Schedule(dg, xgm.CLIQUE5, codes)
    assert s.count() == want
    s.close()
    dg.close()
