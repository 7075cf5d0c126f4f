"""On-device orientation (xgm_graph_orient): the DAG built in HBM equals the
host orientation (GraphMiner Graph::orientation,
codes/GraphMiner/src/common/graph.cc:233-279) array for array, its task
order equals the one built for the uploaded host DAG, and every DAG count
on it equals the oracle's."""
import sys

import numpy as np
import pytest

from conftest import CITESEER, complete_bipartite, complete_graph, random_graph

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(xgm):
    c = xgm.Context(0)
    yield c
    c.close()


def _graphs(xgm):
    star_src = np.zeros(300, np.int32)
    star_dst = np.arange(1, 301, dtype=np.int32)
    return {
        "citeseer": xgm.load_csr(CITESEER),
        "empty": xgm.from_edges(10, np.zeros(0, np.int32), np.zeros(0, np.int32)),
        "path": xgm.from_edges(4, [0, 1, 2], [1, 2, 3]),
        "k9": complete_graph(xgm, 9),           # equal degrees: id tie-break everywhere
        "k5_7": complete_bipartite(xgm, 5, 7),
        "star300": xgm.from_edges(301, star_src, star_dst),  # one hub, 300 leaves (list > 64)
        "random": random_graph(xgm, 3000, 40000, 5),
        "rmat12": xgm.gen_rmat(12, 16, 0.57, 0.19, 0.19, 3),
        "chunglu": xgm.gen_chung_lu(20000, 12.0, 2.1, 2000.0, 7),
    }


def test_orient_equals_host(xgm, orc, ctx):
    for name, g in _graphs(xgm).items():
        _say("orient", name)
        want = g.orient()
        dd = ctx.orient(ctx.upload(g))
        got = dd.download()
        assert got.flags == want.flags, name
        assert np.array_equal(got.rowptr, want.rowptr), name
        assert np.array_equal(got.colidx, want.colidx), name
        ref = ctx.upload(want)
        assert np.array_equal(dd.task_order(), ref.task_order()), name
        for pid, pname in ((xgm.TRIANGLE, "triangle"), (xgm.CLIQUE4, "clique4"), (xgm.CLIQUE5, "clique5")):
            c = ctx.count(dd, pid)
            assert c == orc.count(pname, want.rowptr, want.colidx)[0], (name, pname)
            if pid != xgm.TRIANGLE:  # the rank-ordered LUT reads the reverse lists
                assert ctx.count(dd, pid, xgm.STRAT_LUT_RANK) == c, (name, pname)
        dd.close()
        ref.close()


def test_orient_rejects(xgm, ctx):
    _say("orient rejects")
    g = xgm.gen_rmat(10, 8, 0.57, 0.19, 0.19, 1)
    dag = ctx.upload(g.orient())
    with pytest.raises(xgm.XGMError):
        ctx.orient(dag)  # already a DAG
    with pytest.raises(xgm.XGMError):
        ctx.orient(ctx.upload_batch([g, g]))


def test_orient_rmat18_hash(xgm, ctx):
    _say("orient rmat18")
    g = xgm.gen_rmat(18, 16, 0.57, 0.19, 0.19, 2)
    want = g.orient()
    dd = ctx.orient(ctx.upload(g))
    assert dd.download().hash() == want.hash()
    assert ctx.count(dd, xgm.TRIANGLE) == ctx.count(want, xgm.TRIANGLE)


def test_upload_rejects_asymmetric(xgm, ctx):
    """The symmetry of an undirected upload is checked on the device: an
    entry above the diagonal without its mirror, and one below it."""
    _say("upload asymmetric")
    flags = xgm.CSR_SORTED | xgm.CSR_SYMMETRIC
    for rowptr, colidx in (([0, 1, 1], [1]), ([0, 0, 1], [0]), ([0, 2, 3, 4], [1, 2, 0, 1])):
        g = xgm.Graph.from_arrays(np.array(rowptr), np.array(colidx), flags)
        assert not g.validate()
        with pytest.raises(xgm.XGMError):
            ctx.upload(g)
    ok = xgm.Graph.from_arrays(np.array([0, 2, 3, 4]), np.array([1, 2, 0, 0]), flags)
    assert ok.validate()
    assert ctx.count(ctx.upload(ok), xgm.DIAMOND_EDGE) == 0


def test_device_tc_items_equal_host(xgm, ctx):
    """Triangle task items built in HBM (a DAG from xgm_graph_orient, whose
    CSR never comes back to the host) equal the host builder's items for the
    same DAG uploaded from the host, item for item and tier for tier, under
    every tier split; hubs above 2,048 out-neighbours and items cut at the
    64K-probe slice bound included."""
    graphs = dict(_graphs(xgm))
    graphs["k2200"] = complete_graph(xgm, 2200)        # tier > 2048, many slices per hub
    graphs["rmat16"] = xgm.gen_rmat(16, 16, 0.57, 0.19, 0.19, 1)
    splits = [{}, dict(wave_max_small=32, wave_max_large=128), dict(wave_max_small=16, wave_max_large=64),
              dict(wave_max_small=64, wave_max_large=256)]
    for name, g in graphs.items():
        _say("tc items", name)
        host = ctx.upload(g.orient())
        dev = ctx.orient(ctx.upload(g))
        for opts in splits:
            hi, ht = host.tc_items(**opts)
            di, dt = dev.tc_items(**opts)
            assert ht == dt, (name, opts)
            assert np.array_equal(hi, di), (name, opts)
            assert ctx.count(dev, xgm.TRIANGLE, **opts) == ctx.count(host, xgm.TRIANGLE, **opts)
        dev.close()
        host.close()


def _lists(dg):
    return {nm: dg.task_list(nm) for nm in dg.task_list_names()}


def _same_lists(a, b, tag):
    la, lb = _lists(a), _lists(b)
    assert sorted(la) == sorted(lb), tag
    for nm in la:
        assert la[nm][1] == lb[nm][1], (tag, nm)
        assert np.array_equal(la[nm][0], lb[nm][0]), (tag, nm)


def _routes(xgm, ctx, make):
    """(host-route graph, device-route graph) from make(), each uploaded and
    prepared under its route by the caller's counts."""
    ctx.set_prep(xgm.PREP_HOST)
    h = make()
    ctx.set_prep(xgm.PREP_DEVICE)
    d = make()
    return h, d


def test_device_undirected_lists_equal_host(xgm, ctx):
    """Task lists of undirected graphs built in HBM (diamond owners and their
    hub slices, the 4-cycle relabel with its wedge, range and hub lists, the
    vertex-induced motif orders, the motif-census owners) equal the host
    builders' lists name for name and int for int, and give the same counts;
    hubs above 1024 / 8192 neighbours, 4-cycle hub rounds and batches of
    instances included."""
    graphs = dict(_graphs(xgm))
    leaves = np.arange(1, 9001, dtype=np.int32)
    graphs["star9000"] = xgm.from_edges(9001, np.zeros(9000, np.int32), leaves)  # owner tier above 8192
    graphs["k1100"] = complete_graph(xgm, 1100)   # owners above 1024, cut into 64K-work slices
    cases = [(xgm.DIAMOND_EDGE, {}), (xgm.CYCLE4_EDGE, {}), (xgm.CYCLE4_EDGE, dict(wedge_range=1)),
             (xgm.CYCLE4_EDGE, dict(wedge_range=2, hub_rows=3)), (xgm.STAR3_V, {}),
             (xgm.TAILEDTRI_V, dict(strategy=xgm.STRAT_LIST))]
    try:
        for name, g in graphs.items():
            for pid, opts in cases:
                if name == "k1100" and pid in (xgm.STAR3_V, xgm.TAILEDTRI_V):
                    continue
                _say("undirected lists", name, pid, opts)
                h, d = _routes(xgm, ctx, lambda: ctx.upload(g))
                ctx.set_prep(xgm.PREP_HOST)
                ch = ctx.count(h, pid, **opts)
                ctx.set_prep(xgm.PREP_DEVICE)
                assert ctx.count(d, pid, **opts) == ch, (name, pid, opts)
                _same_lists(h, d, (name, pid, opts))
                h.close()
                d.close()
            for k in (3, 4):
                _say("motif census lists", name, k)
                h, d = _routes(xgm, ctx, lambda: ctx.upload(g))
                ctx.set_prep(xgm.PREP_HOST)
                mh = ctx.motifs(h, k)
                ctx.set_prep(xgm.PREP_DEVICE)
                assert np.array_equal(ctx.motifs(d, k), mh), (name, k)
                _same_lists(h, d, (name, k))
                h.close()
                d.close()
        parts = [graphs["citeseer"], graphs["k9"], graphs["star300"], graphs["empty"], graphs["rmat12"],
                 graphs["chunglu"]]
        for pid, opts in cases[:4]:
            _say("batch lists", pid, opts)
            h, d = _routes(xgm, ctx, lambda: ctx.upload_batch(parts))
            ctx.set_prep(xgm.PREP_HOST)
            ch = ctx.count_instances(h, pid, **opts)
            ctx.set_prep(xgm.PREP_DEVICE)
            assert np.array_equal(ctx.count_instances(d, pid, **opts), ch), (pid, opts)
            _same_lists(h, d, ("batch", pid, opts))
            h.close()
            d.close()
        _say("batch motif census")
        h, d = _routes(xgm, ctx, lambda: ctx.upload_batch(parts))
        ctx.set_prep(xgm.PREP_HOST)
        mh = ctx.motifs(h, 4)
        ctx.set_prep(xgm.PREP_DEVICE)
        assert np.array_equal(ctx.motifs(d, 4), mh)
        h.close()
        d.close()
    finally:
        ctx.set_prep(xgm.PREP_AUTO)


def test_prep_route_host_for_device_dag(xgm, ctx):
    """PREP_HOST on a DAG built in HBM copies it back and builds the triangle
    items on the host: the same items as the device builder's."""
    g = xgm.gen_rmat(14, 16, 0.57, 0.19, 0.19, 5)
    try:
        _say("tc items host route")
        ctx.set_prep(xgm.PREP_HOST)
        h = ctx.orient(ctx.upload(g))
        hi, ht = h.tc_items()
        ctx.set_prep(xgm.PREP_DEVICE)
        d = ctx.orient(ctx.upload(g))
        di, dt = d.tc_items()
        assert ht == dt
        assert np.array_equal(hi, di)
        assert ctx.count(h, xgm.TRIANGLE) == ctx.count(d, xgm.TRIANGLE)
        with pytest.raises(xgm.XGMError):
            ctx.set_prep(7)
    finally:
        ctx.set_prep(xgm.PREP_AUTO)
