"""Vertex-induced 3- and 4-motif censuses on the GPU (xgm_motifs) against
the reference's citeseer counts (codes/GraphMiner/src/motif/README.md:53-66)
and the oracle's restatement of its formula path, through the C ABI."""
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


def test_motifs_citeseer(xgm, ctx):
    _say("motifs citeseer")
    dg = ctx.upload(xgm.load_csr(CITESEER))
    assert ctx.motifs(dg, 4)[0].tolist() == [222630, 111153, 22900, 3094, 2200, 255]
    assert ctx.motifs(dg, 3)[0].tolist() == [23380, 1166]
    dg.close()


def _graphs(xgm):
    from golden.make_golden import SMALL, make_graph
    n = 9000   # one vertex of degree 8999: the motif edge kernel's HBM-set tier
    star = xgm.from_edges(n, np.concatenate([np.zeros(n - 1, np.int64), np.arange(1, 300)]),
                          np.concatenate([np.arange(1, n), np.arange(2, 301)]))
    return [random_graph(xgm, 2000, 30000, 71), star, complete_graph(xgm, 40), complete_bipartite(xgm, 30, 50),
            make_graph(xgm, *SMALL["chunglu20k_s9"][:2]), make_graph(xgm, *SMALL["rmat12_s3"][:2])]


def test_motifs_vs_oracle(xgm, orc, ctx):
    for i, g in enumerate(_graphs(xgm)):
        _say("motifs graph", i)
        dg = ctx.upload(g)
        for k in (3, 4):
            assert ctx.motifs(dg, k)[0].tolist() == orc.motif(g.rowptr, g.colidx, k), (i, k)
        dg.close()


def test_motifs_batch(xgm, orc, ctx):
    _say("motifs batch")
    gs = _graphs(xgm)
    dg = ctx.upload_batch(gs)
    got4, got3 = ctx.motifs(dg, 4), ctx.motifs(dg, 3)
    for i, g in enumerate(gs):
        assert got4[i].tolist() == orc.motif(g.rowptr, g.colidx, 4), i
        assert got3[i].tolist() == orc.motif(g.rowptr, g.colidx, 3), i
    dg.close()


def test_motifs_reject_dag(xgm, ctx):
    g = random_graph(xgm, 100, 400, 72)
    dg = ctx.upload(g.orient())
    with pytest.raises(xgm.XGMError):
        ctx.motifs(dg, 4)
    dg.close()
