"""GPU set operations on neighbour lists (xgm_setop_pairs, under each of its
strategies: binary search, wave64 merge path, hybrid; with ancestors,
xgm_setop_pairs_anc) and the one-shot batch call (xgm_count_batch), through
the C ABI, against the oracle's set primitives (oracle/oracle.cpp, pinned in
test_oracle.py) and numpy's set routines for the materialised sets."""
import sys

import numpy as np
import pytest

from conftest import complete_bipartite, complete_graph, random_graph

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(xgm):
    c = xgm.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def graph(xgm):
    """Random graph plus two hubs of degree 2999 and 2500 (longer than the
    2,048-element LDS stage) and isolated vertices."""
    rng = np.random.default_rng(51)
    n = 3000
    src = [rng.integers(0, n - 10, 30000), np.zeros(n - 1, np.int64), np.full(2500, 1)]
    dst = [rng.integers(0, n - 10, 30000), np.arange(1, n), rng.choice(np.arange(2, n - 10), 2500, replace=False)]
    return xgm.from_edges(n + 2, np.concatenate(src), np.concatenate(dst))  # two isolated vertices


def _pairs(g, rng, k=3000):
    a = rng.integers(0, g.n, k).astype(np.int32)
    b = rng.integers(0, g.n, k).astype(np.int32)
    a[:40] = 0            # hub against everything
    b[40:80] = 1          # against the second hub
    a[80:100] = b[80:100]  # A = B
    a[100:110] = g.n - 1  # isolated (degree 0) on either side
    b[110:120] = g.n - 2
    a[120], b[120] = 0, 1  # hub x hub
    bounds = rng.integers(-2, g.n + 5, k).astype(np.int32)
    bounds[:200] = np.iinfo(np.int32).max
    bounds[200:210] = 0
    return a, b, bounds


def _nbr(g, v):
    return g.colidx[g.rowptr[v]:g.rowptr[v + 1]]


STRATS = [0, 0x100, 0x200]  # binary search, merge path, hybrid


@pytest.mark.parametrize("strat", STRATS)
@pytest.mark.parametrize("bounded", [False, True])
def test_setop_counts_vs_oracle(xgm, orc, ctx, graph, bounded, strat):
    _say("setop counts", bounded, hex(strat))
    g = graph
    assert np.diff(g.rowptr).max() > 2048
    a, b, bounds = _pairs(g, np.random.default_rng(52))
    bd = bounds if bounded else None
    dg = ctx.upload(g)
    ci = ctx.setop_pairs(dg, xgm.SETOP_INTERSECT | strat, a, b, bd)
    cd = ctx.setop_pairs(dg, xgm.SETOP_DIFF | strat, a, b, bd)
    ce = ctx.setop_pairs(dg, xgm.SETOP_DIFF_EXCL | strat, a, b, bd)
    big = np.iinfo(np.int32).max
    for i in range(len(a)):
        A, B = _nbr(g, a[i]), _nbr(g, b[i])
        up = int(bounds[i]) if bounded else big
        assert ci[i] == orc.intersect_num_bound(A, B, up), ("intersect", i)
        assert cd[i] == orc.difference_num_noexcl(A, B, up), ("diff", i)
        assert ce[i] == orc.difference_num(A, B, up, int(b[i])), ("diff_excl", i)
    dg.close()


@pytest.mark.parametrize("strat", STRATS)
@pytest.mark.parametrize("op", [1, 2, 3])
def test_setop_materialise(xgm, orc, ctx, graph, op, strat):
    _say("setop materialise", op, hex(strat))
    g = graph
    a, b, bounds = _pairs(g, np.random.default_rng(53), 1500)
    dg = ctx.upload(g)
    counts, offs, vals = ctx.setop_pairs(dg, op | strat, a, b, bounds, materialise=True)
    assert offs[0] == 0 and np.array_equal(np.diff(offs), counts.astype(np.int64))
    for i in range(len(a)):
        A, B = _nbr(g, a[i]), _nbr(g, b[i])
        A, B = A[A < bounds[i]], B[B < bounds[i]]
        if op == 1:
            want = orc.intersect_set(A, B)
        else:
            want = np.setdiff1d(A, B)
            if op == 3:
                want = want[want != b[i]]
        assert np.array_equal(vals[offs[i]:offs[i + 1]], want), (op, i)
    dg.close()


@pytest.fixture(scope="module")
def hubs(xgm):
    """Hubs of 1,100 to 7,000 neighbours drawn from 9,000 vertices (both
    lists above the hybrid's 1024 threshold for every hub pair, many merge
    tiles per pair), plus a sparse random background."""
    rng = np.random.default_rng(57)
    n, H = 9000, 16
    src, dst = [rng.integers(0, n, 20000)], [rng.integers(0, n, 20000)]
    for h in range(H):
        k = int(rng.integers(1100, 7000))
        src.append(np.full(k, h))
        dst.append(rng.choice(np.arange(H, n), k, replace=False))
    return xgm.from_edges(n, np.concatenate(src), np.concatenate(dst)), H


@pytest.mark.parametrize("strat", STRATS)
def test_setop_hub_pairs(xgm, orc, ctx, hubs, strat):
    """Hub x hub pairs: counts, bounded counts and materialised sets equal
    the oracle / numpy under every strategy (the merge path's tile
    boundaries fall inside runs of equal and unequal keys alike)."""
    _say("setop hub pairs", hex(strat))
    g, H = hubs
    rng = np.random.default_rng(58)
    a = np.repeat(np.arange(H), H).astype(np.int32)
    b = np.tile(np.arange(H), H).astype(np.int32)
    bounds = rng.integers(0, g.n + 1, len(a)).astype(np.int32)
    dg = ctx.upload(g)
    for op in (1, 2, 3):
        c = ctx.setop_pairs(dg, op | strat, a, b)
        cb = ctx.setop_pairs(dg, op | strat, a, b, bounds)
        cnt, offs, vals = ctx.setop_pairs(dg, op | strat, a, b, bounds, materialise=True)
        assert np.array_equal(cnt, cb)
        for i in range(len(a)):
            A, B = _nbr(g, a[i]), _nbr(g, b[i])
            if op == 1:
                assert c[i] == orc.intersect_num(A, B)
                want = orc.intersect_set(A[A < bounds[i]], B[B < bounds[i]])
            else:
                assert c[i] == (orc.difference_num(A, B, np.iinfo(np.int32).max, int(b[i])) if op == 3
                                else orc.difference_num_noexcl(A, B, np.iinfo(np.int32).max))
                want = np.setdiff1d(A[A < bounds[i]], B[B < bounds[i]])
                if op == 3:
                    want = want[want != b[i]]
            assert cb[i] == len(want), (op, i)
            assert np.array_equal(vals[offs[i]:offs[i + 1]], want), (op, i)
    dg.close()


@pytest.mark.parametrize("strat", STRATS)
def test_setop_ancestors(xgm, ctx, graph, strat):
    """xgm_setop_pairs_anc: |N(a) ∩ N(b) ∩ [0, bound)| without up to four
    ancestor vertices (set_intersect.cuh:640-707), and the differences
    likewise, against numpy."""
    _say("setop ancestors", hex(strat))
    g = graph
    rng = np.random.default_rng(59)
    a, b, bounds = _pairs(g, rng, 1200)
    anc = np.full((len(a), 4), -1, np.int32)
    for i in range(len(a)):
        # ancestors drawn from the result itself, so that excluding them matters
        C = np.intersect1d(_nbr(g, a[i]), _nbr(g, b[i]))
        k = int(rng.integers(0, 5))
        pick = rng.choice(C, min(k, len(C)), replace=False) if len(C) else []
        anc[i, :len(pick)] = pick
        if k > len(pick):
            anc[i, len(pick):k] = rng.integers(0, g.n, k - len(pick))
    dg = ctx.upload(g)
    for op in (1, 2, 3):
        got = ctx.setop_pairs_anc(dg, op | strat, a, b, anc, bounds)
        got2 = ctx.setop_pairs_anc(dg, op | strat, a, b, anc[:, :2])
        for i in range(len(a)):
            A, B = _nbr(g, a[i]), _nbr(g, b[i])
            for ub, an, gv in ((bounds[i], anc[i], got[i]), (np.iinfo(np.int32).max, anc[i, :2], got2[i])):
                A1, B1 = A[A < ub], B[B < ub]
                R = np.intersect1d(A1, B1) if op == 1 else np.setdiff1d(A1, B1)
                if op == 3:
                    R = R[R != b[i]]
                R = R[~np.isin(R, an[an >= 0])]
                assert gv == len(R), (op, i)
    with pytest.raises(xgm.XGMError):
        ctx.setop_pairs_anc(dg, 1, a[:2], b[:2], np.zeros((2, 5), np.int32))  # more than 4 ancestors
    with pytest.raises(xgm.XGMError):
        ctx.setop_pairs(dg, 1 | 0x300, a[:2], b[:2])                          # two strategies at once
    dg.close()


def test_setop_on_dag_and_errors(xgm, orc, ctx):
    _say("setop dag + errors")
    d = random_graph(xgm, 500, 5000, 54).orient()
    dg = ctx.upload(d)
    rng = np.random.default_rng(55)
    a = rng.integers(0, d.n, 500).astype(np.int32)
    b = rng.integers(0, d.n, 500).astype(np.int32)
    got = ctx.setop_pairs(dg, xgm.SETOP_INTERSECT, a, b)
    for i in range(len(a)):
        assert got[i] == orc.intersect_num(_nbr(d, a[i]), _nbr(d, b[i]))
    assert len(ctx.setop_pairs(dg, 1, np.zeros(0, np.int32), np.zeros(0, np.int32))) == 0
    with pytest.raises(xgm.XGMError):
        ctx.setop_pairs(dg, 1, [0, d.n], [1, 2])           # vertex outside the graph
    with pytest.raises(xgm.XGMError):
        ctx.setop_pairs(dg, 9, [0], [1])                    # unknown operation
    dg.close()


def test_count_batch_mixed_patterns(xgm, orc, ctx):
    _say("count_batch")
    gs = [random_graph(xgm, 400, 3000, 60 + i) for i in range(6)] + [complete_graph(xgm, 12),
                                                                     complete_bipartite(xgm, 7, 9)]
    pats = [1, 2, 3, 4, 5, 1, 3, 5]
    names = {1: "triangle", 2: "clique4", 3: "clique5", 4: "diamond", 5: "rectangle"}
    got = ctx.count_batch(gs, pats)
    for g, p, c in zip(gs, pats, got):
        h = g.orient() if p <= 3 else g
        assert int(c) == orc.count(names[p], h.rowptr, h.colidx)[0], p
    with pytest.raises(xgm.XGMError):
        ctx.count_batch([gs[0].orient()], [1])              # undirected input required
