"""The allocation-free entry point (xgm_setop_pairs_noalloc) against the
allocating one (xgm_setop_pairs) through the C ABI: bit-identical counts on
two calls in a row, the second with fewer pairs (the static staging array
keeps the first call's contents), for every operation with and without
bounds, and the size limit rejected with a clear error."""
import sys

import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_noalloc_matches_allocating_entry(xgm, orc):
    ctx = xgm.Context(0)
    rng = np.random.default_rng(61)
    n = 3000
    # a random graph plus a hub longer than the 2,048-element LDS stage
    src = np.concatenate([rng.integers(0, n, 30000), np.zeros(n - 1, np.int64)])
    dst = np.concatenate([rng.integers(0, n, 30000), np.arange(1, n)])
    g = xgm.from_edges(n, src, dst)
    dg = ctx.upload(g)
    for op in (xgm.SETOP_INTERSECT, xgm.SETOP_DIFF, xgm.SETOP_DIFF_EXCL):
        for k in (20000, 700):  # second call smaller: stale staging must not leak in
            for bounded in (False, True):
                _say("noalloc op", op, "pairs", k, "bounded", bounded)
                a = rng.integers(0, n, k).astype(np.int32)
                b = rng.integers(0, n, k).astype(np.int32)
                a[:3] = 0  # the hub on both sides
                b[:2] = 0
                bd = rng.integers(0, n + 1, k).astype(np.int32) if bounded else None
                want = ctx.setop_pairs(dg, op, a, b, bd)
                got = ctx.setop_pairs_noalloc(dg, op, a, b, bd)
                assert np.array_equal(got, want), (op, k, bounded)
    # spot check against the oracle's primitives
    a = rng.integers(0, n, 200).astype(np.int32)
    b = rng.integers(0, n, 200).astype(np.int32)
    got = ctx.setop_pairs_noalloc(dg, xgm.SETOP_INTERSECT, a, b)
    N = lambda v: g.colidx[g.rowptr[v]:g.rowptr[v + 1]]
    assert all(int(got[i]) == orc.intersect_num(N(a[i]), N(b[i])) for i in range(200))
    # limits and errors
    assert ctx.setop_pairs_noalloc(dg, xgm.SETOP_INTERSECT, [], []).shape == (0,)
    big = np.zeros(xgm.NOALLOC_MAX_PAIRS + 1, np.int32)
    with pytest.raises(xgm.XGMError) as e:
        ctx.setop_pairs_noalloc(dg, xgm.SETOP_INTERSECT, big, big)
    assert e.value.status == xgm.XGM_EINVAL
    with pytest.raises(xgm.XGMError):
        ctx.setop_pairs_noalloc(dg, xgm.SETOP_INTERSECT, [n], [0])
    dg.close()
    ctx.close()
