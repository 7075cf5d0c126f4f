"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

* set primitives: brute force with numpy set routines on random sorted lists
  and the edge cases the reference's GPU code special-cases (empty operands,
  bounds at or below the minimum, disjoint and identical lists);
* orientation: the reference rule (codes/GraphMiner/src/common/graph.cc:246-247)
  checked edge by edge, and equality with the library's orientation;
* patterns: reference known answers on citeseer
  (codes/GraphMiner/src/triangle/README.md:53, src/motif/README.md:57,
  src/sgl/README.md:53), closed forms on K_n and K_{a,b}, brute force over
  vertex subsets on tiny graphs, the identities between patterns, and the
  counts the reference's own binaries wrote into tests/golden/counts.json.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from conftest import CITESEER, GOLDEN, complete_bipartite, complete_graph, random_graph

# ------------------------------------------------------------------ set ops


def _rand_sorted(rng, n, hi):
    return np.unique(rng.integers(0, hi, n)).astype(np.int32)


def test_intersect_bruteforce(orc):
    rng = np.random.default_rng(0)
    for _ in range(400):
        a = _rand_sorted(rng, rng.integers(0, 60), 100)
        b = _rand_sorted(rng, rng.integers(0, 60), 100)
        ref = np.intersect1d(a, b)
        assert orc.intersect_num(a, b) == len(ref)
        assert np.array_equal(orc.intersect_set(a, b), ref)
        up = int(rng.integers(-5, 110))
        assert orc.intersect_num_bound(a, b, up) == int((ref < up).sum())


def test_intersect_edge_cases(orc):
    e = np.zeros(0, np.int32)
    a = np.array([1, 5, 9], np.int32)
    assert orc.intersect_num(e, a) == 0 and orc.intersect_num(a, e) == 0
    assert orc.intersect_num(a, a) == 3
    assert orc.intersect_num(a, np.array([2, 6], np.int32)) == 0
    assert orc.intersect_num_bound(a, a, 1) == 0   # bound at the minimum
    assert orc.intersect_num_bound(a, a, 0) == 0   # bound below the minimum
    assert orc.intersect_num_bound(a, a, 10) == 3


def test_difference_semantics(orc):
    """GraphMiner excludes the owner vid(B) (src/common/VertexSet.cc:69-88);
    GLUMIN's GPU difference does not (set_difference.cuh:20-35)."""
    rng = np.random.default_rng(1)
    for _ in range(400):
        a = _rand_sorted(rng, rng.integers(0, 50), 80)
        b = _rand_sorted(rng, rng.integers(0, 50), 80)
        up = int(rng.integers(0, 90))
        vid = int(rng.integers(0, 80))
        diff = np.setdiff1d(a, b)
        assert orc.difference_num_noexcl(a, b, up) == int((diff < up).sum())
        # the GraphMiner merge stops at the first element of B >= up, which
        # does not change the set below up
        assert orc.difference_num(a, b, up, vid) == int(((diff < up) & (diff != vid)).sum())
    e = np.zeros(0, np.int32)
    assert orc.difference_num(np.array([3, 4], np.int32), e, 100, 4) == 1
    assert orc.difference_num_noexcl(e, np.array([1], np.int32), 100) == 0


# -------------------------------------------------------------- orientation


def test_orientation_rule(xgm, orc):
    g = random_graph(xgm, 300, 3000, seed=3)
    deg = g.degree()
    rp, ci = orc.orient(g.rowptr, g.colidx)
    assert rp[-1] == g.nnz // 2
    for u in range(g.n):
        out = ci[rp[u]:rp[u + 1]]
        assert np.all(np.diff(out) > 0)
        nbr = g.colidx[g.rowptr[u]:g.rowptr[u + 1]]
        keep = [v for v in nbr if deg[v] > deg[u] or (deg[v] == deg[u] and v > u)]
        assert list(out) == keep
    d = g.orient()
    assert np.array_equal(d.rowptr, rp) and np.array_equal(d.colidx, ci)


# ------------------------------------------------------------------ patterns


def _counts(orc, g):
    d = g.orient()
    out = {}
    for p in ("triangle", "clique4", "clique5"):
        out[p] = orc.count(p, d.rowptr, d.colidx)[0]
    for p in ("diamond", "rectangle"):
        out[p] = orc.count(p, g.rowptr, g.colidx)[0]
    return out


def test_citeseer_known_answers(xgm, orc):
    g = xgm.load_csr(CITESEER)
    c = _counts(orc, g)
    assert c["triangle"] == 1166      # triangle/README.md:53
    assert c["clique4"] == 255        # motif/README.md:57
    assert c["diamond"] == 3730       # sgl/README.md:53
    assert c["rectangle"] == 6059     # sgl/README.md:53
    assert c["clique5"] == 46         # networkx brute force (no reference file publishes it)
    # identities from motif/README.md:53-57: edge-induced = induced + multiples of K4
    assert c["diamond"] == 2200 + 6 * c["clique4"]
    assert c["rectangle"] == 3094 + 2200 + 3 * c["clique4"]


@pytest.mark.parametrize("n", [4, 5, 7, 12])
def test_complete_graph_closed_forms(xgm, orc, n):
    c = _counts(orc, complete_graph(xgm, n))
    assert c["triangle"] == math.comb(n, 3)
    assert c["clique4"] == math.comb(n, 4)
    assert c["clique5"] == math.comb(n, 5)
    assert c["diamond"] == 6 * math.comb(n, 4)
    assert c["rectangle"] == 3 * math.comb(n, 4)


@pytest.mark.parametrize("a,b", [(2, 2), (3, 5), (6, 4)])
def test_bipartite_closed_forms(xgm, orc, a, b):
    c = _counts(orc, complete_bipartite(xgm, a, b))
    assert c["triangle"] == 0 and c["clique4"] == 0 and c["diamond"] == 0
    assert c["rectangle"] == math.comb(a, 2) * math.comb(b, 2)


def _brute(g):
    n = g.n
    adj = [set(g.colidx[g.rowptr[u]:g.rowptr[u + 1]].tolist()) for u in range(n)]
    E = lambda x, y: y in adj[x]
    tri = sum(1 for a, b, c in itertools.combinations(range(n), 3) if E(a, b) and E(a, c) and E(b, c))
    k4 = k5 = dia = cyc = 0
    for q in itertools.combinations(range(n), 4):
        es = [E(x, y) for x, y in itertools.combinations(q, 2)]
        ne = sum(es)
        k4 += ne == 6
        # edge-induced copies inside a 4-set: diamonds = #(5-edge subgraphs
        # with the missing edge's endpoints...) counted directly
        a, b, c, d = q
        # 4-cycles on this vertex set: three possible cyclic orders
        for x, y, z, w in ((a, b, c, d), (a, b, d, c), (a, c, b, d)):
            cyc += E(x, y) and E(y, z) and E(z, w) and E(w, x)
        # diamonds: choose the missing (non-required) pair; remaining 5 pairs must be edges
        pairs = list(itertools.combinations(q, 2))
        for skip in range(6):
            dia += all(E(*pairs[i]) for i in range(6) if i != skip)
    for q in itertools.combinations(range(n), 5):
        k5 += all(E(x, y) for x, y in itertools.combinations(q, 2))
    return dict(triangle=tri, clique4=k4, clique5=k5, diamond=dia, rectangle=cyc)


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_bruteforce_tiny(xgm, orc, seed):
    g = random_graph(xgm, 16, 60, seed)
    assert _counts(orc, g) == _brute(g)


def test_golden_reference_counts(xgm, orc):
    """Counts written by the reference binaries (oracle/_ref) for seeded graphs."""
    from golden.make_golden import SMALL, make_graph
    gold = json.load(open(os.path.join(GOLDEN, "counts.json")))
    for name in ("rmat10_s7", "rmat12_s3", "er2000_d16_s5", "chunglu20k_s9"):
        kind, kw, _ = SMALL[name]
        g = make_graph(xgm, kind, kw)
        assert f"{g.hash():016x}" == gold[name]["hash"], name
        assert _counts(orc, g) == gold[name]["counts"], name


def test_cycle4_wedge_matches_restatement(xgm, orc):
    """The wedge-aggregation 4-cycle counter (large golden counts) equals the
    rectangle.h restatement, the reference's golden counts, brute force and
    closed forms."""
    from golden.make_golden import SMALL, make_graph
    gold = json.load(open(os.path.join(GOLDEN, "counts.json")))
    for name in ("rmat10_s7", "rmat12_s3", "er2000_d16_s5", "chunglu20k_s9"):
        g = make_graph(xgm, *SMALL[name][:2])
        assert orc.cycle4_wedge(g.rowptr, g.colidx, 4) == gold[name]["counts"]["rectangle"], name
    for seed in (11, 12, 13):
        g = random_graph(xgm, 16, 60, seed)
        assert orc.cycle4_wedge(g.rowptr, g.colidx) == _brute(g)["rectangle"]
    for seed in (41, 42):
        g = random_graph(xgm, 2000, 30000, seed)
        assert orc.cycle4_wedge(g.rowptr, g.colidx) == orc.count("rectangle", g.rowptr, g.colidx)[0]
    for a, b in ((2, 2), (3, 5), (40, 70)):
        g = complete_bipartite(xgm, a, b)
        assert orc.cycle4_wedge(g.rowptr, g.colidx) == math.comb(a, 2) * math.comb(b, 2)
    for n in (4, 9, 30):
        g = complete_graph(xgm, n)
        assert orc.cycle4_wedge(g.rowptr, g.colidx) == 3 * math.comb(n, 4)
    assert orc.cycle4_wedge(xgm.load_csr(CITESEER).rowptr, xgm.load_csr(CITESEER).colidx) == 6059


def _brute_motifs(g):
    """Vertex-induced 3- and 4-motifs by classifying every vertex subset."""
    n = g.n
    adj = [set(g.colidx[g.rowptr[u]:g.rowptr[u + 1]].tolist()) for u in range(n)]
    m3 = [0, 0]
    for q in itertools.combinations(range(n), 3):
        ne = sum(y in adj[x] for x, y in itertools.combinations(q, 2))
        if ne == 2:
            m3[0] += 1
        elif ne == 3:
            m3[1] += 1
    m4 = [0] * 6   # 3-star, 4-path, tailed triangle, 4-cycle, diamond, 4-clique
    for q in itertools.combinations(range(n), 4):
        deg = sorted(sum(y in adj[x] for y in q if y != x) for x in q)
        ne = sum(deg) // 2
        if ne == 3 and deg == [1, 1, 1, 3]:
            m4[0] += 1
        elif ne == 3 and deg == [1, 1, 2, 2]:
            m4[1] += 1
        elif ne == 4 and deg == [1, 2, 2, 3]:
            m4[2] += 1
        elif ne == 4 and deg == [2, 2, 2, 2]:
            m4[3] += 1
        elif ne == 5:
            m4[4] += 1
        elif ne == 6:
            m4[5] += 1
    return m3, m4


def test_motif_known_answers_and_bruteforce(xgm, orc):
    """The reference's citeseer motif counts (src/motif/README.md:53-66) and
    brute force over every 3- and 4-vertex subset of small graphs."""
    g = xgm.load_csr(CITESEER)
    assert orc.motif(g.rowptr, g.colidx, 4) == [222630, 111153, 22900, 3094, 2200, 255]
    assert orc.motif(g.rowptr, g.colidx, 3) == [23380, 1166]
    for seed in (11, 12, 13, 14):
        h = random_graph(xgm, 16, 45 + 5 * seed % 20, seed)
        m3, m4 = _brute_motifs(h)
        assert orc.motif(h.rowptr, h.colidx, 3, 2) == m3, seed
        assert orc.motif(h.rowptr, h.colidx, 4, 2) == m4, seed
    for a, b in ((3, 4), (5, 5)):
        h = complete_bipartite(xgm, a, b)
        # K_{a,b}: 3-stars a*C(b,3) + b*C(a,3), 4-paths 0 (every 4-path closes), 4-cycles C(a,2)C(b,2)
        assert orc.motif(h.rowptr, h.colidx, 4) == [a * math.comb(b, 3) + b * math.comb(a, 3), 0, 0,
                                                    math.comb(a, 2) * math.comb(b, 2), 0, 0]


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                     "motif_omp_formula")), reason="oracle/_ref not built")
def test_motif_restatement_matches_reference_binary(xgm, orc):
    for seed in (31, 32):
        g = random_graph(xgm, 600, 9000, seed)
        for k in (3, 4):
            assert orc.motif(g.rowptr, g.colidx, k) == orc.ref_motif_graph(g, k), (seed, k)


def test_clique_bitset_matches_restatement(xgm, orc):
    """The bitset clique counter (large golden counts) equals the automine
    restatement, the reference's golden counts, brute force and K_n."""
    from golden.make_golden import SMALL, make_graph
    gold = json.load(open(os.path.join(GOLDEN, "counts.json")))
    for name in ("rmat10_s7", "rmat12_s3", "er2000_d16_s5", "chunglu20k_s9", "rmat16_s1"):
        d = make_graph(xgm, *SMALL[name][:2]).orient()
        for k in (4, 5):
            assert orc.clique_bitset(d.rowptr, d.colidx, k, threads=4) == gold[name]["counts"][f"clique{k}"], (name, k)
    for seed in (11, 12, 13):
        g = random_graph(xgm, 16, 60, seed)
        d = g.orient()
        b = _brute(g)
        assert orc.clique_bitset(d.rowptr, d.colidx, 4) == b["clique4"]
        assert orc.clique_bitset(d.rowptr, d.colidx, 5) == b["clique5"]
    for n in (5, 9, 70, 130):   # 70 and 130 cross 64-bit words
        d = complete_graph(xgm, n).orient()
        assert orc.clique_bitset(d.rowptr, d.colidx, 4) == math.comb(n, 4)
        assert orc.clique_bitset(d.rowptr, d.colidx, 5) == math.comb(n, 5)
    d = random_graph(xgm, 3000, 60000, 43).orient()
    for k in (4, 5):
        assert orc.clique_bitset(d.rowptr, d.colidx, k) == orc.count(f"clique{k}", d.rowptr, d.colidx)[0]
    c = xgm.load_csr(CITESEER).orient()
    assert orc.clique_bitset(c.rowptr, c.colidx, 4) == 255 and orc.clique_bitset(c.rowptr, c.colidx, 5) == 46


def test_algorithmic_bytes_tc_closed_form(xgm, orc):
    """B_TC = 4*sum_v d+(v)*(d+(v)+d-(v)) + 40*m on the DAG."""
    g = random_graph(xgm, 500, 5000, seed=5)
    d = g.orient()
    _, B = orc.count("triangle", d.rowptr, d.colidx)
    dout = d.degree().astype(np.int64)
    din = np.bincount(d.colidx, minlength=d.n).astype(np.int64)
    assert B == int(4 * (dout * dout).sum() + 4 * (din * dout).sum() + 40 * d.nnz)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                     "tc_omp_base")), reason="oracle/_ref not built")
def test_restatement_matches_reference_binaries(xgm, orc):
    g = random_graph(xgm, 400, 6000, seed=21)
    mine = _counts(orc, g)
    for p in mine:
        assert orc.ref_count_graph(p, g)[0] == mine[p], p


def test_mlp_reference_gradients_vs_finite_differences():
    """Pins oracle/mlp_ref.py: its analytic backward pass (rounding off)
    equals central finite differences of its own loss, parameter by
    parameter on a sample of every tensor."""
    from oracle import mlp_ref
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mlp_ref.NPARAM) * 0.3
    X = rng.standard_normal((40, 16))
    T = rng.standard_normal((40, 5))
    M = (rng.random((40, 5)) < 0.8).astype(np.float64)
    _, g = mlp_ref.loss_and_grads(p, X, T, M, rounding=False)
    h = 1e-6
    o = 0
    for name, shape in mlp_ref.SIZES:
        k = int(np.prod(shape))
        for i in rng.choice(k, size=min(k, 12), replace=False):
            q = p.copy(); q[o + i] += h
            lp = mlp_ref.loss(q, X, T, M, rounding=False)
            q[o + i] -= 2 * h
            lm = mlp_ref.loss(q, X, T, M, rounding=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[o + i]) <= 1e-5 * max(1.0, abs(fd)), (name, i, fd, g[o + i])
        o += k


def test_bf16_rounding_known_values():
    from oracle.mlp_ref import bf16
    # exactly representable, round-half-to-even, and carry into the exponent
    assert bf16(1.0) == 1.0 and bf16(-2.5) == -2.5
    assert bf16(1.0 + 2 ** -8) == 1.0                 # halfway, even neighbour 1.0
    assert bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6   # halfway, rounds up to even
    assert bf16(1.0 + 2 ** -7 + 2 ** -9) == 1.0 + 2 ** -7
