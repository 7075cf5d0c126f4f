// ORACLE - test infrastructure only.  A CPU restatement of the reference's
// GraphMiner OpenMP path, used by tests/, by the smoke check and by the
// benchmark's cpu_baseline leg as the checker.  Nothing in the product
// library links or calls this file.
//
// Every function follows the reference file:line named beside it, with the
// same set semantics (sorted int32 lists, merge-based counting) and the
// same `schedule(dynamic,1)` vertex loop.  Each pattern also totals the
// algorithmic bytes of its set-operation stream:
//   B = sum over set ops of 4*(|A|+|B|)  +  40 per task
// (a task is one DAG or undirected edge; 8 B src/dst + 32 B of rowptr).
// For bounded operations (x < up) only the elements below the bound count.
//
// Pinning (see tests/test_oracle.py): brute force vs std::set_* on random
// lists; reference known answers on citeseer (1166 / 255 / 3730 / 6059);
// closed forms on complete and complete-bipartite graphs; identities
// between patterns; and the reference's own GraphMiner binaries built from
// source into oracle/_ref/ by oracle/Makefile.
#include <omp.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

typedef int32_t vid;
typedef int64_t eid;
typedef unsigned long long u64;

namespace {

// GraphMiner VertexSet::get_intersect_num (include/VertexSet.h:65-76).
inline u64 isect_num(const vid *a, eid na, const vid *b, eid nb) {
  eid i = 0, j = 0;
  u64 c = 0;
  while (i < na && j < nb) {
    vid x = a[i], y = b[j];
    if (x <= y) ++i;
    if (y <= x) ++j;
    if (x == y) ++c;
  }
  return c;
}

// VertexSet::intersect_ns (include/VertexSet.h:110-122): count x < up.
inline u64 isect_num_bound(const vid *a, eid na, const vid *b, eid nb, vid up) {
  eid i = 0, j = 0;
  u64 c = 0;
  while (i < na && j < nb) {
    vid x = a[i], y = b[j];
    if (x >= up || y >= up) break;
    if (x <= y) ++i;
    if (y <= x) ++j;
    if (x == y) ++c;
  }
  return c;
}

// VertexSet::operator& (include/VertexSet.h:53-64): materialise A ∩ B.
inline eid isect_set(const vid *a, eid na, const vid *b, eid nb, vid *out) {
  eid i = 0, j = 0, k = 0;
  while (i < na && j < nb) {
    vid x = a[i], y = b[j];
    if (x <= y) ++i;
    if (y <= x) ++j;
    if (x == y) out[k++] = x;
  }
  return k;
}

// number of entries of sorted a[0..n) below up
inline eid below(const vid *a, eid n, vid up) { return std::lower_bound(a, a + n, up) - a; }

struct G {
  eid n;
  const eid *rp;
  const vid *ci;
  const vid *N(eid v) const { return ci + rp[v]; }
  eid d(eid v) const { return rp[v + 1] - rp[v]; }
};

eid max_degree(const G &g) {
  eid m = 0;
  for (eid v = 0; v < g.n; ++v) m = std::max(m, g.d(v));
  return m;
}

void set_threads(int threads) {
  if (threads > 0) omp_set_num_threads(threads);
}

}  // namespace

extern "C" {

// ---- set-operation primitives (pinned against std::set_* by brute force) ----

u64 orc_intersect_num(const vid *a, eid na, const vid *b, eid nb) { return isect_num(a, na, b, nb); }

u64 orc_intersect_num_bound(const vid *a, eid na, const vid *b, eid nb, vid up) {
  return isect_num_bound(a, na, b, nb, up);
}

eid orc_intersect_set(const vid *a, eid na, const vid *b, eid nb, vid *out) {
  return isect_set(a, na, b, nb, out);
}

// GraphMiner difference: |{x in A : x < up, x not in B, x != vid_b}|, the
// vid(B) exclusion of VertexSet::difference_ns (src/common/VertexSet.cc:69-88).
u64 orc_difference_num(const vid *a, eid na, const vid *b, eid nb, vid up, vid vid_b) {
  eid i = 0, j = 0;
  u64 c = 0;
  while (i < na && j < nb) {
    vid x = a[i], y = b[j];
    if (x >= up || y >= up) break;
    if (x <= y) ++i;
    if (y <= x) ++j;
    if (x < y && x != vid_b) ++c;
  }
  while (i < na) {
    vid x = a[i];
    if (x >= up) break;
    ++i;
    if (x != vid_b) ++c;
  }
  return c;
}

// GLUMIN GPU difference: |{x in A : x < up, x not in B}| with no owner
// exclusion (xgminer/include/glumin/set_difference.cuh:20-35, 110-129).
u64 orc_difference_num_noexcl(const vid *a, eid na, const vid *b, eid nb, vid up) {
  u64 c = 0;
  eid j = 0;
  for (eid i = 0; i < na && a[i] < up; ++i) {
    while (j < nb && b[j] < a[i]) ++j;
    if (!(j < nb && b[j] == a[i])) ++c;
  }
  return c;
}

// ---- orientation: Graph::orientation (src/common/graph.cc:233-279) ----
// out_rp has n+1 entries; out_ci has room for nnz/2 entries.  Returns the
// number of DAG edges.
eid orc_orient(eid n, const eid *rp, const vid *ci, eid *out_rp, vid *out_ci) {
  std::vector<eid> deg(n), nd(n, 0);
  for (eid v = 0; v < n; ++v) deg[v] = rp[v + 1] - rp[v];
  auto keep = [&](eid s, vid t) { return deg[t] > deg[s] || (deg[t] == deg[s] && t > s); };
  for (eid s = 0; s < n; ++s)
    for (eid k = rp[s]; k < rp[s + 1]; ++k) nd[s] += keep(s, ci[k]);
  out_rp[0] = 0;
  for (eid s = 0; s < n; ++s) out_rp[s + 1] = out_rp[s] + nd[s];
  for (eid s = 0; s < n; ++s) {
    eid o = out_rp[s];
    for (eid k = rp[s]; k < rp[s + 1]; ++k)
      if (keep(s, ci[k])) out_ci[o++] = ci[k];
  }
  return out_rp[n];
}

// ---- patterns.  Vertex loop over [v_begin, v_end) (the whole graph when
// v_end < 0), or over the explicit task list set by orc_set_tasks();
// `threads` <= 0 keeps the OpenMP default; *bytes (optional) receives the
// algorithmic byte total. ----

namespace {
const vid *g_tasks = nullptr;  // explicit vertex tasks (orc_set_tasks), or null
eid g_ntasks = 0;
inline eid loop_len(eid b, eid e) { return g_tasks ? g_ntasks : e - b; }
inline eid loop_vertex(eid b, eid i) { return g_tasks ? (eid)g_tasks[i] : b + i; }
}  // namespace

// Restrict the pattern loops to tasks[0..ntasks) (a sample of the vertex
// tasks, e.g. every S-th vertex in decreasing-work order) until called with
// tasks = null.  The array must outlive the counts that use it.
void orc_set_tasks(const vid *tasks, eid ntasks) {
  g_tasks = ntasks > 0 ? tasks : nullptr;
  g_ntasks = ntasks > 0 ? ntasks : 0;
}

// TCSolver (src/triangle/omp_base.cc:15-21) on a DAG.
u64 orc_tc(eid n, const eid *rp, const vid *ci, eid v_begin, eid v_end, int threads, u64 *bytes) {
  G g{n, rp, ci};
  if (v_end < 0 || v_end > n) v_end = n;
  set_threads(threads);
  u64 count = 0, B = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : count, B)
  for (eid i = 0; i < loop_len(v_begin, v_end); ++i) {
    const eid u = loop_vertex(v_begin, i);
    const vid *yu = g.N(u);
    const eid du = g.d(u);
    for (eid k = 0; k < du; ++k) {
      vid v = yu[k];
      count += isect_num(yu, du, g.N(v), g.d(v));
      B += 40 + 4 * (u64)(du + g.d(v));
    }
  }
  if (bytes) *bytes = B;
  return count;
}

// automine_4clique (src/clique/cpu_kernels/automine_omp.h:67-83) on a DAG.
u64 orc_clique4(eid n, const eid *rp, const vid *ci, eid v_begin, eid v_end, int threads, u64 *bytes) {
  G g{n, rp, ci};
  if (v_end < 0 || v_end > n) v_end = n;
  set_threads(threads);
  const eid md = max_degree(g);
  u64 count = 0, B = 0;
#pragma omp parallel reduction(+ : count, B)
  {
    std::vector<vid> y01(md + 1);
#pragma omp for schedule(dynamic, 1)
    for (eid i = 0; i < loop_len(v_begin, v_end); ++i) {
      const eid v0 = loop_vertex(v_begin, i);
      const vid *y0 = g.N(v0);
      const eid d0 = g.d(v0);
      for (eid k = 0; k < d0; ++k) {
        vid v1 = y0[k];
        eid s = isect_set(y0, d0, g.N(v1), g.d(v1), y01.data());
        B += 40 + 4 * (u64)(d0 + g.d(v1));
        for (eid i = 0; i < s; ++i) {
          vid v2 = y01[i];
          count += isect_num(y01.data(), s, g.N(v2), g.d(v2));
          B += 4 * (u64)(s + g.d(v2));
        }
      }
    }
  }
  if (bytes) *bytes = B;
  return count;
}

// automine_5clique (src/clique/cpu_kernels/automine_omp.h:138-157) on a DAG.
u64 orc_clique5(eid n, const eid *rp, const vid *ci, eid v_begin, eid v_end, int threads, u64 *bytes) {
  G g{n, rp, ci};
  if (v_end < 0 || v_end > n) v_end = n;
  set_threads(threads);
  const eid md = max_degree(g);
  u64 count = 0, B = 0;
#pragma omp parallel reduction(+ : count, B)
  {
    std::vector<vid> y12(md + 1), y123(md + 1);
#pragma omp for schedule(dynamic, 1)
    for (eid i = 0; i < loop_len(v_begin, v_end); ++i) {
      const eid v1 = loop_vertex(v_begin, i);
      const vid *y1 = g.N(v1);
      const eid d1 = g.d(v1);
      for (eid k = 0; k < d1; ++k) {
        vid v2 = y1[k];
        eid s2 = isect_set(y1, d1, g.N(v2), g.d(v2), y12.data());
        B += 40 + 4 * (u64)(d1 + g.d(v2));
        for (eid i = 0; i < s2; ++i) {
          vid v3 = y12[i];
          eid s3 = isect_set(y12.data(), s2, g.N(v3), g.d(v3), y123.data());
          B += 4 * (u64)(s2 + g.d(v3));
          for (eid j = 0; j < s3; ++j) {
            vid v4 = y123[j];
            count += isect_num(y123.data(), s3, g.N(v4), g.d(v4));
            B += 4 * (u64)(s3 + g.d(v4));
          }
        }
      }
    }
  }
  if (bytes) *bytes = B;
  return count;
}

// Edge-induced diamond (src/sgl/cpu_kernels/diamond.h:1-14) on the
// undirected graph: sum over v1 < v0 of C(|N(v0) ∩ N(v1)|, 2).
u64 orc_diamond(eid n, const eid *rp, const vid *ci, eid v_begin, eid v_end, int threads, u64 *bytes) {
  G g{n, rp, ci};
  if (v_end < 0 || v_end > n) v_end = n;
  set_threads(threads);
  u64 count = 0, B = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : count, B)
  for (eid i = 0; i < loop_len(v_begin, v_end); ++i) {
    const eid v0 = loop_vertex(v_begin, i);
    const vid *y0 = g.N(v0);
    const eid d0 = g.d(v0);
    for (eid k = 0; k < d0; ++k) {
      vid v1 = y0[k];
      if (v1 >= v0) break;
      u64 c = isect_num(y0, d0, g.N(v1), g.d(v1));
      count += c * (c - (c > 0)) / 2;  // C(c,2); c*(c-1)/2 with c=0 -> 0
      B += 40 + 4 * (u64)(d0 + g.d(v1));
    }
  }
  if (bytes) *bytes = B;
  return count;
}

// Edge-induced 4-cycle (src/sgl/cpu_kernels/rectangle.h:1-11) on the
// undirected graph: for v1 < v0 in N(v0), v2 < v1 in N(v0), count
// |N(v1) ∩ N(v2) ∩ [0, v0)|.
u64 orc_cycle4(eid n, const eid *rp, const vid *ci, eid v_begin, eid v_end, int threads, u64 *bytes) {
  G g{n, rp, ci};
  if (v_end < 0 || v_end > n) v_end = n;
  set_threads(threads);
  u64 count = 0, B = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : count, B)
  for (eid i = 0; i < loop_len(v_begin, v_end); ++i) {
    const eid v0 = loop_vertex(v_begin, i);
    const vid *y0 = g.N(v0);
    const eid d0 = g.d(v0);
    for (eid k1 = 0; k1 < d0; ++k1) {
      vid v1 = y0[k1];
      if (v1 >= v0) break;
      const eid b1 = below(g.N(v1), g.d(v1), (vid)v0);
      B += 40;
      for (eid k2 = 0; k2 < d0; ++k2) {
        vid v2 = y0[k2];
        if (v2 >= v1) break;
        count += isect_num_bound(g.N(v1), g.d(v1), g.N(v2), g.d(v2), (vid)v0);
        B += 4 * (u64)(b1 + below(g.N(v2), g.d(v2), (vid)v0));
      }
    }
  }
  if (bytes) *bytes = B;
  return count;
}


// Edge-induced 4-cycles by wedge aggregation -- a second, independent CPU
// algorithm (not a restatement of rectangle.h) for golden counts on graphs
// where the reference loop order takes hours.  Vertices ranked by ascending
// (degree, id); every 4-cycle is counted once at its top-ranked vertex x
// and the opposite vertex w: C(c, 2) for c = #{v in N(x) ∩ N(w) : rank v <
// rank x}, summed over w ranked below x.  Pinned against orc_cycle4 on every
// small golden graph, brute force and closed forms (tests/test_oracle.py).
uint64_t orc_cycle4_wedge(int64_t n, const int64_t *rp, const vid *ci, int threads) {
  set_threads(threads);
  std::vector<int64_t> rank(n), byrank(n);
  for (int64_t v = 0; v < n; ++v) byrank[v] = v;
  std::stable_sort(byrank.begin(), byrank.end(),
                   [&](int64_t a, int64_t b) { return rp[a + 1] - rp[a] < rp[b + 1] - rp[b]; });
  for (int64_t r = 0; r < n; ++r) rank[byrank[r]] = r;
  // relabelled CSR, lists sorted by rank
  std::vector<int64_t> rrp(n + 1, 0);
  for (int64_t r = 0; r < n; ++r) rrp[r + 1] = rrp[r] + (rp[byrank[r] + 1] - rp[byrank[r]]);
  std::vector<vid> rci(rrp[n]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < n; ++r) {
    const int64_t v = byrank[r];
    vid *dst = rci.data() + rrp[r];
    for (int64_t k = rp[v]; k < rp[v + 1]; ++k) dst[k - rp[v]] = (vid)rank[ci[k]];
    std::sort(dst, dst + (rp[v + 1] - rp[v]));
  }
  u64 count = 0;
#pragma omp parallel reduction(+ : count)
  {
    std::vector<vid> cnt(n, 0), touched;
#pragma omp for schedule(dynamic, 64)
    for (int64_t x = 0; x < n; ++x) {
      touched.clear();
      for (int64_t a = rrp[x]; a < rrp[x + 1] && rci[a] < x; ++a) {
        const vid v = rci[a];
        for (int64_t b = rrp[v]; b < rrp[v + 1] && rci[b] < x; ++b) {
          const vid w = rci[b];
          if (cnt[w]++ == 0) touched.push_back(w);
        }
      }
      for (vid w : touched) {
        const u64 c = (u64)cnt[w];
        count += c * (c - 1) / 2;
        cnt[w] = 0;
      }
    }
  }
  return count;
}


// Vertex-induced k-motif census by the reference's formula path:
// automine_3motif / automine_4motif (src/motif/cpu_kernels/automine_formula.h:1-58)
// and the post-processing of MotifSolver (src/motif/omp_formula.cc:37-47).
// k = 3: out = {wedges, triangles}; k = 4: out = {3-stars, 4-paths,
// tailed triangles, 4-cycles, diamonds, 4-cliques} (the reference's order).
// GraphMiner set semantics: difference_set drops the owner vid of B
// (VertexSet.cc:21-67), bounded / intersection with `up` keep x < up.
int orc_motif(eid n, const eid *rp, const vid *ci, int k, int threads, u64 *out) {
  if (k != 3 && k != 4) return 1;
  G g{n, rp, ci};
  set_threads(threads);
  const eid md = max_degree(g);
  u64 c[6] = {0, 0, 0, 0, 0, 0};
  u64 c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
#pragma omp parallel reduction(+ : c0, c1, c2, c3, c4, c5)
  {
    std::vector<vid> W(md), N01(md), D(md);
#pragma omp for schedule(dynamic, 1)
    for (eid v0 = 0; v0 < n; ++v0) {
      const vid *y0 = g.N(v0);
      const eid d0 = g.d(v0);
      const eid f0 = below(y0, d0, (vid)v0);  // y0f0 = bounded(y0, v0)
      if (k == 3) c0 += (u64)d0 * (u64)(d0 - 1);
      for (eid i1 = 0; i1 < f0; ++i1) {
        const vid v1 = y0[i1];
        const vid *y1 = g.N(v1);
        const eid d1 = g.d(v1);
        if (k == 3) {
          c1 += isect_num_bound(y0, d0, y1, d1, v1);
          continue;
        }
        const u64 tri = isect_num(y0, d0, y1, d1);
        c4 += tri * (tri - 1);
        const u64 staru = (u64)d0 - tri - 1, starv = (u64)d1 - tri - 1;
        c2 += tri * (staru + starv);
        c1 += staru * starv;
        c0 += staru * (staru - 1);
        c0 += starv * (starv - 1);
        // y0f0y1f1 = intersection_set(y0f0, y1, v1)
        eid nw = 0;
        for (eid a = 0, b = 0; a < f0 && b < d1;) {
          const vid x = y0[a], y = y1[b];
          if (x >= v1 || y >= v1) break;
          if (x <= y) ++a;
          if (y <= x) ++b;
          if (x == y) W[nw++] = x;
        }
        // n0f0y1 = difference_set(y1, y0): y1 \ y0, without vid(y0) = v0
        eid nn = 0;
        {
          eid a = 0, b = 0;
          while (a < d1 && b < d0) {
            const vid x = y1[a], y = y0[b];
            if (x <= y) ++a;
            if (y <= x) ++b;
            if (x < y && x != (vid)v0) N01[nn++] = x;
          }
          for (; a < d1; ++a)
            if (y1[a] != (vid)v0) N01[nn++] = y1[a];
        }
        // y0f0n1f1 = difference_set(y0f0, y1, v1): x < v1, x not in y1, x != v1
        eid nd = 0;
        {
          eid a = 0, b = 0;
          bool stop = false;
          while (a < f0 && b < d1) {
            const vid x = y0[a], y = y1[b];
            if (x >= v1 || y >= v1) { stop = x >= v1; break; }
            if (x <= y) ++a;
            if (y <= x) ++b;
            if (x < y && x != v1) D[nd++] = x;
          }
          if (!stop)
            for (; a < f0; ++a) {
              if (y0[a] >= v1) break;
              if (y0[a] != v1) D[nd++] = y0[a];
            }
        }
        for (eid i2 = 0; i2 < nw; ++i2) {
          const vid v2 = W[i2];
          c5 += isect_num_bound(W.data(), nw, g.N(v2), g.d(v2), v2);
        }
        for (eid i2 = 0; i2 < nd; ++i2) {
          const vid v2 = D[i2];
          c3 += isect_num_bound(N01.data(), nn, g.N(v2), g.d(v2), (vid)v0);
        }
      }
    }
  }
  c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3; c[4] = c4; c[5] = c5;
  if (k == 3) {
    out[0] = c[0] / 2 - 3 * c[1];
    out[1] = c[1];
  } else {
    c[4] = c[4] / 2 - c[5] * 6;
    c[2] = c[2] / 2 - c[4] * 2;
    c[1] = c[1] - c[3] * 4;
    c[0] = c[0] / 6 - c[2] / 3;
    for (int i = 0; i < 6; ++i) out[i] = c[i];
  }
  return 0;
}


// k-cliques (k = 4, 5) of a DAG by bitset sub-DAGs -- a second CPU
// algorithm (not a restatement of automine_4clique / automine_5clique) for
// golden counts where the reference loop order takes CPU-days.  For each v0
// the out-neighbours X = N+(v0) (sorted) get rows row_i = {j : X[j] in
// N+(X[i])} as 64-bit words; K4(v0) = sum_{i, j in row_i} popc(row_i &
// row_j), K5(v0) adds one more level through l in row_i & row_j.  Pinned
// against the restatement, the reference's counts, brute force and closed
// forms (tests/test_oracle.py).  Vertices [v_begin, v_end) only.
uint64_t orc_clique_bitset(int64_t n, const int64_t *rp, const vid *ci, int k, int64_t v_begin, int64_t v_end,
                           int threads) {
  if (k != 4 && k != 5) return 0;
  G g{n, rp, ci};
  if (v_end < 0 || v_end > n) v_end = n;
  set_threads(threads);
  const eid md = max_degree(g);
  const eid Wmax = (md + 63) / 64;
  u64 count = 0;
#pragma omp parallel reduction(+ : count)
  {
    std::vector<u64> rows((size_t)std::max<eid>(1, md) * std::max<eid>(1, Wmax)), t1(Wmax);
#pragma omp for schedule(dynamic, 1)
    for (eid v0 = v_begin; v0 < v_end; ++v0) {
      const vid *X = g.N(v0);
      const eid s = g.d(v0);
      if (s < k - 1) continue;
      const eid W = (s + 63) / 64;
      std::fill(rows.begin(), rows.begin() + s * W, 0ull);
      for (eid i = 0; i < s; ++i) {
        const vid *Y = g.N(X[i]);
        const eid dy = g.d(X[i]);
        u64 *ri = rows.data() + i * W;
        for (eid a = 0, b = 0; a < s && b < dy;) {  // positions of X ∩ N+(X[i]) in X
          if (X[a] < Y[b]) ++a;
          else if (Y[b] < X[a]) ++b;
          else { ri[a >> 6] |= 1ull << (a & 63); ++a; ++b; }
        }
      }
      for (eid i = 0; i < s; ++i) {
        const u64 *ri = rows.data() + i * W;
        for (eid wj = 0; wj < W; ++wj)
          for (u64 bj = ri[wj]; bj; bj &= bj - 1) {
            const eid j = wj * 64 + __builtin_ctzll(bj);
            const u64 *rj = rows.data() + j * W;
            if (k == 4) {
              for (eid w = 0; w < W; ++w) count += __builtin_popcountll(ri[w] & rj[w]);
              continue;
            }
            for (eid w = 0; w < W; ++w) t1[w] = ri[w] & rj[w];
            for (eid wl = 0; wl < W; ++wl)
              for (u64 bl = t1[wl]; bl; bl &= bl - 1) {
                const u64 *rl = rows.data() + (wl * 64 + __builtin_ctzll(bl)) * W;
                for (eid w = 0; w < W; ++w) count += __builtin_popcountll(t1[w] & rl[w]);
              }
          }
      }
    }
  }
  return count;
}

}  // extern "C"
