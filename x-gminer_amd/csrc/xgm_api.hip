// C-ABI layer of libxgm: contexts, device graphs, task preparation and the
// pattern dispatch.  Declared in include/xgm.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/xgm.h"
#include "host_util.h"
#include "kernels.h"

using namespace xgm;

namespace {

constexpr int kScratchU64 = 64;   // accumulators
constexpr int kScratchQueues = 64;

}  // namespace

struct xgm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  std::string err;
  u64 *d_out = nullptr;
  int *d_queues = nullptr;
  int32_t *d_noalloc = nullptr;  // this device's static staging array (xgm_setop_pairs_noalloc)
  int prep = XGM_PREP_AUTO;      // where task lists are built (xgm_ctx_set_prep)
  std::vector<hipEvent_t> ev;  // 2 per profiled launch
  xgm_profile prof{};
  int n_marks = 0;

  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
  int hip_fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return XGM_EHIP;
  }
};

namespace {
constexpr int kClqLists = 13;
constexpr int kBinCap[4] = {64, 256, 512, 1024};
}  // namespace

// Lists of a clique schedule: 0-3 id-ordered LUT bins, 4-7 rank-ordered
// LUT bins, 8-9 edge LUT (min out-degree <= 128 / larger), 10-11 edge list,
// 12 rank-ordered LUT for hubs (1024 < d+ <= kHubLutMax; rows in LDS up to
// kTriLutMax, in HBM above).

struct xgm_sched {
  int pattern = 0, k = 0;
  int device = 0;
  int32_t *d_buf = nullptr;
  int64_t off[kClqLists] = {}, cnt[kClqLists] = {};  // in items; edge lists hold 2 ints per item
  std::vector<int32_t> owner[kClqLists];              // item -> vertex (for the timing log)
  ~xgm_sched() {
    (void)hipSetDevice(device);
    (void)hipFree(d_buf);
  }
  CliqueTasks tasks(u64 *log_base) const {
    CliqueTasks ct;
    for (int i = 0; i < 4; ++i) {
      ct.lut[i] = TaskList{d_buf + off[i], cnt[i]};
      ct.rlut[i] = TaskList{d_buf + off[4 + i], cnt[4 + i]};
    }
    for (int i = 0; i < 2; ++i) {
      ct.elut[i] = EdgeItems{d_buf + off[8 + i], cnt[8 + i]};
      ct.list[i] = EdgeItems{d_buf + off[10 + i], cnt[10 + i]};
    }
    ct.rlut_hbm = TaskList{d_buf + off[12], cnt[12]};
    if (log_base) {
      int64_t o = 0;
      u64 **dst[kClqLists] = {&ct.log_lut[0], &ct.log_lut[1], &ct.log_lut[2], &ct.log_lut[3],
                              &ct.log_rlut[0], &ct.log_rlut[1], &ct.log_rlut[2], &ct.log_rlut[3],
                              &ct.log_elut[0], &ct.log_elut[1], &ct.log_list[0], &ct.log_list[1],
                              &ct.log_rlut_hbm};
      for (int i = 0; i < kClqLists; ++i) {
        *dst[i] = log_base + o;
        o += cnt[i];
      }
    }
    return ct;
  }
  int64_t items() const {
    int64_t t = 0;
    for (int i = 0; i < kClqLists; ++i) t += cnt[i];
    return t;
  }
};

struct xgm_dgraph {
  xgm_ctx *ctx = nullptr;
  int device = 0;
  int64_t n = 0, nnz = 0;
  uint32_t flags = 0;
  int64_t *d_rowptr = nullptr;
  int32_t *d_colidx = nullptr;
  int64_t *d_rrowptr = nullptr;  // reverse DAG (in-neighbour lists), DAG graphs only
  int32_t *d_rcolidx = nullptr;
  std::vector<int64_t> h_rrowptr;
  std::vector<int32_t> h_rcolidx;
  int max_out = 0;
  std::vector<int64_t> h_rowptr;  // host copy of the CSR (task preparation)
  std::vector<int32_t> h_colidx;
  // DAG vertex tasks (d+ >= 2) by decreasing work estimate, and d+ per vertex
  std::vector<int32_t> order;
  std::vector<int32_t> dplus;
  // device task lists, cached per key (pattern family / strategy / threshold)
  std::map<std::string, std::pair<int32_t *, std::vector<int64_t>>> task_cache;
  int *d_scratch = nullptr;
  size_t scratch_ints = 0;
  // undirected graphs: degree per vertex; (degree, id)-relabelled copy with
  // per-vertex wedge work, built on first 4-cycle count
  std::vector<int64_t> work;
  bool relabelled = false;
  std::map<std::string, std::unique_ptr<xgm_sched>> scheds;  // cached clique schedules
  u64 *d_hub_lut = nullptr;        // per-block HBM LUT rows for clique hubs
  int *d_dense = nullptr;          // per-block 4-cycle counters, zero between counts
  size_t dense_ints = 0;
  int64_t *d_rl_rowptr = nullptr;
  int32_t *d_rl_colidx = nullptr;
  int32_t *d_rl_cut = nullptr;     // per entry: |N(v) ∩ [0, x)| (DevCSR::cut)
  int32_t *d_rl_nlow = nullptr;    // per vertex: |N(x) ∩ [0, x)|
  std::vector<int32_t> rl_order;   // relabelled task ids by decreasing wedge work
  std::vector<int64_t> rl_work;
  std::vector<int32_t> rl_nlow;    // lower-neighbour counts (relabelled in HBM; else empty)
  std::vector<int64_t> rl_rowptr;  // host copy of the relabelled CSR
  std::vector<int32_t> rl_colidx;
  int *d_hub = nullptr;            // 4-cycle hub counter rows (hub_rows * stride ints)
  // bucket-bitmap triangle strategy: bucket-major colidx, bucket offsets
  // (64 per vertex), non-empty bucket masks; built on first use
  int32_t *d_bcol = nullptr;
  uint16_t *d_boff = nullptr;
  u64 *d_bmap = nullptr;
  std::map<std::string, int> hub_count;  // hub tasks per cached hub list
  std::map<std::string, int64_t> range_nlow;  // largest lower-neighbour count per cached range list
  std::map<std::string, int64_t> list_md;     // largest degree of the list tasks per cached motif list
  int4 *d_cursor = nullptr;        // 4-cycle range items: per-block list cursors
  size_t cursor_n = 0;
  // scratch buffers replaced by larger ones: a plan captured earlier still
  // launches kernels on them, so they live as long as the graph
  std::vector<void *> retired;
  // batched instances (disjoint union): instance of every vertex, first id
  // of every instance (n_inst + 1 bounds), largest instance
  int n_inst = 1;
  std::vector<int32_t> h_slot;
  std::vector<int64_t> inst_base;
  int64_t max_inst_n = 0;
  int32_t *d_slot = nullptr;
  int64_t *d_inst_lo = nullptr;
  u64 *d_inst_out = nullptr;       // per-instance counts
  u64 *d_motif = nullptr;          // per-instance 4-motif edge sums (4 per instance)
  std::unique_ptr<xgm_dgraph> dag; // undirected graphs: degree-oriented copy for motif censuses
  size_t hub_ints = 0;
  // plans whose captured kernels read this graph's buffers: xgm_graph_free
  // defers the release until the last of them is freed
  int plan_refs = 0;
  bool free_pending = false;
  // cliques under XGM_STRAT_AUTO use the rank-ordered LUT until a count
  // finds the DAG outside (degree, id) orientation; then the id-ordered LUT
  bool rank_ok = true;
  ~xgm_dgraph() {
    (void)hipFree(d_dense);
    (void)hipFree(d_cursor);
    (void)hipFree(d_hub);
    (void)hipFree(d_bcol);
    (void)hipFree(d_boff);
    (void)hipFree(d_bmap);
    (void)hipFree(d_hub_lut);
    (void)hipFree(d_slot);
    (void)hipFree(d_inst_lo);
    (void)hipFree(d_inst_out);
    (void)hipFree(d_motif);
    (void)hipFree(d_rl_rowptr);
    (void)hipFree(d_rl_colidx);
    (void)hipFree(d_rl_cut);
    (void)hipFree(d_rl_nlow);
    for (auto &kv : task_cache) (void)hipFree(kv.second.first);
    (void)hipFree(d_rowptr);
    (void)hipFree(d_colidx);
    (void)hipFree(d_rrowptr);
    (void)hipFree(d_rcolidx);
    (void)hipFree(d_scratch);
    for (void *x : retired) (void)hipFree(x);
  }
};

static int ensure_host_mirror(xgm_ctx *ctx, xgm_dgraph *g);

// Task lists of g from its host copy (true) or in HBM (false): the host
// builders under XGM_PREP_HOST or, under XGM_PREP_AUTO, when a host copy of
// the CSR is already at hand.
static bool host_prep(const xgm_ctx *ctx, const xgm_dgraph *g) {
  if (ctx->prep == XGM_PREP_HOST) return true;
  if (ctx->prep == XGM_PREP_DEVICE) return false;
  return (int64_t)g->h_rowptr.size() == g->n + 1;
}

namespace {

#define HIP_TRY(ctx, call)                                  \
  do {                                                      \
    hipError_t _e = (call);                                 \
    if (_e != hipSuccess) return (ctx)->hip_fail(_e, #call); \
  } while (0)

Launch make_launch(xgm_ctx *ctx, int slot) {
  Launch l;
  l.stream = ctx->stream;
  l.num_cu = ctx->num_cu;
  l.d_out = ctx->d_out + slot;
  l.d_err = ctx->d_out + (kScratchU64 - 1);
  l.d_queue = ctx->d_queues;
  l.begin = [ctx](const char *name, int64_t tasks) {
    int i = ctx->prof.n_kernels;
    if (i >= XGM_MAX_PROFILE_KERNELS) return;
    std::snprintf(ctx->prof.k[i].name, sizeof(ctx->prof.k[i].name), "%s", name);
    ctx->prof.k[i].tasks = tasks;
    hipEventRecord(ctx->ev[2 * i], ctx->stream);
  };
  l.end = [ctx]() {
    int i = ctx->prof.n_kernels;
    if (i >= XGM_MAX_PROFILE_KERNELS) return;
    hipEventRecord(ctx->ev[2 * i + 1], ctx->stream);
    ctx->prof.n_kernels = i + 1;
  };
  return l;
}

// DAG vertex tasks with d+ >= 2, ordered by decreasing work estimate
// sum_{v in N+(u)} d+(v) (ties by id): dynamic claiming then approximates
// longest-processing-time-first inside every tier.
void build_dag_order(const xgm_csr &g, std::vector<int32_t> &order, std::vector<int32_t> &dplus, int &max_out) {
  const int64_t n = g.n;
  std::vector<int64_t> work(n, 0);
  dplus.assign(n, 0);
  parallel_for(n, 4096, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) {
      int64_t du = g.rowptr[u + 1] - g.rowptr[u];
      dplus[u] = (int32_t)du;
      if (du < 2) continue;
      int64_t w = 0;
      for (int64_t k = g.rowptr[u]; k < g.rowptr[u + 1]; ++k) {
        int32_t v = g.colidx[k];
        w += g.rowptr[v + 1] - g.rowptr[v];
      }
      work[u] = w;
    }
  });
  order.clear();
  max_out = 0;
  for (int64_t u = 0; u < n; ++u) {
    max_out = std::max(max_out, dplus[u]);
    if (dplus[u] >= 2) order.push_back((int32_t)u);
  }
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return work[a] > work[b]; });
}

// Split `order` into consecutive degree ranges (lo_i, hi_i] (inclusive hi)
// for vertices with d+ >= min_deg; cached on the device graph under `key`.
int tasks_by_ranges(xgm_ctx *ctx, xgm_dgraph *g, const std::string &key, int min_deg,
                    const std::vector<int> &hi, std::vector<TaskList> &out) {
  auto it = g->task_cache.find(key);
  if (it == g->task_cache.end()) {
    std::vector<std::vector<int32_t>> parts(hi.size());
    for (int32_t u : g->order) {
      int d = g->dplus[u];
      if (d < min_deg) continue;
      size_t t = 0;
      while (t + 1 < hi.size() && d > hi[t]) ++t;
      parts[t].push_back(u);
    }
    std::vector<int32_t> all;
    std::vector<int64_t> counts;
    for (auto &p : parts) {
      counts.push_back((int64_t)p.size());
      all.insert(all.end(), p.begin(), p.end());
    }
    int32_t *d = nullptr;
    HIP_TRY(ctx, hipMalloc(&d, std::max<size_t>(1, all.size()) * sizeof(int32_t)));
    if (!all.empty()) {
      hipError_t e = hipMemcpy(d, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(d);
        return ctx->hip_fail(e, "H2D tasks");
      }
    }
    it = g->task_cache.emplace(key, std::make_pair(d, counts)).first;
  }
  out.clear();
  int64_t off = 0;
  for (int64_t c : it->second.second) {
    TaskList t;
    t.tasks = it->second.first + off;
    t.count = c;
    out.push_back(t);
    off += c;
  }
  return XGM_OK;
}

// Grow a per-graph scratch buffer to `need` elements of `elem` bytes; sets
// *grown when it allocated.  The old buffer is retired, not freed: a plan
// captured earlier still launches kernels on it.
int grow_scratch(xgm_ctx *ctx, xgm_dgraph *g, void **buf, size_t &have, size_t need, size_t elem,
                 bool *grown = nullptr) {
  if (grown) *grown = false;
  if (*buf && have >= need) return XGM_OK;
  if (*buf) g->retired.push_back(*buf);
  *buf = nullptr;
  have = 0;
  HIP_TRY(ctx, hipMalloc(buf, std::max<size_t>(1, need) * elem));
  have = need;
  if (grown) *grown = true;
  return XGM_OK;
}

int ensure_scratch(xgm_ctx *ctx, xgm_dgraph *g, size_t ints) {
  return grow_scratch(ctx, g, (void **)&g->d_scratch, g->scratch_ints, ints, sizeof(int));
}

// Undirected task order: vertices with degree >= 3 by decreasing owned
// probe work sum_{p in N(t), (deg p, p) < (deg t, t)} deg(p).
void build_undirected_order(const xgm_csr &g, std::vector<int32_t> &order, std::vector<int32_t> &deg,
                            std::vector<int64_t> &work) {
  const int64_t n = g.n;
  deg.assign(n, 0);
  work.assign(n, 0);
  for (int64_t u = 0; u < n; ++u) deg[u] = (int32_t)(g.rowptr[u + 1] - g.rowptr[u]);
  parallel_for(n, 4096, [&](int64_t b, int64_t e) {
    for (int64_t t = b; t < e; ++t) {
      int64_t w = 0;
      for (int64_t k = g.rowptr[t]; k < g.rowptr[t + 1]; ++k) {
        int32_t p = g.colidx[k];
        if (deg[p] < deg[t] || (deg[p] == deg[t] && p < t)) w += deg[p];
      }
      work[t] = w;
    }
  });
  order.clear();
  for (int64_t u = 0; u < n; ++u)
    if (deg[u] >= 3) order.push_back((int32_t)u);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return work[a] > work[b]; });
}

// Relabel vertices by ascending (degree, id) inside each id range
// [bounds[i], bounds[i+1]) (one range per instance of a batch; empty bounds =
// the whole graph), so every instance keeps its contiguous ids: the returned
// CSR lists are sorted in the new ids; wedge work of x = sum over lower
// neighbours v < x of |N(v) ∩ [0, x)|.  Counts of isomorphism-invariant
// patterns are unchanged.
void relabel_by_degree(const xgm_csr &g, const std::vector<int64_t> &bounds, std::vector<int64_t> &rp,
                       std::vector<int32_t> &ci, std::vector<int32_t> &order, std::vector<int64_t> &wwork) {
  const int64_t n = g.n;
  std::vector<int32_t> rank(n), byrank(n);
  std::iota(byrank.begin(), byrank.end(), 0);
  auto deg = [&](int32_t v) { return g.rowptr[v + 1] - g.rowptr[v]; };
  const std::vector<int64_t> whole{0, n};
  const auto &bd = bounds.empty() ? whole : bounds;
  parallel_for((int64_t)bd.size() - 1, 1, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      std::stable_sort(byrank.begin() + bd[i], byrank.begin() + bd[i + 1],
                       [&](int32_t a, int32_t c) { return deg(a) < deg(c); });
  });
  for (int64_t r = 0; r < n; ++r) rank[byrank[r]] = (int32_t)r;
  std::vector<int64_t> cnt(n);
  for (int64_t r = 0; r < n; ++r) cnt[r] = deg(byrank[r]);
  exclusive_scan(cnt, rp);
  ci.resize(g.nnz);
  parallel_for(n, 2048, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; ++r) {
      int32_t v = byrank[r];
      int32_t *dst = ci.data() + rp[r];
      int64_t k0 = g.rowptr[v], k1 = g.rowptr[v + 1];
      for (int64_t k = k0; k < k1; ++k) dst[k - k0] = rank[g.colidx[k]];
      std::sort(dst, dst + (k1 - k0));
    }
  });
  wwork.assign(n, 0);
  parallel_for(n, 2048, [&](int64_t b, int64_t e) {
    for (int64_t x = b; x < e; ++x) {
      const int32_t *nx = ci.data() + rp[x];
      int64_t lo = std::lower_bound(nx, nx + (rp[x + 1] - rp[x]), (int32_t)x) - nx;
      if (lo < 2) continue;
      int64_t w = 0;
      for (int64_t k = 0; k < lo; ++k) {
        int32_t v = nx[k];
        const int32_t *nv = ci.data() + rp[v];
        w += std::lower_bound(nv, nv + (rp[v + 1] - rp[v]), (int32_t)x) - nv;
      }
      wwork[x] = w;
    }
  });
  order.clear();
  for (int64_t x = 0; x < n; ++x)
    if (wwork[x] > 0) order.push_back((int32_t)x);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return wwork[a] > wwork[b]; });
}

// Upload a task list split by thresholds on `key_of(u)` (inclusive upper
// bounds, last = everything else), cached under `key`.
template <class KeyOf>
int tasks_split(xgm_ctx *ctx, xgm_dgraph *g, const std::string &key, const std::vector<int32_t> &order,
                KeyOf key_of, const std::vector<int64_t> &hi, std::vector<TaskList> &out) {
  auto it = g->task_cache.find(key);
  if (it == g->task_cache.end()) {
    std::vector<std::vector<int32_t>> parts(hi.size());
    for (int32_t u : order) {
      int64_t kv = key_of(u);
      size_t t = 0;
      while (t + 1 < hi.size() && kv > hi[t]) ++t;
      parts[t].push_back(u);
    }
    std::vector<int32_t> all;
    std::vector<int64_t> counts;
    for (auto &p : parts) {
      counts.push_back((int64_t)p.size());
      all.insert(all.end(), p.begin(), p.end());
    }
    int32_t *d = nullptr;
    HIP_TRY(ctx, hipMalloc(&d, std::max<size_t>(1, all.size()) * sizeof(int32_t)));
    if (!all.empty()) {
      hipError_t e = hipMemcpy(d, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(d);
        return ctx->hip_fail(e, "H2D tasks");
      }
    }
    it = g->task_cache.emplace(key, std::make_pair(d, counts)).first;
  }
  out.clear();
  int64_t off = 0;
  for (int64_t c : it->second.second) {
    TaskList t;
    t.tasks = it->second.first + off;
    t.count = c;
    out.push_back(t);
    off += c;
  }
  return XGM_OK;
}

// Upload parts (each `stride` ints per item) as one cached device array;
// returns one TaskList per part (count = items).
int cache_parts(xgm_ctx *ctx, xgm_dgraph *g, const std::string &key, const std::vector<std::vector<int32_t>> &parts,
                int stride, std::vector<TaskList> &out) {
  auto it = g->task_cache.find(key);
  if (it == g->task_cache.end()) {
    std::vector<int32_t> all;
    std::vector<int64_t> counts;
    for (auto &p : parts) {
      counts.push_back((int64_t)p.size() / stride);
      all.insert(all.end(), p.begin(), p.end());
    }
    int32_t *d = nullptr;
    HIP_TRY(ctx, hipMalloc(&d, std::max<size_t>(4, all.size()) * sizeof(int32_t)));
    if (!all.empty()) {
      hipError_t e = hipMemcpy(d, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(d);
        return ctx->hip_fail(e, "H2D items");
      }
    }
    it = g->task_cache.emplace(key, std::make_pair(d, counts)).first;
  }
  out.clear();
  int64_t off = 0;
  for (int64_t c : it->second.second) {
    TaskList t;
    t.tasks = it->second.first + off;
    t.count = c;
    out.push_back(t);
    off += c * stride;
  }
  return XGM_OK;
}

bool cached(const xgm_dgraph *g, const std::string &key) { return g->task_cache.count(key) != 0; }

// Reverse (in-neighbour) lists of a DAG, each sorted.
void build_reverse(const std::vector<int64_t> &rp, const std::vector<int32_t> &ci, int64_t n,
                   std::vector<int64_t> &rrp, std::vector<int32_t> &rci) {
  std::vector<int64_t> indeg(n, 0);
  parallel_for((int64_t)ci.size(), 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k) __atomic_fetch_add(&indeg[ci[k]], 1, __ATOMIC_RELAXED);
  });
  exclusive_scan(indeg, rrp);
  rci.resize(ci.size());
  std::vector<int64_t> pos(rrp.begin(), rrp.end() - 1);
  parallel_for(n, 4096, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u)
      for (int64_t k = rp[u]; k < rp[u + 1]; ++k)
        rci[__atomic_fetch_add(&pos[ci[k]], 1, __ATOMIC_RELAXED)] = (int32_t)u;
  });
  parallel_for(n, 4096, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) std::sort(rci.begin() + rrp[v], rci.begin() + rrp[v + 1]);
  });
}

// Triangle-count items (t, k0, k1, 0): every DAG edge is charged to the
// endpoint with the longer out-list (ties to the head); an item covers the
// positions [k0, k1) of N+(t) ++ N-(t), cut at ~kSlice probes or at the
// wave tiers' segment capacity.  Tiers by d+(t); items ordered by
// decreasing work inside a tier.
void build_tc_items(const xgm_dgraph *g, int max_small, int max_large, std::vector<std::vector<int32_t>> &parts) {
  const int64_t n = g->n;
  const auto &dp = g->dplus;
  const int64_t kSlice = kTcSlice;
  struct It { int64_t work; int32_t t, k0, k1; };
  std::vector<std::vector<It>> tiers(NUM_TIERS);
  for (int64_t t = 0; t < n; ++t) {
    const int d = dp[t];
    if (d < 1) continue;
    const int tier = d <= max_small ? TIER_W64 : d <= max_large ? TIER_W256 : d <= kTierMaxDeg[TIER_B2048] ? TIER_B2048
                                                                                                       : TIER_BIG;
    const int cap = tier < 2 ? kTcWaveCap[tier] : INT32_MAX;
    const int32_t *out = g->h_colidx.data() + g->h_rowptr[t];
    const int32_t *in = g->h_rcolidx.data() + g->h_rrowptr[t];
    const int din = (int)(g->h_rrowptr[t + 1] - g->h_rrowptr[t]);
    const int P = d + din;
    int64_t acc = 0;
    int k0 = 0;
    for (int k = 0; k < P; ++k) {
      const int x = k < d ? out[k] : in[k - d];
      const bool charged = k < d ? dp[x] < d : dp[x] <= d;
      if (charged) acc += dp[x];
      if (acc >= kSlice || k + 1 - k0 >= cap) {
        if (acc) tiers[tier].push_back({acc, (int32_t)t, k0, k + 1});
        k0 = k + 1;
        acc = 0;
      }
    }
    if (acc) tiers[tier].push_back({acc, (int32_t)t, k0, P});
  }
  parts.assign(NUM_TIERS, {});
  for (int t = 0; t < NUM_TIERS; ++t) {
    auto &v = tiers[t];
    std::stable_sort(v.begin(), v.end(), [](const It &a, const It &b) { return a.work > b.work; });
    for (auto &it : v) parts[t].insert(parts[t].end(), {it.t, it.k0, it.k1, 0});
  }
}

// codes[v]: 0 LUT, 1 rank LUT, 2 edge LUT, 3 edge list, 4 rank LUT one bin
// up; vertices with d+ < k-1 are not tasks.  Rank codes on hubs
// (1024 < d+ <= kHubLutMax) use the hub rank LUT; other LUT codes on hubs and
// unknown codes fall back to the edge LUT.
int build_clique_sched(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const int8_t *codes, xgm_sched *sc) {
  const int k = pattern == XGM_CLIQUE4 ? 4 : 5;
  // edge items (codes 2, 3 and LUT codes out of the LUT range) read the
  // forward lists on the host; a DAG built in HBM copies them back only then
  bool edges = false;
  for (int32_t v : g->order) {
    const int d = g->dplus[v];
    if (d < k - 1) continue;
    int c = codes ? codes[v] : 0;
    if (c < 0 || c > 4) c = 2;
    edges |= c == 2 || c == 3 || (d > kBinCap[3] && !((c == 1 || c == 4) && d <= kHubLutMax));
  }
  if (edges)
    if (int rc = ensure_host_mirror(ctx, const_cast<xgm_dgraph *>(g))) return rc;
  sc->pattern = pattern;
  sc->k = k;
  sc->device = ctx->device;
  std::vector<int32_t> L[kClqLists];
  for (int32_t v : g->order) {
    const int d = g->dplus[v];
    if (d < k - 1) continue;
    int c = codes ? codes[v] : 0;
    if (c < 0 || c > 4) c = 2;
    if (d > kBinCap[3] && (c == 1 || c == 4) && d <= kHubLutMax) {
      L[12].push_back(v);
      sc->owner[12].push_back(v);
      continue;
    }
    if (d > kBinCap[3] && (c == 0 || c == 1 || c == 4)) c = 2;
    int bin = 0;
    while (d > kBinCap[bin]) ++bin;
    if (c == 0) { L[bin].push_back(v); sc->owner[bin].push_back(v); continue; }
    if (c == 1 || c == 4) {
      const int b = (c == 4) ? std::min(bin + 1, 3) : bin;
      L[4 + b].push_back(v);
      sc->owner[4 + b].push_back(v);
      continue;
    }
    const int base = (c == 2) ? 8 : 10;
    for (int64_t e = g->h_rowptr[v]; e < g->h_rowptr[v + 1]; ++e) {
      const int32_t v1 = g->h_colidx[e];
      if (g->dplus[v1] < k - 2) continue;
      const int li = base + (std::min(d, g->dplus[v1]) <= 128 ? 0 : 1);
      L[li].push_back(v);
      L[li].push_back(v1);
      sc->owner[li].push_back(v);
    }
  }
  std::vector<int32_t> all;
  for (int i = 0; i < kClqLists; ++i) {
    const int stride = (i >= 8 && i <= 11) ? 2 : 1;  // edge lists hold (v0, v1) pairs
    sc->off[i] = (int64_t)all.size();
    sc->cnt[i] = (int64_t)L[i].size() / stride;
    all.insert(all.end(), L[i].begin(), L[i].end());
  }
  HIP_TRY(ctx, hipMalloc(&sc->d_buf, std::max<size_t>(2, all.size()) * sizeof(int32_t)));
  if (!all.empty()) HIP_TRY(ctx, hipMemcpy(sc->d_buf, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return XGM_OK;
}

// Degree-threshold codes of a strategy (no model).  AUTO and LUT_RANK: the
// rank-ordered LUT for every task up to lut_max out-degree and for hubs
// (1024 < d+ <= kHubLutMax, the hub rank LUT), the edge LUT between lut_max
// and 1024 and above kHubLutMax; AUTO takes the id-ordered LUT for tasks of
// at most 64 out-neighbours, whose packed kernel (8 tasks per wave) runs
// them faster than one rank-LUT block per task.  LUT: the id-ordered LUT up
// to lut_max, the edge LUT above (the reference's fixed-threshold scheme,
// GLUMIN's LUT tasks up to 1024 out-neighbours, xgminer/include/common.h:61).
void heuristic_codes(const xgm_dgraph *g, int k, int strategy, int lut_max, int8_t *codes) {
  for (int64_t v = 0; v < g->n; ++v) {
    const int d = g->dplus[v];
    int8_t c = -1;
    if (d >= k - 1) {
      switch (strategy) {
        case XGM_STRAT_LIST: c = 3; break;
        case XGM_STRAT_ELUT: c = 2; break;
        case XGM_STRAT_LUT: c = d <= lut_max ? 0 : 2; break;
        case XGM_STRAT_AUTO:
          if (d <= kBinCap[0] && d <= lut_max) {
            c = 0;
            break;
          }
          [[fallthrough]];
        default: c = (d <= lut_max || d > 1024) ? 1 : 2; break;  // rank LUT; hubs: hub rank LUT
      }
    }
    codes[v] = c;
  }
}

int run_clique_sched(xgm_ctx *ctx, xgm_dgraph *gm, const xgm_sched *sc, const DevCSR &dg, Launch &l, u64 *log) {
  const int64_t edges = sc->cnt[8] + sc->cnt[9] + sc->cnt[10] + sc->cnt[11];
  if (edges) {
    size_t need = (size_t)clique_edge_grid(ctx->num_cu) * 2 * std::max(1, gm->max_out);
    int rc = ensure_scratch(ctx, gm, need);
    if (rc) return rc;
  }
  if (sc->cnt[12] && !gm->d_hub_lut) {
    const size_t words = (size_t)clique_hub_slots(ctx->num_cu) * ((size_t)kHubLutMax * kHubLutMax / 64);
    HIP_TRY(ctx, hipMalloc(&gm->d_hub_lut, words * sizeof(u64)));
  }
  CliqueTasks ct = sc->tasks(log);
  ct.hub_slots = gm->d_hub_lut;
  hipError_t e = launch_clique(dg, ct, sc->k, gm->d_scratch, std::max(1, gm->max_out), l, l.d_queue);
  if (e != hipSuccess) return ctx->hip_fail(e, "clique launch");
  return XGM_OK;
}

int check_device(xgm_ctx *ctx) {
  hipDeviceProp_t p;
  HIP_TRY(ctx, hipGetDeviceProperties(&p, ctx->device));
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return ctx->fail(XGM_ENODEV, std::string("device is ") + p.gcnArchName + ", need gfx950");
  ctx->num_cu = p.multiProcessorCount;
  return XGM_OK;
}

}  // namespace

namespace {

// A device allocation released at scope exit.
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// Degree orientation of the undirected device graph u, in HBM (k_prep.hip):
// degrees; an out/in flag per entry and its exclusive scan, which place the
// out/in lists and give both row pointers; task sort keys and a stable
// descending radix sort.  The host mirrors that the
// task preparation reads (CSR, reverse CSR, d+, order) are copied back.
// Equal to uploading xgm_orient_dag() of the host CSR.
int orient_device(xgm_ctx *ctx, const xgm_dgraph *u, std::unique_ptr<xgm_dgraph> &out) {
  const int64_t n = u->n;
  const hipStream_t s = ctx->stream;
  const int cu = ctx->num_cu;
  std::unique_ptr<xgm_dgraph> d(new (std::nothrow) xgm_dgraph);
  if (!d) return XGM_ENOMEM;
  d->ctx = ctx;
  d->device = ctx->device;
  d->n = n;
  d->flags = XGM_CSR_SORTED | XGM_CSR_DAG;
  const size_t n1 = (size_t)std::max<int64_t>(1, n);
  const int64_t m = u->nnz;
  DevBuf deg, flag, pos, keys, ids, dplus, tmp;
  HIP_TRY(ctx, hipMalloc(&deg.p, n1 * sizeof(int32_t)));
  HIP_TRY(ctx, hipMalloc(&flag.p, (m + 1) * sizeof(int32_t)));
  HIP_TRY(ctx, hipMalloc(&pos.p, (m + 1) * sizeof(int32_t)));
  HIP_TRY(ctx, hipMalloc(&d->d_rowptr, (n + 1) * sizeof(int64_t)));
  HIP_TRY(ctx, hipMalloc(&d->d_rrowptr, (n + 1) * sizeof(int64_t)));
  HIP_TRY(ctx, launch_prep_degree(u->d_rowptr, n, deg.as<int32_t>(), cu, s));
  HIP_TRY(ctx, launch_prep_split(u->d_rowptr, u->d_colidx, deg.as<int32_t>(), n, m, flag.as<int32_t>(), cu, s));
  size_t scan_bytes = 0, sort_bytes = 0;
  HIP_TRY(ctx, prep_scan(nullptr, scan_bytes, flag.as<int32_t>(), pos.as<int32_t>(), m + 1, s));
  HIP_TRY(ctx, prep_sort(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, n, s));
  size_t tmp_bytes = std::max<size_t>(16, std::max(scan_bytes, sort_bytes));
  HIP_TRY(ctx, hipMalloc(&tmp.p, tmp_bytes));
  {
    size_t b = tmp_bytes;
    HIP_TRY(ctx, prep_scan(tmp.p, b, flag.as<int32_t>(), pos.as<int32_t>(), m + 1, s));
  }
  int32_t n_out = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&n_out, pos.as<int32_t>() + m, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  // In a symmetric graph every edge {u, w} is one out- and one in-entry.
  if (2 * (int64_t)n_out != m) return ctx->fail(XGM_EINVAL, "graph is not symmetric (out and in totals differ)");
  d->nnz = n_out;
  const size_t m1 = (size_t)std::max<int64_t>(1, d->nnz);
  HIP_TRY(ctx, hipMalloc(&d->d_colidx, m1 * sizeof(int32_t)));
  HIP_TRY(ctx, hipMalloc(&d->d_rcolidx, m1 * sizeof(int32_t)));
  HIP_TRY(ctx, launch_prep_scatter(u->d_rowptr, u->d_colidx, flag.as<int32_t>(), pos.as<int32_t>(), n, m, d->d_rowptr,
                                   d->d_colidx, d->d_rrowptr, d->d_rcolidx, cu, s));
  HIP_TRY(ctx, hipMalloc(&keys.p, 2 * n1 * sizeof(uint32_t)));
  HIP_TRY(ctx, hipMalloc(&ids.p, 2 * n1 * sizeof(int32_t)));
  HIP_TRY(ctx, hipMalloc(&dplus.p, n1 * sizeof(int32_t)));
  uint32_t *k_in = keys.as<uint32_t>(), *k_out = k_in + n1;
  int32_t *i_in = ids.as<int32_t>(), *i_out = i_in + n1;
  HIP_TRY(ctx, launch_prep_work(d->d_rowptr, d->d_colidx, n, k_in, i_in, dplus.as<int32_t>(), cu, s));
  if (n) {
    size_t b = tmp_bytes;
    HIP_TRY(ctx, prep_sort(tmp.p, b, k_in, k_out, i_in, i_out, n, s));
  }
  // only d+ and the task order come back (2 ints per vertex): the task
  // lists are built from them or in HBM, so the CSR stays on the device
  std::vector<int32_t> sorted;
  try {
    d->dplus.resize(n);
    sorted.resize(n);
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  if (n) {
    HIP_TRY(ctx, hipMemcpyAsync(d->dplus.data(), dplus.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(sorted.data(), i_out, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(ctx, hipStreamSynchronize(s));
  int64_t n_tasks = 0;
  d->max_out = 0;
  for (int64_t v = 0; v < n; ++v) {
    d->max_out = std::max(d->max_out, d->dplus[v]);
    n_tasks += d->dplus[v] >= 2;
  }
  d->order.assign(sorted.begin(), sorted.begin() + n_tasks);
  out = std::move(d);
  return XGM_OK;
}

}  // namespace

extern "C" {

int xgm_ctx_create(int device, xgm_ctx **out) {
  if (!out) return XGM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return XGM_ENODEV;
  if (device < 0 || device >= ndev) return XGM_EINVAL;
  std::unique_ptr<xgm_ctx> ctx(new (std::nothrow) xgm_ctx);
  if (!ctx) return XGM_ENOMEM;
  ctx->device = device;
  HIP_TRY(ctx.get(), hipSetDevice(device));
  int rc = check_device(ctx.get());
  if (rc) return rc;
  HIP_TRY(ctx.get(), hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIP_TRY(ctx.get(), hipMalloc(&ctx->d_out, kScratchU64 * sizeof(u64)));
  HIP_TRY(ctx.get(), hipMalloc(&ctx->d_queues, kScratchQueues * sizeof(int)));
  ctx->ev.resize(2 * XGM_MAX_PROFILE_KERNELS);
  for (auto &e : ctx->ev) HIP_TRY(ctx.get(), hipEventCreate(&e));
  HIP_TRY(ctx.get(), setop_scratch_address(device, &ctx->d_noalloc));
  *out = ctx.release();
  return XGM_OK;
}

void xgm_ctx_destroy(xgm_ctx *ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (auto e : ctx->ev) hipEventDestroy(e);
  hipFree(ctx->d_out);
  hipFree(ctx->d_queues);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *xgm_ctx_last_error(const xgm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int xgm_ctx_set_prep(xgm_ctx *ctx, int route) {
  if (!ctx) return XGM_EINVAL;
  if (route != XGM_PREP_AUTO && route != XGM_PREP_HOST && route != XGM_PREP_DEVICE)
    return ctx->fail(XGM_EINVAL, "prep route must be XGM_PREP_AUTO, _HOST or _DEVICE");
  ctx->prep = route;
  return XGM_OK;
}

int xgm_graph_upload(xgm_ctx *ctx, const xgm_csr *g, xgm_dgraph **out) {
  if (!ctx || !g || !out) return XGM_EINVAL;
  *out = nullptr;
  // the O(m log d) symmetry check runs on the device after the copy
  if (validate_csr(g, false) != XGM_OK) return ctx->fail(XGM_EINVAL, "CSR failed validation");
  if (!(g->flags & XGM_CSR_SORTED)) return ctx->fail(XGM_EINVAL, "neighbour lists must be sorted");
  if (g->nnz >= (int64_t)INT32_MAX) return ctx->fail(XGM_EUNSUPPORTED, "nnz >= 2^31 not supported");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<xgm_dgraph> d(new (std::nothrow) xgm_dgraph);
  if (!d) return XGM_ENOMEM;
  d->ctx = ctx;
  d->device = ctx->device;
  d->n = g->n;
  d->nnz = g->nnz;
  d->flags = g->flags;
  auto cleanup = [&](int rc) { return rc; };  // ~xgm_dgraph frees the buffers
  hipError_t e;
  if ((e = hipMalloc(&d->d_rowptr, (g->n + 1) * sizeof(int64_t))) != hipSuccess)
    return cleanup(ctx->hip_fail(e, "hipMalloc(rowptr)"));
  if ((e = hipMalloc(&d->d_colidx, std::max<int64_t>(1, g->nnz) * sizeof(int32_t))) != hipSuccess)
    return cleanup(ctx->hip_fail(e, "hipMalloc(colidx)"));
  if ((e = hipMemcpyAsync(d->d_rowptr, g->rowptr, (g->n + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                          ctx->stream)) != hipSuccess)
    return cleanup(ctx->hip_fail(e, "H2D rowptr"));
  if (g->nnz && (e = hipMemcpyAsync(d->d_colidx, g->colidx, g->nnz * sizeof(int32_t),
                                    hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
    return cleanup(ctx->hip_fail(e, "H2D colidx"));
  if (g->flags & XGM_CSR_SYMMETRIC) {
    DevBuf bad;
    u64 h_bad[3] = {0, 0, 0};
    HIP_TRY(ctx, hipMalloc(&bad.p, 4 * sizeof(u64)));
    HIP_TRY(ctx, launch_zero(bad.p, 4, ctx->stream));
    HIP_TRY(ctx, launch_prep_mirror(d->d_rowptr, d->d_colidx, g->n, g->nnz, bad.as<u64>(), ctx->num_cu, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(h_bad, bad.p, 3 * sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (h_bad[0] || h_bad[1] != h_bad[2]) return ctx->fail(XGM_EINVAL, "CSR failed validation (not symmetric)");
  }
  // The host copy feeds task preparation: DAG task lists need it at once;
  // an undirected graph copies it back from HBM on first use.
  if (g->flags & XGM_CSR_DAG) {
    try {
      d->h_rowptr.assign(g->rowptr, g->rowptr + g->n + 1);
      d->h_colidx.assign(g->colidx, g->colidx + g->nnz);
    } catch (const std::bad_alloc &) {
      return XGM_ENOMEM;
    }
  }
  if (g->flags & XGM_CSR_DAG) {
    build_dag_order(*g, d->order, d->dplus, d->max_out);
    try {
      build_reverse(d->h_rowptr, d->h_colidx, g->n, d->h_rrowptr, d->h_rcolidx);
    } catch (const std::bad_alloc &) {
      return XGM_ENOMEM;
    }
    if ((e = hipMalloc(&d->d_rrowptr, (g->n + 1) * sizeof(int64_t))) != hipSuccess)
      return cleanup(ctx->hip_fail(e, "hipMalloc(rrowptr)"));
    if ((e = hipMalloc(&d->d_rcolidx, std::max<int64_t>(1, g->nnz) * sizeof(int32_t))) != hipSuccess)
      return cleanup(ctx->hip_fail(e, "hipMalloc(rcolidx)"));
    if ((e = hipMemcpyAsync(d->d_rrowptr, d->h_rrowptr.data(), (g->n + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                            ctx->stream)) != hipSuccess)
      return cleanup(ctx->hip_fail(e, "H2D rrowptr"));
    if (g->nnz && (e = hipMemcpyAsync(d->d_rcolidx, d->h_rcolidx.data(), g->nnz * sizeof(int32_t),
                                      hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
      return cleanup(ctx->hip_fail(e, "H2D rcolidx"));
  }
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return cleanup(ctx->hip_fail(e, "upload sync"));
  *out = d.release();
  return XGM_OK;
}

int xgm_graph_orient(xgm_ctx *ctx, const xgm_dgraph *und, xgm_dgraph **out) {
  if (!ctx || !und || !out) return XGM_EINVAL;
  *out = nullptr;
  if (und->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
  if ((und->flags & (XGM_CSR_SORTED | XGM_CSR_SYMMETRIC)) != (XGM_CSR_SORTED | XGM_CSR_SYMMETRIC) ||
      (und->flags & XGM_CSR_DAG))
    return ctx->fail(XGM_EINVAL, "orientation needs a sorted, symmetric (undirected) graph");
  if (und->d_slot) return ctx->fail(XGM_EINVAL, "orient a batch through its counts, not xgm_graph_orient");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<xgm_dgraph> d;
  const int rc = orient_device(ctx, und, d);
  if (rc) return rc;
  *out = d.release();
  return XGM_OK;
}

int xgm_graph_download(xgm_ctx *ctx, const xgm_dgraph *g, xgm_hgraph **out) {
  if (!ctx || !g || !out) return XGM_EINVAL;
  *out = nullptr;
  if (g->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<xgm_hgraph> h(new (std::nothrow) xgm_hgraph);
  if (!h) return XGM_ENOMEM;
  try {
    h->rowptr.resize(g->n + 1);
    h->colidx.resize(g->nnz);
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  h->n = g->n;
  h->nnz = g->nnz;
  h->flags = g->flags;
  HIP_TRY(ctx, hipMemcpy(h->rowptr.data(), g->d_rowptr, (g->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (g->nnz) HIP_TRY(ctx, hipMemcpy(h->colidx.data(), g->d_colidx, g->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
  *out = h.release();
  return XGM_OK;
}

int xgm_graph_tc_items(xgm_ctx *ctx, const xgm_dgraph *g, const xgm_opts *opts, int32_t *out, int64_t cap,
                       int64_t *len, int64_t *tier_counts) {
  if (!ctx || !g || !len) return XGM_EINVAL;
  if (!(g->flags & XGM_CSR_DAG)) return ctx->fail(XGM_EINVAL, "triangle items need a DAG");
  if (opts && opts->strategy != XGM_STRAT_AUTO && opts->strategy != XGM_STRAT_HASH &&
      opts->strategy != XGM_STRAT_BSEARCH)
    return ctx->fail(XGM_EINVAL, "items of the LDS-staged triangle strategies only");
  uint64_t c = 0;
  int rc = xgm_count(ctx, g, XGM_TRIANGLE, opts, &c);  // builds and caches the items
  if (rc) return rc;
  const int ws = (opts && opts->wave_max_small > 0) ? std::min(opts->wave_max_small, kTcWaveCap[0]) : kTcWaveCap[0];
  const int wl = (opts && opts->wave_max_large > 0) ? std::max(ws, std::min(opts->wave_max_large, kTcWaveCap[1]))
                                                    : std::max(ws, kTcWaveLargeDefault);
  auto it = g->task_cache.find("tc_items_" + std::to_string(ws) + "_" + std::to_string(wl));
  if (it == g->task_cache.end()) return ctx->fail(XGM_EINVAL, "no triangle items cached");
  int64_t items = 0;
  for (size_t t = 0; t < it->second.second.size(); ++t) {
    if (tier_counts && t < (size_t)NUM_TIERS) tier_counts[t] = it->second.second[t];
    items += it->second.second[t];
  }
  *len = 4 * items;
  if (out && cap > 0 && items)
    HIP_TRY(ctx, hipMemcpy(out, it->second.first, (size_t)std::min(cap, *len) * sizeof(int32_t), hipMemcpyDeviceToHost));
  return XGM_OK;
}

// ints per task of a cached list
static int list_stride(const std::string &name) {
  for (const char *p : {"tc_items", "diamond", "motif_edges", "cycle4_range", "cycle4_hubs"})
    if (name.rfind(p, 0) == 0) return 4;
  return 1;
}

int xgm_graph_task_list_names(const xgm_dgraph *g, char *names, int64_t cap, int64_t *len) {
  if (!g || !len || cap < 0 || (cap > 0 && !names)) return XGM_EINVAL;
  std::string all;
  for (const auto &kv : g->task_cache) all += kv.first + "\n";
  *len = (int64_t)all.size();
  std::copy(all.begin(), all.begin() + std::min<int64_t>(cap, *len), names);
  return XGM_OK;
}

int xgm_graph_task_list(xgm_ctx *ctx, const xgm_dgraph *g, const char *name, int32_t *out, int64_t cap,
                        int64_t *len, int64_t *counts, int counts_cap, int *n_counts) {
  if (!ctx || !g || !name || !len || cap < 0 || (cap > 0 && !out)) return XGM_EINVAL;
  auto it = g->task_cache.find(name);
  if (it == g->task_cache.end()) return ctx->fail(XGM_EINVAL, std::string("no cached task list ") + name);
  const auto &c = it->second.second;
  int64_t tasks = 0;
  for (size_t i = 0; i < c.size(); ++i) {
    if (counts && (int)i < counts_cap) counts[i] = c[i];
    tasks += c[i];
  }
  if (n_counts) *n_counts = (int)c.size();
  *len = tasks * list_stride(name);
  if (cap > 0 && *len > 0) {
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpy(out, it->second.first, (size_t)std::min(cap, *len) * sizeof(int32_t),
                           hipMemcpyDeviceToHost));
  }
  return XGM_OK;
}

int xgm_graph_task_order(const xgm_dgraph *g, int32_t *out, int64_t cap, int64_t *len) {
  if (!g || !len || cap < 0 || (cap > 0 && !out)) return XGM_EINVAL;
  if (!(g->flags & XGM_CSR_DAG)) return XGM_EINVAL;
  *len = (int64_t)g->order.size();
  std::copy(g->order.begin(), g->order.begin() + std::min<int64_t>(cap, *len), out);
  return XGM_OK;
}

void xgm_graph_free(xgm_dgraph *g) {
  if (!g) return;
  // a plan still replays kernels on this graph's buffers: release them with
  // the last such plan (xgm_plan_free)
  if (g->plan_refs > 0) {
    g->free_pending = true;
    return;
  }
  // hipFree waits for outstanding work on the device; the context is not
  // touched, so a graph may be freed after its context.
  (void)hipSetDevice(g->device);
  delete g;
}

static DevCSR dev_view(const xgm_dgraph *g);
static int clique_k(int pattern);

// Host copy of a device graph's CSR (task preparation of the diamond,
// 4-cycle and motif paths, and of clique edge items), copied from HBM on
// first use.
static int ensure_host_mirror(xgm_ctx *ctx, xgm_dgraph *g) {
  if ((int64_t)g->h_rowptr.size() == g->n + 1) return XGM_OK;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  try {
    rp.resize(g->n + 1);
    ci.resize(g->nnz);
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  HIP_TRY(ctx, hipMemcpy(rp.data(), g->d_rowptr, (g->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (g->nnz) HIP_TRY(ctx, hipMemcpy(ci.data(), g->d_colidx, g->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
  g->h_rowptr = std::move(rp);
  g->h_colidx = std::move(ci);
  return XGM_OK;
}

// Diamond / motif edge items: every undirected edge owned by its (degree,
// id)-larger endpoint, owners in degree tiers (<= 64, <= 1024, <= 8192,
// larger), owners above 1024 split into slices of ~64K probed elements so
// hubs spread over the whole GPU.  Cached as "diamond".
// Owners of degree <= 2 hold no diamond (c <= 1) and are left out; the
// motif sums (with_deg2) also need the degree-2 owners.  Built on the host
// from its copy of the CSR, or in HBM (build_dia_items_device: the same
// items).
static int diamond_lists(xgm_ctx *ctx, xgm_dgraph *gm, std::vector<TaskList> &tl, bool with_deg2 = false) {
  std::vector<std::vector<int32_t>> parts(4);
  const char *key = with_deg2 ? "motif_edges" : "diamond";
  if (!cached(gm, key) && !host_prep(ctx, gm)) {
    int32_t *d_items = nullptr;
    int64_t counts[4] = {};
    HIP_TRY(ctx, build_dia_items_device(dev_view(gm), with_deg2 ? 2 : 3, ctx->num_cu, ctx->stream, &d_items, counts));
    gm->task_cache.emplace(key, std::make_pair(d_items, std::vector<int64_t>(counts, counts + 4)));
  }
  if (!cached(gm, key)) {
    if (int rc = ensure_host_mirror(ctx, gm)) return rc;
    if ((int64_t)gm->dplus.size() != gm->n) {  // owner order, built on first use
      const xgm_csr hv{gm->n, gm->nnz, gm->h_rowptr.data(), gm->h_colidx.data(), gm->flags};
      try {
        build_undirected_order(hv, gm->order, gm->dplus, gm->work);
      } catch (const std::bad_alloc &) {
        return XGM_ENOMEM;
      }
    }
    const auto &deg = gm->dplus;
    std::vector<int32_t> owners = gm->order;  // degree >= 3, by decreasing work
    if (with_deg2)
      for (int64_t u = 0; u < gm->n; ++u)
        if (deg[u] == 2) owners.push_back((int32_t)u);
    for (int32_t t : owners) {
      const int d = deg[t];
      const int tier = d <= 64 ? 0 : d <= 1024 ? 1 : d <= 8192 ? 2 : 3;
      auto &P = parts[tier];
      if (tier <= 1) {
        P.insert(P.end(), {t, 0, d, 0});
        continue;
      }
      const int32_t *nt = gm->h_colidx.data() + gm->h_rowptr[t];
      int64_t acc = 0;
      int k0 = 0;
      for (int k = 0; k < d; ++k) {
        const int32_t p = nt[k];
        if (deg[p] < d || (deg[p] == d && p < t)) acc += deg[p];
        if (acc >= kDiaSlice || k + 1 - k0 >= kDiaSlicePositions) {
          P.insert(P.end(), {t, k0, k + 1, 0});
          k0 = k + 1;
          acc = 0;
        }
      }
      if (k0 < d) P.insert(P.end(), {t, k0, d, 0});
    }
  }
  return cache_parts(ctx, gm, key, parts, 4, tl);
}

// Vertex-induced motif tasks (degree >= 3, by decreasing degree): LUT
// tiers by degree (<= 64, <= 256, <= 1024) and the set-difference list path
// above 1024, or every task in the list path (list_all); cached as
// "vmotif" / "vmotif_list", with the largest list-task degree in list_md.
static int vmotif_lists(xgm_ctx *ctx, xgm_dgraph *gm, bool list_all, std::vector<TaskList> &tl) {
  const std::string key = list_all ? "vmotif_list" : "vmotif";
  int rc = XGM_OK;
  std::vector<int32_t> ord;
  if (!cached(gm, key) && !host_prep(ctx, gm)) {  // the same lists, built in HBM
    int32_t *d_ids = nullptr;
    int64_t counts[4] = {}, mx = 0;
    HIP_TRY(ctx, build_deg3_order_device(dev_view(gm), list_all, ctx->num_cu, ctx->stream, &d_ids, counts, &mx));
    gm->task_cache.emplace(key, std::make_pair(d_ids, std::vector<int64_t>(counts, counts + 4)));
    gm->list_md[key] = mx;
  }
  if (!cached(gm, key)) {
    if ((rc = ensure_host_mirror(ctx, gm))) return rc;
    for (int64_t v = 0; v < gm->n; ++v)
      if (gm->h_rowptr[v + 1] - gm->h_rowptr[v] >= 3) ord.push_back((int32_t)v);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
      return gm->h_rowptr[a + 1] - gm->h_rowptr[a] > gm->h_rowptr[b + 1] - gm->h_rowptr[b];
    });
  }
  const int64_t *hrp = gm->h_rowptr.data();
  auto deg_of = [hrp](int32_t v) { return hrp[v + 1] - hrp[v]; };
  const std::vector<int64_t> bounds = list_all ? std::vector<int64_t>{-1, -1, -1, INT64_MAX}
                                               : std::vector<int64_t>{64, 256, 1024, INT64_MAX};
  if ((rc = tasks_split(ctx, gm, key, ord, deg_of, bounds, tl))) return rc;
  if (!ord.empty()) {
    // largest degree of a list task bounds its difference / intersection slots
    int64_t m = 0;
    for (int32_t v : ord)
      if (deg_of(v) > (list_all ? 2 : 1024)) m = std::max<int64_t>(m, deg_of(v));
    gm->list_md[key] = m;
  }
  return XGM_OK;
}

// Enqueue the kernels of one pattern count on ctx->stream (no host sync,
// no D2H): task lists, scratch and relabelled copies are built and cached
// on first use, so a second call enqueues kernels only (capturable).
static int dispatch_pattern(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const xgm_opts *opts, Launch &l) {
  const int strategy = opts ? opts->strategy : XGM_STRAT_AUTO;
  if (g->d_slot && l.slots != g->d_slot) return ctx->fail(XGM_EINVAL, "a batch counts with its instance slots");
  DevCSR dg;
  dg.rowptr = g->d_rowptr;
  dg.colidx = g->d_colidx;
  dg.n = g->n;
  dg.nnz = g->nnz;
  dg.rrowptr = g->d_rrowptr;
  dg.rcolidx = g->d_rcolidx;
  hipError_t e = hipSuccess;
  xgm_dgraph *gm = const_cast<xgm_dgraph *>(g);
  std::vector<TaskList> tl;
  int rc = XGM_OK;
  switch (pattern) {
    case XGM_TRIANGLE: {
      if (!(g->flags & XGM_CSR_DAG)) return ctx->fail(XGM_EINVAL, "TRIANGLE needs a DAG-oriented CSR");
      if (strategy == XGM_STRAT_BUCKET) {
        if (gm->max_out >= 65536) return ctx->fail(XGM_EUNSUPPORTED, "bucket strategy needs out-degrees < 65536");
        if (!gm->d_bcol) {
          HIP_TRY(ctx, hipMalloc(&gm->d_bcol, std::max<int64_t>(1, g->nnz) * sizeof(int32_t)));
          HIP_TRY(ctx, hipMalloc(&gm->d_boff, std::max<int64_t>(1, g->n) * 64 * sizeof(uint16_t)));
          HIP_TRY(ctx, hipMalloc(&gm->d_bmap, std::max<int64_t>(1, g->n) * sizeof(u64)));
          HIP_TRY(ctx, launch_bucket_build(gm->d_rowptr, gm->d_colidx, g->n, gm->d_bcol, gm->d_boff, gm->d_bmap,
                                           ctx->num_cu, l.stream));
        }
        if ((rc = tasks_by_ranges(ctx, gm, "tc_bucket", 2, {INT32_MAX}, tl))) return rc;
        e = launch_tc_bucket(dg, gm->d_bcol, gm->d_boff, gm->d_bmap, tl[0], l, l.d_queue);
        break;
      }
      if (strategy == XGM_STRAT_MERGE) {
        if ((rc = tasks_by_ranges(ctx, gm, "tc_merge", 2, {INT32_MAX}, tl))) return rc;
        e = launch_tc_merge(dg, tl[0], l, l.d_queue);
        break;
      }
      if (strategy != XGM_STRAT_AUTO && strategy != XGM_STRAT_HASH && strategy != XGM_STRAT_BSEARCH)
        return ctx->fail(XGM_EUNSUPPORTED, "strategy not available for TRIANGLE");
      int s = (strategy == XGM_STRAT_BSEARCH) ? 2 : 1;
      const int hs = (opts && opts->hash_shift == 1) ? 1 : 2;
      const int ws = (opts && opts->wave_max_small > 0) ? std::min(opts->wave_max_small, kTcWaveCap[0]) : kTcWaveCap[0];
      const int wl = (opts && opts->wave_max_large > 0) ? std::max(ws, std::min(opts->wave_max_large, kTcWaveCap[1]))
                                                        : std::max(ws, kTcWaveLargeDefault);
      const std::string key = "tc_items_" + std::to_string(ws) + "_" + std::to_string(wl);
      std::vector<std::vector<int32_t>> parts;
      if (!cached(gm, key)) {
        const bool have_host = (int64_t)gm->h_rrowptr.size() == g->n + 1;  // a host upload
        if (ctx->prep == XGM_PREP_HOST || (ctx->prep == XGM_PREP_AUTO && have_host)) {
          if (!have_host) {  // a DAG built in HBM: copy it back, reverse lists on the host
            if ((rc = ensure_host_mirror(ctx, gm))) return rc;
            try {
              build_reverse(gm->h_rowptr, gm->h_colidx, g->n, gm->h_rrowptr, gm->h_rcolidx);
            } catch (const std::bad_alloc &) {
              return XGM_ENOMEM;
            }
          }
          build_tc_items(gm, ws, wl, parts);
        } else {
          // DAG built in HBM (xgm_graph_orient): the same items, built there
          int32_t *d_items = nullptr;
          int64_t counts[NUM_TIERS] = {};
          HIP_TRY(ctx, build_tc_items_device(dg, ws, wl, ctx->num_cu, ctx->stream, &d_items, counts));
          gm->task_cache.emplace(key, std::make_pair(d_items, std::vector<int64_t>(counts, counts + NUM_TIERS)));
        }
      }
      if ((rc = cache_parts(ctx, gm, key, parts, 4, tl))) return rc;
      TaskList tiers[NUM_TIERS];
      for (int t = 0; t < NUM_TIERS; ++t) tiers[t] = tl[t];
      e = launch_tc(dg, tiers, s, hs, l, l.d_queue);
      break;
    }
    case XGM_CLIQUE4:
    case XGM_CLIQUE5: {
      if (!(g->flags & XGM_CSR_DAG)) return ctx->fail(XGM_EINVAL, "CLIQUE needs a DAG-oriented CSR");
      if (strategy != XGM_STRAT_AUTO && strategy != XGM_STRAT_LUT && strategy != XGM_STRAT_LIST &&
          strategy != XGM_STRAT_ELUT && strategy != XGM_STRAT_LUT_RANK)
        return ctx->fail(XGM_EUNSUPPORTED, "strategy not available for CLIQUE");
      const int k = (pattern == XGM_CLIQUE4) ? 4 : 5;
      const int lut_max = (opts && opts->lut_max_degree > 0) ? std::min(opts->lut_max_degree, 1024) : 1024;
      // AUTO: the rank-ordered LUT (the id-ordered one for the smallest
      // tasks), or the id-ordered one everywhere on a DAG that a rank-LUT
      // count found outside (degree, id) orientation
      const int strat = (strategy == XGM_STRAT_AUTO && !gm->rank_ok) ? XGM_STRAT_LUT : strategy;
      const std::string key = "csched" + std::to_string(k) + "_" + std::to_string(strat) + "_" +
                              std::to_string(lut_max);
      auto it = gm->scheds.find(key);
      if (it == gm->scheds.end()) {
        std::vector<int8_t> codes(g->n);
        heuristic_codes(gm, k, strat, lut_max, codes.data());
        std::unique_ptr<xgm_sched> sc(new (std::nothrow) xgm_sched);
        if (!sc) return XGM_ENOMEM;
        if ((rc = build_clique_sched(ctx, gm, pattern, codes.data(), sc.get()))) return rc;
        it = gm->scheds.emplace(key, std::move(sc)).first;
      }
      if ((rc = run_clique_sched(ctx, gm, it->second.get(), dg, l, nullptr))) return rc;
      break;
    }
    case XGM_STAR3_V:
    case XGM_DIAMOND_V:
    case XGM_TAILEDTRI_V: {
      if (!(g->flags & XGM_CSR_SYMMETRIC)) return ctx->fail(XGM_EINVAL, "vertex-induced motifs need an undirected CSR");
      if (strategy == XGM_STRAT_BUCKET) {
        // X-GMiner's bucket-bitmap P2 (diamonds only, as in the reference)
        if (pattern != XGM_DIAMOND_V)
          return ctx->fail(XGM_EUNSUPPORTED, "the bucket prefilter counts vertex-induced diamonds only");
        if ((rc = vmotif_lists(ctx, gm, true, tl))) return rc;  // every vertex of degree >= 3
        const int64_t md = gm->list_md["vmotif_list"];
        if (md >= 65536) return ctx->fail(XGM_EUNSUPPORTED, "bucket strategy needs degrees < 65536");
        if (!gm->d_bcol) {
          HIP_TRY(ctx, hipMalloc(&gm->d_bcol, std::max<int64_t>(1, g->nnz) * sizeof(int32_t)));
          HIP_TRY(ctx, hipMalloc(&gm->d_boff, std::max<int64_t>(1, g->n) * 64 * sizeof(uint16_t)));
          HIP_TRY(ctx, hipMalloc(&gm->d_bmap, std::max<int64_t>(1, g->n) * sizeof(u64)));
          HIP_TRY(ctx, launch_bucket_build(gm->d_rowptr, gm->d_colidx, g->n, gm->d_bcol, gm->d_boff, gm->d_bmap,
                                           ctx->num_cu, l.stream));
        }
        const size_t need = (size_t)vmotif_bucket_grid(ctx->num_cu) * 4 * (size_t)std::max<int64_t>(1, md);
        if ((rc = ensure_scratch(ctx, gm, need))) return rc;
        e = launch_vmotif_diamond_bucket(dg, gm->d_bcol, gm->d_boff, gm->d_bmap, tl[3], gm->d_scratch,
                                         (int)std::max<int64_t>(1, md), l, l.d_queue);
        break;
      }
      if (strategy != XGM_STRAT_AUTO && strategy != XGM_STRAT_LUT && strategy != XGM_STRAT_LIST)
        return ctx->fail(XGM_EUNSUPPORTED, "strategy not available for vertex-induced motifs");
      // vertex tasks of degree >= 3 by decreasing degree: LUT tiers by
      // degree (<= 64, <= 256, <= 1024) and the set-difference list path
      // above 1024 (every task under XGM_STRAT_LIST)
      const bool list_all = strategy == XGM_STRAT_LIST;
      const std::string key = list_all ? "vmotif_list" : "vmotif";
      if ((rc = vmotif_lists(ctx, gm, list_all, tl))) return rc;
      TaskList t4[4] = {tl[0], tl[1], tl[2], tl[3]};
      int md = 0;
      if (t4[3].count) {
        md = (int)gm->list_md[key];
        const size_t need = (size_t)vmotif_list_grid(ctx->num_cu) * 8 * (size_t)std::max(1, md);
        if ((rc = ensure_scratch(ctx, gm, need))) return rc;
      }
      const int which = pattern == XGM_STAR3_V ? 0 : pattern == XGM_DIAMOND_V ? 1 : 2;
      e = launch_vmotif(dg, which, t4, gm->d_scratch, std::max(1, md), l, l.d_queue);
      break;
    }
    case XGM_DIAMOND_EDGE: {
      if (!(g->flags & XGM_CSR_SYMMETRIC)) return ctx->fail(XGM_EINVAL, "DIAMOND_EDGE needs an undirected CSR");
      if (strategy != XGM_STRAT_AUTO && strategy != XGM_STRAT_HASH)
        return ctx->fail(XGM_EUNSUPPORTED, "strategy not available for DIAMOND_EDGE");
      if ((rc = diamond_lists(ctx, gm, tl))) return rc;
      TaskList t4[4] = {tl[0], tl[1], tl[2], tl[3]};
      e = launch_diamond(dg, t4, l, l.d_queue);
      break;
    }
    case XGM_CYCLE4_EDGE: {
      if (!(g->flags & XGM_CSR_SYMMETRIC)) return ctx->fail(XGM_EINVAL, "CYCLE4_EDGE needs an undirected CSR");
      if (strategy == XGM_STRAT_LIST) {
        // reference loop order: tasks v0 with >= 2 lower neighbours (original ids)
        std::vector<int32_t> ord;
        if (!cached(gm, "cycle4_pair")) {
          if ((rc = ensure_host_mirror(ctx, gm))) return rc;
          const auto &rp = gm->h_rowptr;
          const auto &ci = gm->h_colidx;
          for (int64_t v = 0; v < g->n; ++v) {
            int64_t lo = std::lower_bound(ci.data() + rp[v], ci.data() + rp[v + 1], (int32_t)v) - (ci.data() + rp[v]);
            if (lo >= 2) ord.push_back((int32_t)v);
          }
        }
        if ((rc = tasks_split(ctx, gm, "cycle4_pair", ord, [](int32_t) { return (int64_t)0; }, {INT64_MAX}, tl)))
          return rc;
        e = launch_cycle4_pair(dg, tl[0], l, l.d_queue);
        break;
      }
      if (strategy != XGM_STRAT_AUTO && strategy != XGM_STRAT_HASH)
        return ctx->fail(XGM_EUNSUPPORTED, "strategy not available for CYCLE4_EDGE");
      if (!gm->relabelled && !host_prep(ctx, gm)) {
        // the same relabelled CSR, order and wedge work, built in HBM; only
        // the order, the work and the lower-neighbour counts come back
        HIP_TRY(ctx, relabel_by_degree_device(dg, g->d_slot, ctx->num_cu, ctx->stream, &gm->d_rl_rowptr,
                                              &gm->d_rl_colidx, gm->rl_order, gm->rl_work, gm->rl_nlow));
        gm->relabelled = true;
      }
      if (!gm->relabelled) {
        if ((rc = ensure_host_mirror(ctx, gm))) return rc;
        xgm_csr hv{g->n, g->nnz, gm->h_rowptr.data(), gm->h_colidx.data(), g->flags};
        // instances keep their id ranges, so the vertex -> instance slots
        // hold for the relabelled ids too
        relabel_by_degree(hv, gm->inst_base, gm->rl_rowptr, gm->rl_colidx, gm->rl_order, gm->rl_work);
        HIP_TRY(ctx, hipMalloc(&gm->d_rl_rowptr, gm->rl_rowptr.size() * 8));
        HIP_TRY(ctx, hipMalloc(&gm->d_rl_colidx, std::max<size_t>(1, gm->rl_colidx.size()) * 4));
        HIP_TRY(ctx, hipMemcpy(gm->d_rl_rowptr, gm->rl_rowptr.data(), gm->rl_rowptr.size() * 8, hipMemcpyHostToDevice));
        if (!gm->rl_colidx.empty())
          HIP_TRY(ctx, hipMemcpy(gm->d_rl_colidx, gm->rl_colidx.data(), gm->rl_colidx.size() * 4, hipMemcpyHostToDevice));
        gm->relabelled = true;
      }
      // vertex tasks by wedge work: <=512, <=4096, <=2^18 (per-block HBM
      // counter rows); above that hubs, split into slices of ~2^16 wedges
      // that share one counter row per hub.  Hubs take the rows in rounds
      // of at most hub_rows (kHubBudgetInts counters in total), the rows
      // zeroed before every round.
      // counter rows span one instance (the largest) or the whole graph
      const int64_t stride = gm->n_inst > 1 ? gm->max_inst_n : g->n;
      const int64_t kHubWork = 1 << 18, kSlice = 1 << 16, kHubBudgetInts = (int64_t)1 << 31;
      const int64_t kRangeGroupWork = 1 << 22;
      const int64_t row_cap = std::max<int64_t>(1, kHubBudgetInts / std::max<int64_t>(1, stride));
      const int64_t rows_max = (opts && opts->hub_rows > 0) ? std::min<int64_t>(opts->hub_rows, row_cap) : row_cap;
      // tasks above 4096 wedges (above 512 when forced: modes 1, 3) may run as LDS range
      // items instead: P = ceil((x - lo) / span) ranges, each cutting the
      // nlow lower lists with two binary searches; taken when those P * nlow
      // cuts are at most as many as the wedges.  span = 2 * kRangeCounters
      // (16-bit counters) when nlow < 65536 bounds every count, else (and
      // always in mode 3) kRangeCounters.  Consecutive spans of one task
      // share an item up to about kRangeGroupWork wedges: inside an item
      // each list cut is carried into the next span, so only the item's
      // first span binary-searches
      const int rmode = opts ? opts->wedge_range : 0;
      const std::string sfx = "_" + std::to_string(rmode);
      const std::string hkey = "cycle4_hubs_" + std::to_string(rows_max) + sfx;
      const std::string wkey = "cycle4_wedge" + sfx, rkey = "cycle4_range" + sfx;
      std::vector<std::vector<int32_t>> parts(4), hubp(1), rngp(1);
      if (!cached(gm, wkey) || !cached(gm, hkey) || !cached(gm, rkey)) {
        const auto &rp = gm->rl_rowptr;
        const auto &ci = gm->rl_colidx;
        // |N(x) ∩ [0, x)| in the relabelled ids
        auto nlow_of = [&](int32_t x) -> int64_t {
          if (!gm->rl_nlow.empty()) return gm->rl_nlow[x];
          const int32_t *nx = ci.data() + rp[x];
          return std::lower_bound(nx, nx + (rp[x + 1] - rp[x]), x) - nx;
        };
        std::vector<int32_t> hubs;
        int64_t range_nlow = 0;
        for (int32_t x : gm->rl_order) {
          const int64_t w = gm->rl_work[x];
          const bool forced = rmode == 1 || rmode == 3;
          if (rmode != 2 && w > (forced ? 512 : 4096)) {
            const int64_t lo = gm->h_slot.empty() ? 0 : gm->inst_base[gm->h_slot[x]];
            const int64_t nlow = nlow_of(x);
            const bool half = rmode != 3 && nlow < 65536;
            const int64_t span = half ? 2 * (int64_t)kRangeCounters : kRangeCounters;
            const int64_t P = (x - lo + span - 1) / span;
            if (forced || P * nlow <= w) {
              const int64_t per_span = std::max<int64_t>(1, w / P);
              const int64_t group = std::max<int64_t>(1, std::min<int64_t>(P, kRangeGroupWork / per_span));
              for (int64_t r0 = lo; r0 < x; r0 += group * span)
                rngp[0].insert(rngp[0].end(),
                               {x, (int32_t)r0, (int32_t)std::min<int64_t>(x, r0 + group * span), half ? 1 : 0});
              range_nlow = std::max(range_nlow, nlow);
              continue;
            }
          }
          if (w > kHubWork) {
            hubs.push_back(x);
          } else if (w <= 128) {
            parts[nlow_of(x) <= 64 ? 3 : 0].push_back(x);
          } else {
            parts[w <= 512 ? 0 : w <= 4096 ? 1 : 2].push_back(x);
          }
        }
        const int64_t R = std::min<int64_t>((int64_t)hubs.size(), rows_max);
        if (R > 0) hubp.assign((hubs.size() + R - 1) / R, {});
        // per hub x and lower neighbour v_k: |N(v_k) ∩ [0, x)|, from the host
        // copy or, for a CSR relabelled in HBM, counted there
        std::vector<int64_t> hoff(hubs.size() + 1, 0);
        for (size_t h = 0; h < hubs.size(); ++h) hoff[h + 1] = hoff[h] + nlow_of(hubs[h]);
        std::vector<int32_t> hcnt;
        if (!hubs.empty() && rp.empty())
          HIP_TRY(ctx, cycle4_hub_counts_device(gm->d_rl_rowptr, gm->d_rl_colidx, hubs, hoff, ctx->num_cu,
                                                ctx->stream, hcnt));
        for (size_t h = 0; h < hubs.size(); ++h) {
          const int32_t x = hubs[h];
          auto &P = hubp[h / R];
          const int32_t slot = (int32_t)(h % R);
          const int nlow = (int)(hoff[h + 1] - hoff[h]);
          int64_t acc = 0;
          int k0 = 0;
          for (int k = 0; k < nlow; ++k) {
            if (rp.empty()) {
              acc += hcnt[hoff[h] + k];
            } else {
              const int32_t v = ci[rp[x] + k];
              const int32_t *nv = ci.data() + rp[v];
              acc += std::lower_bound(nv, nv + (rp[v + 1] - rp[v]), x) - nv;
            }
            if (acc >= kSlice || k + 1 - k0 >= 4096) {
              P.insert(P.end(), {x, k0, k + 1, slot});
              k0 = k + 1;
              acc = 0;
            }
          }
          if (k0 < nlow) P.insert(P.end(), {x, k0, nlow, slot});
        }
        gm->hub_count[hkey] = (int)hubs.size();
        gm->range_nlow[rkey] = range_nlow;
      }
      const int nhubs = gm->hub_count.count(hkey) ? gm->hub_count[hkey] : 0;
      std::vector<TaskList> hl, rl;
      if ((rc = cache_parts(ctx, gm, wkey, parts, 1, tl))) return rc;
      if ((rc = cache_parts(ctx, gm, hkey, hubp, 4, hl))) return rc;
      if ((rc = cache_parts(ctx, gm, rkey, rngp, 4, rl))) return rc;
      TaskList t3[4] = {tl[0], tl[1], tl[2], tl[3]};
      const int64_t hub_rows = std::min<int64_t>(nhubs, rows_max);
      const size_t hub_round_u64 = ((size_t)hub_rows * (size_t)stride + 1) / 2;
      if (nhubs) {
        const size_t need = 2 * hub_round_u64;
        if ((rc = grow_scratch(ctx, gm, (void **)&gm->d_hub, gm->hub_ints, need, sizeof(int)))) return rc;
      }
      const int64_t cursor_stride = gm->range_nlow.count(rkey) ? gm->range_nlow[rkey] : 0;
      if (rl[0].count) {
        const size_t need = (size_t)cycle4_range_grid(ctx->num_cu, rl[0].count) * (size_t)cursor_stride;
        if ((rc = grow_scratch(ctx, gm, (void **)&gm->d_cursor, gm->cursor_n, need, sizeof(int4)))) return rc;
      }
      if (t3[2].count) {
        size_t need = (size_t)cycle4_dense_grid(ctx->num_cu, t3[2].count) * (size_t)stride;
        need = (need + 1) & ~(size_t)1;  // whole 64-bit words for launch_zero
        bool grown = false;
        if ((rc = grow_scratch(ctx, gm, (void **)&gm->d_dense, gm->dense_ints, need, sizeof(int), &grown))) return rc;
        if (grown) HIP_TRY(ctx, launch_zero(gm->d_dense, need / 2, ctx->stream));
      }
      if (!gm->d_rl_cut)  // cut positions of every lower list, once per graph
        HIP_TRY(ctx, build_rl_cuts_device(gm->d_rl_rowptr, gm->d_rl_colidx, g->n, g->nnz, ctx->num_cu, ctx->stream,
                                          &gm->d_rl_cut, &gm->d_rl_nlow));
      DevCSR rg = dg;
      rg.rowptr = gm->d_rl_rowptr;
      rg.colidx = gm->d_rl_colidx;
      rg.cut = gm->d_rl_cut;
      rg.nlow = gm->d_rl_nlow;
      CounterRows rows;
      rows.stride = stride;
      rows.inst_lo = l.slots ? gm->d_inst_lo : nullptr;
      rows.cursor = gm->d_cursor;
      rows.cursor_stride = cursor_stride;
      e = launch_cycle4_wedge(rg, t3, gm->d_dense, hl.data(), (int)hl.size(), gm->d_hub, hub_round_u64, rl[0], rows,
                              l, l.d_queue);
      break;
    }
    default:
      return ctx->fail(XGM_EUNSUPPORTED, "pattern not supported");
  }
  if (e != hipSuccess) return ctx->hip_fail(e, "kernel launch");
  return XGM_OK;
}

static int count_into(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const xgm_opts *opts, uint64_t *out,
                      bool per_inst);
static int vmotif_formula(xgm_ctx *ctx, xgm_dgraph *gm, int pattern, u64 *out);

int xgm_count(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const xgm_opts *opts, uint64_t *out) {
  return count_into(ctx, g, pattern, opts, out, false);
}

int xgm_count_csr(xgm_ctx *ctx, const xgm_csr *g, int pattern, const xgm_opts *opts, uint64_t *out) {
  xgm_dgraph *d = nullptr;
  int rc = xgm_graph_upload(ctx, g, &d);
  if (rc) return rc;
  rc = xgm_count(ctx, d, pattern, opts, out);
  xgm_graph_free(d);
  return rc;
}

int xgm_copy_bandwidth(xgm_ctx *ctx, size_t bytes, int reps, float *gbps) {
  if (!ctx || !gbps || reps < 1 || bytes < 16) return XGM_EINVAL;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  HIP_TRY(ctx, hipMalloc(&a, n16 * 16));
  hipError_t e = hipMalloc(&b, n16 * 16);
  float best = 0.f;
  if (e == hipSuccess) e = launch_zero(a, n16 * 2, ctx->stream);
  for (int r = 0; r <= reps && e == hipSuccess; ++r) {  // first launch is a warm-up
    e = hipEventRecord(ctx->ev[0], ctx->stream);
    if (e == hipSuccess) e = launch_copy(b, a, n16, ctx->num_cu, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev[1]);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    if (e == hipSuccess && r > 0 && (best == 0.f || ms < best)) best = ms;
  }
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return ctx->hip_fail(e, "copy bandwidth");
  *gbps = (float)(2.0 * (double)(n16 * 16) / (best * 1e-3) / 1e9);
  return XGM_OK;
}

int xgm_last_profile(const xgm_ctx *ctx, xgm_profile *out) {
  if (!ctx || !out) return XGM_EINVAL;
  *out = ctx->prof;
  return XGM_OK;
}

// ------------------------------------------------------------ batches

int xgm_graph_upload_batch(xgm_ctx *ctx, const xgm_csr *graphs, int n_inst, xgm_dgraph **out) {
  if (!ctx || !graphs || !out || n_inst < 1) return XGM_EINVAL;
  *out = nullptr;
  const uint32_t flags = graphs[0].flags;
  int64_t n = 0, nnz = 0, max_n = 0;
  for (int i = 0; i < n_inst; ++i) {
    if (graphs[i].flags != flags) return ctx->fail(XGM_EINVAL, "instances must share their CSR flags");
    if (xgm_csr_validate(&graphs[i]) != XGM_OK)
      return ctx->fail(XGM_EINVAL, "instance " + std::to_string(i) + " failed CSR validation");
    n += graphs[i].n;
    nnz += graphs[i].nnz;
    max_n = std::max(max_n, graphs[i].n);
  }
  if (n >= (int64_t)INT32_MAX || nnz >= (int64_t)INT32_MAX)
    return ctx->fail(XGM_EUNSUPPORTED, "batch with >= 2^31 vertices or edges");
  // disjoint union: instance i's ids shift by the vertices before it, which
  // keeps every list sorted and every degree (so the orientation) unchanged
  std::vector<int64_t> rp, base(n_inst + 1, 0);
  std::vector<int32_t> ci, slot;
  try {
    rp.resize(n + 1);
    ci.resize(nnz);
    slot.resize(n);
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  int64_t vo = 0, eo = 0;
  rp[0] = 0;
  for (int i = 0; i < n_inst; ++i) {
    const xgm_csr &g = graphs[i];
    base[i] = vo;
    for (int64_t v = 0; v < g.n; ++v) {
      rp[vo + v + 1] = eo + g.rowptr[v + 1];
      slot[vo + v] = i;
    }
    const int32_t off = (int32_t)vo;
    for (int64_t k = 0; k < g.nnz; ++k) ci[eo + k] = g.colidx[k] + off;
    vo += g.n;
    eo += g.nnz;
  }
  base[n_inst] = vo;
  xgm_csr u{n, nnz, rp.data(), ci.data(), flags};
  xgm_dgraph *d = nullptr;
  int rc = xgm_graph_upload(ctx, &u, &d);
  if (rc) return rc;
  std::unique_ptr<xgm_dgraph> dg(d);
  dg->n_inst = n_inst;
  dg->h_slot = std::move(slot);
  dg->inst_base = std::move(base);
  dg->max_inst_n = max_n;
  HIP_TRY(ctx, hipMalloc(&dg->d_slot, std::max<int64_t>(1, n) * sizeof(int32_t)));
  HIP_TRY(ctx, hipMalloc(&dg->d_inst_lo, (n_inst + 1) * sizeof(int64_t)));
  HIP_TRY(ctx, hipMalloc(&dg->d_inst_out, ((n_inst + 1) & ~1) * sizeof(u64) + sizeof(u64)));
  if (n) HIP_TRY(ctx, hipMemcpy(dg->d_slot, dg->h_slot.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(dg->d_inst_lo, dg->inst_base.data(), (n_inst + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  *out = dg.release();
  return XGM_OK;
}

int xgm_graph_instances(const xgm_dgraph *g) { return g ? g->n_inst : 0; }

}  // extern "C"

// Zero ctx scratch + the per-instance outputs, dispatch, copy back.  A
// batch always counts per instance (its 4-cycle counter rows are indexed
// within an instance); per_inst = false returns the sum.
static int count_into(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const xgm_opts *opts, uint64_t *out,
                      bool per_inst) {
  if (!ctx || !g || !out) return XGM_EINVAL;
  if (g->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
  if (opts && opts->flags != 0) return ctx->fail(XGM_EINVAL, "xgm_opts.flags must be 0");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int ni = g->d_slot ? g->n_inst : 1;
  const bool vmotif = pattern == XGM_STAR3_V || pattern == XGM_DIAMOND_V || pattern == XGM_TAILEDTRI_V;
  if (vmotif && (!opts || opts->strategy == XGM_STRAT_AUTO) && (g->flags & XGM_CSR_SYMMETRIC)) {
    xgm_dgraph *gm = const_cast<xgm_dgraph *>(g);
    std::vector<TaskList> tl;
    if (int rc = vmotif_lists(ctx, gm, false, tl)) return rc;
    if (tl[3].count) {  // tasks above the LUT tiers: the census identities
      std::vector<u64> h(ni);
      if (int rc = vmotif_formula(ctx, gm, pattern, h.data())) return rc;
      if (per_inst) std::memcpy(out, h.data(), ni * sizeof(u64));
      else out[0] = std::accumulate(h.begin(), h.end(), (u64)0);
      return XGM_OK;
    }
  }
  ctx->prof = xgm_profile{};
  HIP_TRY(ctx, launch_zero(ctx->d_out, kScratchU64, ctx->stream));
  HIP_TRY(ctx, launch_zero(ctx->d_queues, kScratchQueues / 2, ctx->stream));
  Launch l = make_launch(ctx, 0);
  if (g->d_slot) {
    HIP_TRY(ctx, launch_zero(g->d_inst_out, ni, ctx->stream));
    l.d_out = g->d_inst_out;
    l.slots = g->d_slot;
  }
  int rc = dispatch_pattern(ctx, g, pattern, opts, l);
  if (rc) return rc;
  u64 err = 0;
  std::vector<u64> h(ni);
  HIP_TRY(ctx, hipMemcpyAsync(h.data(), l.d_out, ni * sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(&err, l.d_err, sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (err) {
    // the rank-ordered LUT needs a (degree, id)-oriented DAG; AUTO falls
    // back to the id-ordered LUT for this graph from now on
    const bool clique = pattern == XGM_CLIQUE4 || pattern == XGM_CLIQUE5;
    if (clique && (!opts || opts->strategy == XGM_STRAT_AUTO) && g->rank_ok) {
      const_cast<xgm_dgraph *>(g)->rank_ok = false;
      return count_into(ctx, g, pattern, opts, out, per_inst);
    }
    return ctx->fail(XGM_EINVAL, "DAG is not in (degree, id) orientation; use XGM_STRAT_LUT");
  }
  float total = 0.f;
  for (int i = 0; i < ctx->prof.n_kernels; ++i) {
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[2 * i], ctx->ev[2 * i + 1]));
    ctx->prof.k[i].ms = ms;
  }
  if (ctx->prof.n_kernels > 0)
    HIP_TRY(ctx, hipEventElapsedTime(&total, ctx->ev[0], ctx->ev[2 * ctx->prof.n_kernels - 1]));
  ctx->prof.total_ms = total;
  if (per_inst) std::memcpy(out, h.data(), ni * sizeof(u64));
  else out[0] = std::accumulate(h.begin(), h.end(), (u64)0);
  return XGM_OK;
}

struct xgm_plan {
  int device = 0;
  int n_inst = 1;
  std::vector<int> patterns;
  u64 *d_out = nullptr;  // n_pat * n_inst counts, then the error word
  int *d_queues = nullptr;
  size_t out_words = 0, queue_words = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // graphs the captured kernels read (each holds one plan reference)
  xgm_dgraph *uses[2] = {nullptr, nullptr};
  // copies of the explicit clique schedules (their task lists are read by
  // the captured kernels), one slot per pattern
  std::vector<std::unique_ptr<xgm_sched>> scheds;
  ~xgm_plan() {
    (void)hipSetDevice(device);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    scheds.clear();
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    (void)hipFree(d_out);
    (void)hipFree(d_queues);
    for (xgm_dgraph *g : uses) {
      if (!g) continue;
      if (--g->plan_refs == 0 && g->free_pending) {
        g->free_pending = false;
        xgm_graph_free(g);
      }
    }
  }
};

extern "C" {

int xgm_count_instances(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const xgm_opts *opts, uint64_t *out) {
  return count_into(ctx, g, pattern, opts, out, true);
}

static bool dag_pattern(int p) { return p == XGM_TRIANGLE || p == XGM_CLIQUE4 || p == XGM_CLIQUE5; }

int xgm_count_batch(xgm_ctx *ctx, const xgm_csr *gs, const int *patterns, int n_inst, uint64_t *out) {
  if (!ctx || !gs || !patterns || !out || n_inst < 1) return XGM_EINVAL;
  std::map<int, std::vector<int>> by_pat;
  for (int i = 0; i < n_inst; ++i) {
    if ((gs[i].flags & (XGM_CSR_SORTED | XGM_CSR_SYMMETRIC)) != (XGM_CSR_SORTED | XGM_CSR_SYMMETRIC))
      return ctx->fail(XGM_EINVAL, "xgm_count_batch takes undirected sorted CSRs");
    if (patterns[i] < XGM_TRIANGLE || patterns[i] > XGM_CYCLE4_EDGE)
      return ctx->fail(XGM_EUNSUPPORTED, "pattern not supported");
    by_pat[patterns[i]].push_back(i);
  }
  for (auto &kv : by_pat) {
    const int p = kv.first;
    const std::vector<int> &idx = kv.second;
    std::vector<xgm_csr> views(idx.size());
    std::vector<std::unique_ptr<xgm_hgraph, void (*)(xgm_hgraph *)>> owned;
    for (size_t j = 0; j < idx.size(); ++j) {
      if (!dag_pattern(p)) {
        views[j] = gs[idx[j]];
        continue;
      }
      xgm_hgraph *h = nullptr;
      int rc = xgm_orient_dag(&gs[idx[j]], &h);
      if (rc) return ctx->fail(rc, "orientation of instance " + std::to_string(idx[j]) + " failed");
      owned.emplace_back(h, xgm_hgraph_free);
      xgm_hgraph_view(h, &views[j]);
    }
    xgm_dgraph *dg = nullptr;
    int rc = xgm_graph_upload_batch(ctx, views.data(), (int)views.size(), &dg);
    if (rc) return rc;
    std::unique_ptr<xgm_dgraph, void (*)(xgm_dgraph *)> hold(dg, xgm_graph_free);
    std::vector<uint64_t> tmp(idx.size());
    if ((rc = count_into(ctx, dg, p, nullptr, tmp.data(), true))) return rc;
    for (size_t j = 0; j < idx.size(); ++j) out[idx[j]] = tmp[j];
  }
  return XGM_OK;
}

// Degree-oriented copy of an undirected device graph (instances keep their
// slots), built on first use and owned by g.
static int ensure_dag(xgm_ctx *ctx, xgm_dgraph *g) {
  if (g->dag) return XGM_OK;
  std::unique_ptr<xgm_dgraph> dd;
  const int rc = orient_device(ctx, g, dd);
  if (rc) return rc;
  if (g->d_slot) {
    dd->n_inst = g->n_inst;
    dd->h_slot = g->h_slot;
    dd->inst_base = g->inst_base;
    dd->max_inst_n = g->max_inst_n;
    const size_t ni = (size_t)g->n_inst;
    HIP_TRY(ctx, hipMalloc(&dd->d_slot, std::max<int64_t>(1, g->n) * sizeof(int32_t)));
    HIP_TRY(ctx, hipMalloc(&dd->d_inst_lo, (ni + 1) * sizeof(int64_t)));
    HIP_TRY(ctx, hipMalloc(&dd->d_inst_out, ((ni + 1) & ~(size_t)1) * sizeof(u64) + sizeof(u64)));
    if (g->n) HIP_TRY(ctx, hipMemcpy(dd->d_slot, g->d_slot, g->n * sizeof(int32_t), hipMemcpyDeviceToDevice));
    HIP_TRY(ctx, hipMemcpy(dd->d_inst_lo, g->d_inst_lo, (ni + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice));
  }
  g->dag = std::move(dd);
  return XGM_OK;
}

// Per instance: the sums S0..S3 over the edges of the motif edge kernel
// (S[4 i .. 4 i + 3]) and the 4-cliques of the degree-oriented copy.
static int motif4_sums(xgm_ctx *ctx, xgm_dgraph *gm, std::vector<u64> &S, std::vector<uint64_t> &k4) {
  const int ni = gm->d_slot ? gm->n_inst : 1;
  int rc = ensure_dag(ctx, gm);
  if (rc) return rc;
  // per-edge sums on the GPU (motif mode of the diamond kernels)
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (!gm->d_motif) HIP_TRY(ctx, hipMalloc(&gm->d_motif, 4 * (size_t)ni * sizeof(u64)));
  std::vector<TaskList> tl;
  if ((rc = diamond_lists(ctx, gm, tl, true))) return rc;
  ctx->prof = xgm_profile{};
  HIP_TRY(ctx, launch_zero(ctx->d_out, kScratchU64, ctx->stream));
  HIP_TRY(ctx, launch_zero(ctx->d_queues, kScratchQueues / 2, ctx->stream));
  HIP_TRY(ctx, launch_zero(gm->d_motif, 4 * (size_t)ni, ctx->stream));
  Launch l = make_launch(ctx, 0);
  l.d_out = gm->d_motif;
  l.slots = gm->d_slot;
  TaskList t4[4] = {tl[0], tl[1], tl[2], tl[3]};
  HIP_TRY(ctx, launch_motif_edges(dev_view(gm), t4, l, l.d_queue));
  S.assign(4 * (size_t)ni, 0);
  k4.assign(ni, 0);
  HIP_TRY(ctx, hipMemcpyAsync(S.data(), gm->d_motif, S.size() * sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return count_into(ctx, gm->dag.get(), XGM_CLIQUE4, nullptr, k4.data(), true);
}

// Vertex-induced 3-stars, diamonds or tailed triangles per instance by the
// census identities (src/motif/omp_formula.cc:40-45): induced diamonds =
// S3 / 2 - 6 K4, tailed triangles = S2 / 2 - 2 diamonds, 3-stars = S0 / 6 -
// tailed / 3.  XGM_STRAT_AUTO takes it when a task lies above the LUT
// tiers (degree > 1024), where the set-difference list path grows with the
// cube of the degree.
static int vmotif_formula(xgm_ctx *ctx, xgm_dgraph *gm, int pattern, u64 *out) {
  std::vector<u64> S;
  std::vector<uint64_t> k4;
  if (int rc = motif4_sums(ctx, gm, S, k4)) return rc;
  for (size_t i = 0; i < k4.size(); ++i) {
    const u64 *s4 = &S[4 * i];
    const u64 dia = s4[3] / 2 - 6 * k4[i];
    const u64 tail = s4[2] / 2 - 2 * dia;
    out[i] = pattern == XGM_DIAMOND_V ? dia : pattern == XGM_TAILEDTRI_V ? tail : s4[0] / 6 - tail / 3;
  }
  return XGM_OK;
}

int xgm_motifs(xgm_ctx *ctx, const xgm_dgraph *g, int k, uint64_t *out) {
  if (!ctx || !g || !out) return XGM_EINVAL;
  if (g->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
  if (!(g->flags & XGM_CSR_SYMMETRIC)) return ctx->fail(XGM_EINVAL, "motif census needs an undirected CSR");
  if (k != 3 && k != 4) return ctx->fail(XGM_EUNSUPPORTED, "motif census for k = 3 or 4");
  xgm_dgraph *gm = const_cast<xgm_dgraph *>(g);
  const int ni = g->d_slot ? g->n_inst : 1;
  int rc = ensure_dag(ctx, gm);
  if (rc) return rc;
  std::vector<uint64_t> tri(ni), k4(ni), c4(ni);
  if (k == 3) {
    // wedges = sum_v d(d-1) / 2 - 3 * triangles (src/motif/omp_formula.cc:38-39)
    std::vector<u64> dd(ni, 0);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, degree_pair_sums_device(dev_view(g), g->d_slot, ni, ctx->stream, dd.data()));
    if ((rc = count_into(ctx, gm->dag.get(), XGM_TRIANGLE, nullptr, tri.data(), true))) return rc;
    for (int i = 0; i < ni; ++i) {
      out[2 * i] = dd[i] / 2 - 3 * tri[i];
      out[2 * i + 1] = tri[i];
    }
    return XGM_OK;
  }
  std::vector<u64> S;
  if ((rc = motif4_sums(ctx, gm, S, k4))) return rc;
  if ((rc = count_into(ctx, gm, XGM_CYCLE4_EDGE, nullptr, c4.data(), true))) return rc;
  // solve for the vertex-induced counts (src/motif/omp_formula.cc:40-45); the
  // induced 4-cycles come from the edge-induced count:
  // C4_edge = I(C4) + I(diamond) + 3 K4
  for (int i = 0; i < ni; ++i) {
    const u64 *s4 = &S[4 * (size_t)i];
    const u64 K4 = k4[i];
    const u64 dia = s4[3] / 2 - 6 * K4;
    const u64 tail = s4[2] / 2 - 2 * dia;
    const u64 cyc = c4[i] - dia - 3 * K4;
    const u64 path = s4[1] - 4 * cyc;
    const u64 star = s4[0] / 6 - tail / 3;
    uint64_t *o = out + 6 * (size_t)i;
    o[0] = star;
    o[1] = path;
    o[2] = tail;
    o[3] = cyc;
    o[4] = dia;
    o[5] = K4;
  }
  return XGM_OK;
}

}  // extern "C"

// op: XGM_SETOP_{INTERSECT, DIFF, DIFF_EXCL} | XGM_SETOP_{MERGE, HYBRID}; the
// kernel operation and strategy, or false for anything else.
static bool split_setop(int op, int &kop, int &strategy) {
  kop = op & 0xff;
  const int flags = op & ~0xff;
  strategy = flags == 0 ? SETOP_STRAT_BSEARCH : flags == XGM_SETOP_MERGE ? SETOP_STRAT_MERGE
           : flags == XGM_SETOP_HYBRID ? SETOP_STRAT_HYBRID : -1;
  return strategy >= 0 && (kop == XGM_SETOP_INTERSECT || kop == XGM_SETOP_DIFF || kop == XGM_SETOP_DIFF_EXCL);
}

static int setop_pairs_impl(xgm_ctx *ctx, const xgm_dgraph *g, int op, const int32_t *a, const int32_t *b,
                            const int32_t *bounds, const int32_t *anc, int n_anc, int64_t n_pairs, uint32_t *counts,
                            int64_t *offsets, int32_t *values, int64_t values_cap) {
  if (!ctx || !g || n_pairs < 0 || (n_pairs && (!a || !b || !counts))) return XGM_EINVAL;
  if (g->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
  int kop = 0, strategy = 0;
  if (!split_setop(op, kop, strategy)) return ctx->fail(XGM_EINVAL, "unknown set operation");
  if (n_anc < 0 || n_anc > XGM_SETOP_MAX_ANCESTORS || (n_anc && !anc))
    return ctx->fail(XGM_EINVAL, "ancestors: 0 to 4 per pair");
  if (values && !offsets) return ctx->fail(XGM_EINVAL, "values need offsets");
  for (int64_t i = 0; i < n_pairs; ++i)
    if (a[i] < 0 || a[i] >= g->n || b[i] < 0 || b[i] >= g->n)
      return ctx->fail(XGM_EINVAL, "pair " + std::to_string(i) + " names a vertex outside the graph");
  if (n_pairs == 0) {
    if (offsets) offsets[0] = 0;
    return XGM_OK;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  // device copies of the pair arrays, freed on every return path
  struct Bufs {
    std::vector<void *> p;
    ~Bufs() {
      for (void *x : p) (void)hipFree(x);
    }
  } bufs;
  auto dev = [&](void **d, size_t bytes) {
    hipError_t e = hipMalloc(d, std::max<size_t>(bytes, 4));
    if (e == hipSuccess) bufs.p.push_back(*d);
    return e;
  };
  int32_t *da = nullptr, *db = nullptr, *dbound = nullptr, *dval = nullptr, *danc = nullptr;
  uint32_t *dcnt = nullptr;
  int64_t *doff = nullptr;
  const size_t n = (size_t)n_pairs;
  HIP_TRY(ctx, dev((void **)&da, n * 4));
  HIP_TRY(ctx, dev((void **)&db, n * 4));
  HIP_TRY(ctx, dev((void **)&dcnt, n * 4));
  HIP_TRY(ctx, hipMemcpyAsync(da, a, n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(db, b, n * 4, hipMemcpyHostToDevice, ctx->stream));
  if (bounds) {
    HIP_TRY(ctx, dev((void **)&dbound, n * 4));
    HIP_TRY(ctx, hipMemcpyAsync(dbound, bounds, n * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  if (n_anc) {
    HIP_TRY(ctx, dev((void **)&danc, n * n_anc * 4));
    HIP_TRY(ctx, hipMemcpyAsync(danc, anc, n * n_anc * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  ctx->prof = xgm_profile{};
  Launch l = make_launch(ctx, 0);
  const DevCSR dg = dev_view(g);
  HIP_TRY(ctx, launch_setop_pairs(dg, kop, strategy, da, db, dbound, danc, n_anc, n_pairs, dcnt, nullptr, nullptr, l));
  HIP_TRY(ctx, hipMemcpyAsync(counts, dcnt, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (values) {
    offsets[0] = 0;
    for (size_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + counts[i];
    const int64_t total = offsets[n];
    if (total > values_cap)
      return ctx->fail(XGM_EINVAL, "values_cap " + std::to_string(values_cap) + " < " + std::to_string(total));
    HIP_TRY(ctx, dev((void **)&doff, (n + 1) * 8));
    HIP_TRY(ctx, dev((void **)&dval, (size_t)total * 4));
    HIP_TRY(ctx, hipMemcpyAsync(doff, offsets, (n + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, launch_setop_pairs(dg, kop, strategy, da, db, dbound, danc, n_anc, n_pairs, dcnt, doff, dval, l));
    if (total) HIP_TRY(ctx, hipMemcpyAsync(values, dval, (size_t)total * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  for (int i = 0; i < ctx->prof.n_kernels; ++i) {
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[2 * i], ctx->ev[2 * i + 1]));
    ctx->prof.k[i].ms = ms;
    ctx->prof.total_ms += ms;
  }
  return XGM_OK;
}

extern "C" {

int xgm_setop_pairs(xgm_ctx *ctx, const xgm_dgraph *g, int op, const int32_t *a, const int32_t *b,
                    const int32_t *bounds, int64_t n_pairs, uint32_t *counts, int64_t *offsets, int32_t *values,
                    int64_t values_cap) {
  return setop_pairs_impl(ctx, g, op, a, b, bounds, nullptr, 0, n_pairs, counts, offsets, values, values_cap);
}

int xgm_setop_pairs_anc(xgm_ctx *ctx, const xgm_dgraph *g, int op, const int32_t *a, const int32_t *b,
                        const int32_t *bounds, const int32_t *ancestors, int n_anc, int64_t n_pairs, uint32_t *counts) {
  return setop_pairs_impl(ctx, g, op, a, b, bounds, ancestors, n_anc, n_pairs, counts, nullptr, nullptr, 0);
}

static_assert(XGM_NOALLOC_MAX_PAIRS == kNoAllocPairs, "allocation-free pair limit");

int xgm_setop_pairs_noalloc(xgm_ctx *ctx, const xgm_dgraph *g, int op, const int32_t *a, const int32_t *b,
                            const int32_t *bounds, int64_t n_pairs, uint32_t *counts) {
  if (!ctx || !g || n_pairs < 0 || (n_pairs && (!a || !b || !counts))) return XGM_EINVAL;
  if (g->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
  int kop = 0, strategy = 0;
  if (!split_setop(op, kop, strategy)) return ctx->fail(XGM_EINVAL, "unknown set operation");
  if (n_pairs > XGM_NOALLOC_MAX_PAIRS)
    return ctx->fail(XGM_EINVAL, std::to_string(n_pairs) + " pairs exceed the allocation-free maximum of " +
                                     std::to_string((int64_t)XGM_NOALLOC_MAX_PAIRS));
  for (int64_t i = 0; i < n_pairs; ++i)
    if (a[i] < 0 || a[i] >= g->n || b[i] < 0 || b[i] >= g->n)
      return ctx->fail(XGM_EINVAL, "pair " + std::to_string(i) + " names a vertex outside the graph");
  if (n_pairs == 0) return XGM_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  // the staging regions of this device's static array; every input region
  // is overwritten before the launch and the kernel writes every count, so
  // nothing left by an earlier call is read
  int32_t *da = ctx->d_noalloc, *db = da + kNoAllocPairs, *dbound = db + kNoAllocPairs;
  uint32_t *dcnt = reinterpret_cast<uint32_t *>(dbound + kNoAllocPairs);
  const size_t n = (size_t)n_pairs;
  HIP_TRY(ctx, hipMemcpyAsync(da, a, n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(db, b, n * 4, hipMemcpyHostToDevice, ctx->stream));
  if (bounds) HIP_TRY(ctx, hipMemcpyAsync(dbound, bounds, n * 4, hipMemcpyHostToDevice, ctx->stream));
  ctx->prof = xgm_profile{};
  Launch l = make_launch(ctx, 0);
  HIP_TRY(ctx, launch_setop_pairs(dev_view(g), kop, strategy, da, db, bounds ? dbound : nullptr, nullptr, 0, n_pairs,
                                  dcnt, nullptr, nullptr, l));
  HIP_TRY(ctx, hipMemcpyAsync(counts, dcnt, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < ctx->prof.n_kernels; ++i) {
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[2 * i], ctx->ev[2 * i + 1]));
    ctx->prof.k[i].ms = ms;
    ctx->prof.total_ms += ms;
  }
  return XGM_OK;
}

static int plan_create_impl(xgm_ctx *ctx, const xgm_dgraph *dag, const xgm_dgraph *und, const int *patterns,
                            int n_pat, const xgm_opts *opts, const xgm_sched *const *scheds, xgm_plan **out) {
  if (!ctx || !patterns || n_pat < 1 || n_pat > 16 || !out) return XGM_EINVAL;
  *out = nullptr;
  for (int i = 0; scheds && i < n_pat; ++i) {
    const xgm_sched *s = scheds[i];
    if (!s) continue;
    if (!dag || s->pattern != patterns[i] || !clique_k(patterns[i]))
      return ctx->fail(XGM_EINVAL, "plan schedule " + std::to_string(i) + " is not a clique schedule of its pattern");
    if (s->device != ctx->device) return ctx->fail(XGM_EINVAL, "plan schedule belongs to another device");
  }
  int n_inst = -1;
  for (int i = 0; i < n_pat; ++i) {
    const xgm_dgraph *g = dag_pattern(patterns[i]) ? dag : und;
    if (!g) return ctx->fail(XGM_EINVAL, "plan pattern " + std::to_string(patterns[i]) + " has no graph");
    if (g->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
    if (n_inst >= 0 && g->n_inst != n_inst) return ctx->fail(XGM_EINVAL, "batches differ in instance count");
    n_inst = g->n_inst;
    if (opts && opts[i].flags != 0) return ctx->fail(XGM_EINVAL, "xgm_opts.flags must be 0");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  // first counts build and upload every task list and scratch buffer, so
  // the captured dispatch enqueues kernels only
  std::vector<uint64_t> tmp(n_inst);
  for (int i = 0; i < n_pat; ++i) {
    const xgm_dgraph *g = dag_pattern(patterns[i]) ? dag : und;
    int rc = (scheds && scheds[i]) ? xgm_count_sched(ctx, g, scheds[i], nullptr, tmp.data())
                                   : count_into(ctx, g, patterns[i], opts ? &opts[i] : nullptr, tmp.data(), true);
    if (rc) return rc;
  }
  std::unique_ptr<xgm_plan> p(new (std::nothrow) xgm_plan);
  if (!p) return XGM_ENOMEM;
  p->device = ctx->device;
  p->scheds.resize(n_pat);
  for (int i = 0; scheds && i < n_pat; ++i) {
    const xgm_sched *s = scheds[i];
    if (!s) continue;
    std::unique_ptr<xgm_sched> c(new (std::nothrow) xgm_sched);
    if (!c) return XGM_ENOMEM;
    c->pattern = s->pattern;
    c->k = s->k;
    c->device = s->device;
    std::copy(s->off, s->off + kClqLists, c->off);
    std::copy(s->cnt, s->cnt + kClqLists, c->cnt);
    // the lists are laid out back to back; the last (hub tasks) is 1 int per item
    const size_t ints = std::max<size_t>(2, (size_t)(s->off[kClqLists - 1] + s->cnt[kClqLists - 1]));
    HIP_TRY(ctx, hipMalloc(&c->d_buf, ints * sizeof(int32_t)));
    HIP_TRY(ctx, hipMemcpy(c->d_buf, s->d_buf, ints * sizeof(int32_t), hipMemcpyDeviceToDevice));
    p->scheds[i] = std::move(c);
  }
  p->n_inst = n_inst;
  p->patterns.assign(patterns, patterns + n_pat);
  p->out_words = (((size_t)n_pat * n_inst + 1) + 1) & ~(size_t)1;
  p->queue_words = (size_t)n_pat * kScratchQueues / 2;
  HIP_TRY(ctx, hipMalloc(&p->d_out, p->out_words * sizeof(u64)));
  HIP_TRY(ctx, hipMalloc(&p->d_queues, p->queue_words * sizeof(u64)));
  HIP_TRY(ctx, hipEventCreate(&p->ev[0]));
  HIP_TRY(ctx, hipEventCreate(&p->ev[1]));
  HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = XGM_OK;
  hipError_t e = launch_zero(p->d_out, p->out_words, ctx->stream);
  if (e == hipSuccess) e = launch_zero(p->d_queues, p->queue_words, ctx->stream);
  if (e != hipSuccess) rc = ctx->hip_fail(e, "plan zero");
  for (int i = 0; i < n_pat && rc == XGM_OK; ++i) {
    const xgm_dgraph *g = dag_pattern(patterns[i]) ? dag : und;
    Launch l = make_launch(ctx, 0);
    l.begin = [](const char *, int64_t) {};
    l.end = []() {};
    l.d_out = p->d_out + (size_t)i * n_inst;
    l.d_err = p->d_out + (size_t)n_pat * n_inst;
    l.d_queue = p->d_queues + (size_t)i * kScratchQueues;
    l.slots = g->d_slot;
    if (p->scheds[i])
      rc = run_clique_sched(ctx, const_cast<xgm_dgraph *>(g), p->scheds[i].get(), dev_view(g), l, nullptr);
    else
      rc = dispatch_pattern(ctx, g, patterns[i], opts ? &opts[i] : nullptr, l);
  }
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(ctx->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return ctx->hip_fail(e, "plan capture");
  p->graph = graph;
  HIP_TRY(ctx, hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  // pin the graphs the capture reads until the plan is freed
  const xgm_dgraph *used[2] = {dag, und};
  for (int j = 0; j < 2; ++j) {
    bool needed = false;
    for (int i = 0; i < n_pat; ++i) needed |= (dag_pattern(patterns[i]) ? dag : und) == used[j];
    if (needed && used[j]) {
      p->uses[j] = const_cast<xgm_dgraph *>(used[j]);
      ++p->uses[j]->plan_refs;
    }
  }
  *out = p.release();
  return XGM_OK;
}

int xgm_plan_create(xgm_ctx *ctx, const xgm_dgraph *dag, const xgm_dgraph *und, const int *patterns, int n_pat,
                    const xgm_opts *opts, xgm_plan **out) {
  return plan_create_impl(ctx, dag, und, patterns, n_pat, opts, nullptr, out);
}

int xgm_plan_create_sched(xgm_ctx *ctx, const xgm_dgraph *dag, const xgm_dgraph *und, const int *patterns,
                          int n_pat, const xgm_opts *opts, const xgm_sched *const *scheds, xgm_plan **out) {
  return plan_create_impl(ctx, dag, und, patterns, n_pat, opts, scheds, out);
}

int xgm_plan_run(xgm_ctx *ctx, xgm_plan *plan, uint64_t *out, float *ms) {
  if (!ctx || !plan || !out) return XGM_EINVAL;
  if (plan->device != ctx->device) return ctx->fail(XGM_EINVAL, "plan belongs to another device");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t nc = plan->patterns.size() * (size_t)plan->n_inst;
  std::vector<u64> h(nc + 1);
  HIP_TRY(ctx, hipEventRecord(plan->ev[0], ctx->stream));
  HIP_TRY(ctx, hipGraphLaunch(plan->exec, ctx->stream));
  HIP_TRY(ctx, hipEventRecord(plan->ev[1], ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(h.data(), plan->d_out, (nc + 1) * sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (h[nc]) return ctx->fail(XGM_EINVAL, "DAG is not in (degree, id) orientation; use XGM_STRAT_LUT");
  std::memcpy(out, h.data(), nc * sizeof(u64));
  if (ms) HIP_TRY(ctx, hipEventElapsedTime(ms, plan->ev[0], plan->ev[1]));
  return XGM_OK;
}

void xgm_plan_free(xgm_plan *plan) { delete plan; }

int xgm_autotune(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, int reps, xgm_opts *best, float *best_ms) {
  if (!ctx || !g || !best || reps < 1) return XGM_EINVAL;
  std::vector<xgm_opts> cand;
  auto mk = [](int strat, int lut, int hs, int ws, int wl) {
    xgm_opts o{};
    o.strategy = strat;
    o.lut_max_degree = lut;
    o.hash_shift = hs;
    o.wave_max_small = ws;
    o.wave_max_large = wl;
    return o;
  };
  switch (pattern) {
    case XGM_TRIANGLE:
      for (int hs : {2, 1})
        for (auto t : {std::make_pair(64, 256), std::make_pair(32, 256), std::make_pair(64, 128),
                       std::make_pair(32, 128), std::make_pair(16, 64)})
          cand.push_back(mk(XGM_STRAT_HASH, 0, hs, t.first, t.second));
      cand.push_back(mk(XGM_STRAT_BSEARCH, 0, 0, 0, 0));
      cand.push_back(mk(XGM_STRAT_MERGE, 0, 0, 0, 0));
      if (g->max_out < 65536) cand.push_back(mk(XGM_STRAT_BUCKET, 0, 0, 0, 0));
      break;
    case XGM_CLIQUE4:
    case XGM_CLIQUE5:
      cand.push_back(mk(XGM_STRAT_AUTO, 0, 0, 0, 0));  // rank LUT (id LUT off degree orientation)
      cand.push_back(mk(XGM_STRAT_LUT_RANK, 256, 0, 0, 0));
      cand.push_back(mk(XGM_STRAT_LUT, 1024, 0, 0, 0));
      cand.push_back(mk(XGM_STRAT_LUT, 256, 0, 0, 0));
      cand.push_back(mk(XGM_STRAT_ELUT, 0, 0, 0, 0));
      break;
    case XGM_DIAMOND_EDGE:
      cand.push_back(mk(XGM_STRAT_AUTO, 0, 0, 0, 0));
      break;
    case XGM_CYCLE4_EDGE:
      // the pair strategy (reference loop order) does one bounded
      // intersection per pair of lower neighbours, never less work than the
      // wedge aggregation and O(d^2) intersections on a hub: not a candidate
      cand.push_back(mk(XGM_STRAT_AUTO, 0, 0, 0, 0));
      for (int rm : {1, 2}) {
        cand.push_back(mk(XGM_STRAT_AUTO, 0, 0, 0, 0));
        cand.back().wedge_range = rm;
      }
      break;
    default:
      return ctx->fail(XGM_EUNSUPPORTED, "pattern not supported");
  }
  const int ni = g->n_inst;
  std::vector<uint64_t> ref, got(ni);
  float bms = 0.f;
  int bi = -1;
  for (size_t c = 0; c < cand.size(); ++c) {
    int rc = count_into(ctx, g, pattern, &cand[c], got.data(), true);  // builds the candidate's task lists
    if (rc == XGM_EINVAL && cand[c].strategy == XGM_STRAT_LUT_RANK) continue;  // foreign orientation
    if (rc) return rc;
    if (ref.empty()) ref = got;
    else if (got != ref)
      return ctx->fail(XGM_EINVAL, "autotune candidates disagree (candidate " + std::to_string(c) + ")");
    float t = 0.f;
    for (int r = 0; r < reps; ++r) {
      if ((rc = count_into(ctx, g, pattern, &cand[c], got.data(), true))) return rc;
      const float ms = ctx->prof.total_ms;
      t = r == 0 ? ms : std::min(t, ms);
    }
    if (bi < 0 || t < bms) {
      bi = (int)c;
      bms = t;
    }
  }
  if (bi < 0) return ctx->fail(XGM_EINVAL, "no autotune candidate ran");
  *best = cand[bi];
  if (best_ms) *best_ms = bms;
  return XGM_OK;
}


// ------------------------------------------------------------ schedules

static DevCSR dev_view(const xgm_dgraph *g) {
  DevCSR dg;
  dg.rowptr = g->d_rowptr;
  dg.colidx = g->d_colidx;
  dg.n = g->n;
  dg.nnz = g->nnz;
  dg.rrowptr = g->d_rrowptr;
  dg.rcolidx = g->d_rcolidx;
  return dg;
}

static int clique_k(int pattern) { return pattern == XGM_CLIQUE4 ? 4 : pattern == XGM_CLIQUE5 ? 5 : 0; }

int xgm_sched_heuristic(const xgm_dgraph *g, int pattern, int strategy, int lut_max_degree, int8_t *codes) {
  if (!g || !codes || !clique_k(pattern) || !(g->flags & XGM_CSR_DAG)) return XGM_EINVAL;
  const int lut_max = lut_max_degree > 0 ? std::min(lut_max_degree, 1024) : 1024;
  heuristic_codes(g, clique_k(pattern), strategy, lut_max, codes);
  return XGM_OK;
}

int xgm_sched_features(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, float *feats) {
  if (!ctx || !g || !feats || !clique_k(pattern) || !(g->flags & XGM_CSR_DAG)) return XGM_EINVAL;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  float *dF = nullptr;
  HIP_TRY(ctx, hipMalloc(&dF, std::max<int64_t>(1, g->n) * 16 * sizeof(float)));
  hipError_t e = launch_clique_features(dev_view(g), clique_k(pattern), dF, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(feats, dF, g->n * 16 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dF);
  if (e != hipSuccess) return ctx->hip_fail(e, "features");
  return XGM_OK;
}

// List a vertex task runs in under code c (kernel and bin), or -1 for the
// edge codes, whose items are single DAG edges.  Mirrors build_clique_sched.
static int vertex_list(int d, int c) {
  if (c == 2 || c == 3) return -1;
  if (d > kBinCap[3]) return (c == 1 || c == 4) && d <= kHubLutMax ? 12 : -1;
  int bin = 0;
  while (d > kBinCap[bin]) ++bin;
  if (c == 0) return bin;
  return 4 + (c == 4 ? std::min(bin + 1, 3) : bin);
}

// The model predicts each task's share of its kernel's time (log2(1 +
// 1000 us)); the argmin minimises the sum of shares.  The lists run as
// consecutive persistent kernels, so a kernel lasts about max(sum of its
// shares, longest task), a task's wall time being its share x the kernel's
// resident blocks.  Per vertex-task list, while the longest task is the
// kernel's tail, moving it to the edge LUT (single-edge items, no tail)
// is taken when it lowers the estimate: the list's time falls to max(rest
// of the shares, next-longest task) and the edge kernel gains the task's
// edge-LUT share.
static void critical_path_guard(xgm_ctx *ctx, const xgm_dgraph *g, const std::vector<float> &Y, int8_t *codes) {
  auto share = [&](int64_t v, int c) { return (std::exp2((double)Y[(size_t)v * 5 + c]) - 1.0) / 1000.0; };
  struct ListTime {
    double sum = 0.0;
    std::vector<std::pair<double, int32_t>> walls;  // (wall, vertex)
  } lists[kClqLists];
  for (int64_t v = 0; v < g->n; ++v) {
    const int c = codes[v];
    if (c < 0 || c > 4) continue;
    const int li = vertex_list(g->dplus[v], c);
    if (li < 0) continue;
    const double sh = share(v, c);
    lists[li].sum += sh;
    lists[li].walls.push_back({sh * clique_list_capacity(ctx->num_cu, li), (int32_t)v});
  }
  for (int li = 0; li < kClqLists; ++li) {
    // hub tasks (list 12) are split into row pieces over many blocks by
    // their kernel, so they are not one block's tail
    if (li == 12) continue;
    auto &L = lists[li];
    std::sort(L.walls.begin(), L.walls.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
    for (size_t i = 0; i < L.walls.size(); ++i) {
      const double top = L.walls[i].first;
      if (top <= L.sum) break;  // no longer the tail
      const int32_t v = L.walls[i].second;
      const double next = i + 1 < L.walls.size() ? L.walls[i + 1].first : 0.0;
      const double rest = L.sum - share(v, codes[v]);
      const double gain = top - std::max(rest, next) - share(v, 2);
      if (gain <= 0.0) break;
      codes[v] = 2;
      L.sum = rest;
    }
  }
}

int xgm_sched_learned(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, xgm_mlp *m, int8_t *codes) {
  if (!ctx || !g || !m || !codes || !clique_k(pattern) || !(g->flags & XGM_CSR_DAG)) return XGM_EINVAL;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  float *dF = nullptr, *dY = nullptr;
  int8_t *dC = nullptr;
  auto done = [&](int rc) {
    (void)hipFree(dF);
    (void)hipFree(dY);
    (void)hipFree(dC);
    return rc;
  };
  const int64_t n = std::max<int64_t>(1, g->n);
  hipError_t e;
  if ((e = hipMalloc(&dF, n * 16 * sizeof(float))) != hipSuccess) return done(ctx->hip_fail(e, "hipMalloc(F)"));
  if ((e = hipMalloc(&dY, n * 5 * sizeof(float))) != hipSuccess) return done(ctx->hip_fail(e, "hipMalloc(Y)"));
  if ((e = hipMalloc(&dC, n)) != hipSuccess) return done(ctx->hip_fail(e, "hipMalloc(codes)"));
  const DevCSR dg = dev_view(g);
  if ((e = launch_clique_features(dg, clique_k(pattern), dF, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return done(ctx->hip_fail(e, "features"));
  int rc = mlp_predict_device(m, dF, g->n, dY);
  if (rc) return done(ctx->fail(rc, std::string("model: ") + xgm_mlp_last_error(m)));
  std::vector<float> Y((size_t)g->n * 5);
  if ((e = launch_sched_argmin(dg, dF, dY, dC, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(codes, dC, g->n, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(Y.data(), dY, Y.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return done(ctx->hip_fail(e, "argmin"));
  critical_path_guard(ctx, g, Y, codes);
  return done(XGM_OK);
}

int xgm_sched_create(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const int8_t *codes, xgm_sched **out) {
  if (!ctx || !g || !codes || !out || !clique_k(pattern) || !(g->flags & XGM_CSR_DAG)) return XGM_EINVAL;
  *out = nullptr;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<xgm_sched> sc(new (std::nothrow) xgm_sched);
  if (!sc) return XGM_ENOMEM;
  int rc = build_clique_sched(ctx, g, pattern, codes, sc.get());
  if (rc) return rc;
  *out = sc.release();
  return XGM_OK;
}

void xgm_sched_free(xgm_sched *s) { delete s; }

int xgm_count_sched(xgm_ctx *ctx, const xgm_dgraph *g, const xgm_sched *s, float *vertex_us, uint64_t *out) {
  if (!ctx || !g || !s || !out) return XGM_EINVAL;
  if (g->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  xgm_dgraph *gm = const_cast<xgm_dgraph *>(g);
  ctx->prof = xgm_profile{};
  HIP_TRY(ctx, launch_zero(ctx->d_out, kScratchU64, ctx->stream));
  HIP_TRY(ctx, launch_zero(ctx->d_queues, kScratchQueues / 2, ctx->stream));
  u64 *dlog = nullptr;
  const int64_t items = s->items();
  if (vertex_us) {
    HIP_TRY(ctx, hipMalloc(&dlog, std::max<int64_t>(1, items) * sizeof(u64)));
    HIP_TRY(ctx, launch_zero(dlog, std::max<int64_t>(1, items), ctx->stream));
  }
  Launch l = make_launch(ctx, 0);
  int rc = run_clique_sched(ctx, gm, s, dev_view(g), l, dlog);
  if (rc) {
    (void)hipFree(dlog);
    return rc;
  }
  u64 hv[kScratchU64];
  hipError_t e = hipMemcpyAsync(hv, ctx->d_out, sizeof(hv), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(dlog);
    return ctx->hip_fail(e, "count");
  }
  float total = 0.f;
  for (int i = 0; i < ctx->prof.n_kernels; ++i) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev[2 * i], ctx->ev[2 * i + 1]);
    ctx->prof.k[i].ms = ms;
  }
  if (ctx->prof.n_kernels > 0)
    (void)hipEventElapsedTime(&total, ctx->ev[0], ctx->ev[2 * ctx->prof.n_kernels - 1]);
  ctx->prof.total_ms = total;
  if (vertex_us) {
    std::vector<u64> h(std::max<int64_t>(1, items));
    e = hipMemcpy(h.data(), dlog, h.size() * sizeof(u64), hipMemcpyDeviceToHost);
    (void)hipFree(dlog);
    if (e != hipSuccess) return ctx->hip_fail(e, "D2H log");
    std::fill(vertex_us, vertex_us + g->n, 0.f);
    int64_t o = 0;
    for (int i = 0; i < kClqLists; ++i) {
      // 100 MHz ticks -> microseconds, divided by the kernel's resident
      // blocks: the task's share of its kernel's duration
      const float scale = 0.01f / (float)std::max(1, clique_list_capacity(ctx->num_cu, i));
      for (int64_t j = 0; j < s->cnt[i]; ++j) vertex_us[s->owner[i][j]] += (float)h[o + j] * scale;
      o += s->cnt[i];
    }
  }
  if (hv[kScratchU64 - 1])
    return ctx->fail(XGM_EINVAL, "DAG is not in (degree, id) orientation; use XGM_STRAT_LUT");
  *out = (uint64_t)hv[0];
  return XGM_OK;
}

}  // extern "C"
