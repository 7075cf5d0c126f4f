// C-ABI layer of libxgm: contexts, device graphs, task preparation and the
// pattern dispatch.  Declared in include/xgm.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
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

struct xgm_dgraph {
  xgm_ctx *ctx = nullptr;
  int64_t n = 0, nnz = 0;
  uint32_t flags = 0;
  int64_t *d_rowptr = nullptr;
  int32_t *d_colidx = nullptr;
  // DAG vertex-task tiers by out-degree (see TcTier)
  int32_t *d_tasks = nullptr;
  TaskList tiers[NUM_TIERS];
};

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

// Vertex tasks of a DAG, split into out-degree tiers, each sorted by
// decreasing work estimate sum_{v in N+(u)} d+(v) (ties by id).
void build_dag_tiers(const xgm_csr &g, std::vector<int32_t> &all, int64_t counts[NUM_TIERS]) {
  const int64_t n = g.n;
  std::vector<int64_t> work(n, 0);
  std::vector<int8_t> tier(n, -1);
  parallel_for(n, 4096, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) {
      int64_t du = g.rowptr[u + 1] - g.rowptr[u];
      if (du < 2) continue;
      int64_t w = 0;
      for (int64_t k = g.rowptr[u]; k < g.rowptr[u + 1]; ++k) {
        int32_t v = g.colidx[k];
        w += g.rowptr[v + 1] - g.rowptr[v];
      }
      work[u] = w;
      int t = 0;
      while (t < NUM_TIERS - 1 && du > kTierMaxDeg[t]) ++t;
      tier[u] = (int8_t)t;
    }
  });
  std::vector<std::vector<int32_t>> lists(NUM_TIERS);
  for (int64_t u = 0; u < n; ++u)
    if (tier[u] >= 0) lists[tier[u]].push_back((int32_t)u);
  all.clear();
  for (int t = 0; t < NUM_TIERS; ++t) {
    auto &L = lists[t];
    std::stable_sort(L.begin(), L.end(), [&](int32_t a, int32_t b) { return work[a] > work[b]; });
    counts[t] = (int64_t)L.size();
    all.insert(all.end(), L.begin(), L.end());
  }
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

int xgm_graph_upload(xgm_ctx *ctx, const xgm_csr *g, xgm_dgraph **out) {
  if (!ctx || !g || !out) return XGM_EINVAL;
  *out = nullptr;
  if (xgm_csr_validate(g) != XGM_OK) return ctx->fail(XGM_EINVAL, "CSR failed validation");
  if (!(g->flags & XGM_CSR_SORTED)) return ctx->fail(XGM_EINVAL, "neighbour lists must be sorted");
  if (g->nnz >= (int64_t)INT32_MAX) return ctx->fail(XGM_EUNSUPPORTED, "nnz >= 2^31 not supported");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<xgm_dgraph> d(new (std::nothrow) xgm_dgraph);
  if (!d) return XGM_ENOMEM;
  d->ctx = ctx;
  d->n = g->n;
  d->nnz = g->nnz;
  d->flags = g->flags;
  auto cleanup = [&](int rc) {
    hipFree(d->d_rowptr);
    hipFree(d->d_colidx);
    hipFree(d->d_tasks);
    return rc;
  };
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
  if (g->flags & XGM_CSR_DAG) {
    std::vector<int32_t> all;
    int64_t counts[NUM_TIERS];
    build_dag_tiers(*g, all, counts);
    if ((e = hipMalloc(&d->d_tasks, std::max<size_t>(1, all.size()) * sizeof(int32_t))) != hipSuccess)
      return cleanup(ctx->hip_fail(e, "hipMalloc(tasks)"));
    if (!all.empty() && (e = hipMemcpyAsync(d->d_tasks, all.data(), all.size() * sizeof(int32_t),
                                            hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
      return cleanup(ctx->hip_fail(e, "H2D tasks"));
    int64_t off = 0;
    for (int t = 0; t < NUM_TIERS; ++t) {
      d->tiers[t].tasks = d->d_tasks + off;
      d->tiers[t].count = counts[t];
      off += counts[t];
    }
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
      return cleanup(ctx->hip_fail(e, "upload sync"));
  } else {
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
      return cleanup(ctx->hip_fail(e, "upload sync"));
  }
  *out = d.release();
  return XGM_OK;
}

void xgm_graph_free(xgm_dgraph *g) {
  if (!g) return;
  hipSetDevice(g->ctx->device);
  hipStreamSynchronize(g->ctx->stream);
  hipFree(g->d_rowptr);
  hipFree(g->d_colidx);
  hipFree(g->d_tasks);
  delete g;
}

int xgm_count(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const xgm_opts *opts, uint64_t *out) {
  if (!ctx || !g || !out) return XGM_EINVAL;
  if (g->ctx != ctx) return ctx->fail(XGM_EINVAL, "graph belongs to another context");
  if (opts && opts->flags != 0) return ctx->fail(XGM_EINVAL, "xgm_opts.flags must be 0");
  const int strategy = opts ? opts->strategy : XGM_STRAT_AUTO;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ctx->prof = xgm_profile{};
  DevCSR dg;
  dg.rowptr = g->d_rowptr;
  dg.colidx = g->d_colidx;
  dg.n = g->n;
  dg.nnz = g->nnz;
  HIP_TRY(ctx, launch_zero(ctx->d_out, kScratchU64, ctx->stream));
  HIP_TRY(ctx, launch_zero(ctx->d_queues, kScratchQueues / 2, ctx->stream));
  Launch l = make_launch(ctx, 0);
  hipError_t e = hipSuccess;
  switch (pattern) {
    case XGM_TRIANGLE: {
      if (!(g->flags & XGM_CSR_DAG)) return ctx->fail(XGM_EINVAL, "TRIANGLE needs a DAG-oriented CSR");
      int s = (strategy == XGM_STRAT_BSEARCH) ? 2 : 1;
      if (strategy != XGM_STRAT_AUTO && strategy != XGM_STRAT_HASH && strategy != XGM_STRAT_BSEARCH)
        return ctx->fail(XGM_EUNSUPPORTED, "strategy not available for TRIANGLE");
      e = launch_tc(dg, g->tiers, s, l, ctx->d_queues);
      break;
    }
    default:
      return ctx->fail(XGM_EUNSUPPORTED, "pattern not supported");
  }
  if (e != hipSuccess) return ctx->hip_fail(e, "kernel launch");
  u64 h = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&h, ctx->d_out, sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  float total = 0.f;
  for (int i = 0; i < ctx->prof.n_kernels; ++i) {
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[2 * i], ctx->ev[2 * i + 1]));
    ctx->prof.k[i].ms = ms;
  }
  if (ctx->prof.n_kernels > 0)
    HIP_TRY(ctx, hipEventElapsedTime(&total, ctx->ev[0], ctx->ev[2 * ctx->prof.n_kernels - 1]));
  ctx->prof.total_ms = total;
  *out = (uint64_t)h;
  return XGM_OK;
}

int xgm_count_csr(xgm_ctx *ctx, const xgm_csr *g, int pattern, const xgm_opts *opts, uint64_t *out) {
  xgm_dgraph *d = nullptr;
  int rc = xgm_graph_upload(ctx, g, &d);
  if (rc) return rc;
  rc = xgm_count(ctx, d, pattern, opts, out);
  xgm_graph_free(d);
  return rc;
}

int xgm_last_profile(const xgm_ctx *ctx, xgm_profile *out) {
  if (!ctx || !out) return XGM_EINVAL;
  *out = ctx->prof;
  return XGM_OK;
}

}  // extern "C"
