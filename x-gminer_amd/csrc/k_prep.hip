// On-device task preparation: degree orientation of an undirected CSR into
// the out-neighbour DAG and its reverse (in-neighbour) lists, and the DAG
// vertex task order.
//
// Orientation rule (codes/GraphMiner/src/common/graph.cc:246-247): the edge
// {u, w} is kept in N+(u) when w follows u in (degree, id) order.  Because
// that order is total, N-(u) = N(u) \ N+(u) (self-loops dropped from both),
// and both halves inherit the id order of N(u): one flag per entry and an
// exclusive scan of the flags place the DAG and its reverse, already
// sorted.  The
// reference builds the reverse lists by scattering and sorting
// (graph.cc:233-279 builds only the DAG; the in-lists are this library's).
//
// Task order: DAG vertices with d+ >= 2 by decreasing work estimate
// sum_{v in N+(u)} d+(v), ties by increasing id -- a stable descending
// radix sort of (work + 1) keys over ids 0..n-1.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "device_common.h"
#include "kernels.h"

namespace xgm {
namespace {

constexpr int kPrepThreads = 256;

__global__ __launch_bounds__(kPrepThreads) void k_prep_degree(const int64_t *__restrict__ rowptr, int64_t n,
                                                              int32_t *__restrict__ deg) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride)
    deg[v] = (int32_t)(rowptr[v + 1] - rowptr[v]);
}

// w follows u in (degree, id) order
__device__ __forceinline__ bool follows(int du, int64_t u, int dw, int w) { return dw > du || (dw == du && w > u); }

// Row of entry k: the last u with rowptr[u] <= k (neighbouring lanes hold
// neighbouring k and walk the same search path).
__device__ __forceinline__ int64_t entry_row(const int64_t *__restrict__ rowptr, int64_t n, int64_t k) {
  int64_t a = 0, b = n;  // rowptr[a] <= k < rowptr[b]
  while (b - a > 1) {
    const int64_t mid = (a + b) >> 1;
    if (rowptr[mid] <= k) a = mid;
    else b = mid;
  }
  return a;
}

// flag[k] = 1 if entry k of the undirected CSR goes to N+ of its row, 0 if
// to N-; flag[nnz] = 0.  One thread per entry.  An exclusive scan of the
// flags gives every out-entry its DAG position and, as k - scan, every
// in-entry its reverse position: the halves keep list order, so no per-row
// counting or compaction is needed.
__global__ __launch_bounds__(kPrepThreads) void k_prep_flags(const int64_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ colidx,
                                                             const int32_t *__restrict__ deg, int64_t n, int64_t nnz,
                                                             int32_t *__restrict__ flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k <= nnz; k += stride) {
    if (k == nnz) {
      flag[k] = 0;
      continue;
    }
    const int64_t u = entry_row(rowptr, n, k);
    const int w = colidx[k];
    flag[k] = follows(deg[u], u, deg[w], w) ? 1 : 0;
  }
}

// ocol[pos[k]] = w for out-entries, icol[k - pos[k]] = w for in-entries.
__global__ __launch_bounds__(kPrepThreads) void k_prep_scatter(const int32_t *__restrict__ colidx,
                                                               const int32_t *__restrict__ flag,
                                                               const int32_t *__restrict__ pos, int64_t nnz,
                                                               int32_t *__restrict__ ocol, int32_t *__restrict__ icol) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += stride) {
    const int w = colidx[k];
    const int64_t p = pos[k];
    if (flag[k]) ocol[p] = w;
    else icol[k - p] = w;
  }
}

// orow[u] = pos[rowptr[u]], irow[u] = rowptr[u] - orow[u], u = 0..n.
__global__ __launch_bounds__(kPrepThreads) void k_prep_rows(const int64_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ pos, int64_t n,
                                                            int64_t *__restrict__ orow, int64_t *__restrict__ irow) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u <= n; u += stride) {
    const int64_t r = rowptr[u];
    const int64_t o = pos[r];
    orow[u] = o;
    irow[u] = r - o;
  }
}

// key[u] = (d+(u) >= 2) ? 1 + sum_{v in N+(u)} d+(v) : 0; ids[u] = u;
// dplus[u] = d+(u).  One wave per vertex.
__global__ __launch_bounds__(kPrepThreads) void k_prep_work(const int64_t *__restrict__ orow,
                                                            const int32_t *__restrict__ ocol, int64_t n,
                                                            uint32_t *__restrict__ key, int32_t *__restrict__ ids,
                                                            int32_t *__restrict__ dplus) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t u = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; u < n; u += nw) {
    const int64_t b = orow[u], e = orow[u + 1];
    u64 s = 0;
    for (int64_t k = b + lane; k < e; k += 64) {
      const int v = ocol[k];
      s += (u64)(orow[v + 1] - orow[v]);
    }
    s = wave_sum_u64(s);
    if (lane == 0) {
      key[u] = (e - b >= 2) ? (uint32_t)(s + 1) : 0u;
      ids[u] = (int32_t)u;
      dplus[u] = (int32_t)(e - b);
    }
  }
}

// Symmetry check of a sorted CSR: every entry v of N(u) with v > u must
// find u in N(v) (binary search), and the entries above and below the
// diagonal must be equally many.  Since lists are strictly increasing the
// found mirrors are distinct, so together the two conditions make every
// entry's mirror present.  bad[0] |= 1 on a missing mirror; bad[1] / bad[2]
// accumulate the upper / lower entry counts.  One thread per entry k (grid
// stride): its row is the last u with rowptr[u] <= k, found by a binary
// search that neighbouring lanes walk along the same path.
__global__ __launch_bounds__(kPrepThreads) void k_prep_mirror(const int64_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ colidx, int64_t n,
                                                              int64_t nnz, u64 *__restrict__ bad) {
  const int lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  u64 up = 0, lo = 0;
  bool miss = false;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += stride) {
    int64_t a = 0, b = n;  // rowptr[a] <= k < rowptr[b]
    while (b - a > 1) {
      const int64_t mid = (a + b) >> 1;
      if (rowptr[mid] <= k) a = mid;
      else b = mid;
    }
    const int64_t u = a;
    const int v = colidx[k];
    if (v > u) {
      ++up;
      const int64_t vb = rowptr[v];
      const int dv = (int)(rowptr[v + 1] - vb);
      miss |= !sorted_contains(colidx + vb, dv, (int)u);
    } else if (v < u) {
      ++lo;
    }
  }
  up = wave_sum_u64(up);
  lo = wave_sum_u64(lo);
  const u64 anym = __ballot(miss);
  if (lane == 0) {
    if (anym) atomicOr(&bad[0], 1ull);
    if (up) atomicAdd(&bad[1], up);
    if (lo) atomicAdd(&bad[2], lo);
  }
}

// ---- triangle task items (xgm_api.hip build_tc_items, on the device) ----
// Every DAG edge is charged to the endpoint t with the longer out-list (ties
// to the head): position k of N+(t) ++ N-(t) holds partner x, charged to t
// iff d+(x) < d+(t) (out-partner) or d+(x) <= d+(t) (in-partner), with work
// d+(x).  An item (t, k0, k1, 0) ends where its work reaches kTcSlice or
// where it has cap positions (64 / 256 in the wave tiers); items of zero
// work are dropped.  One wave per vertex resolves that greedy cut 64
// positions at a time: an inclusive scan of the works, then one ballot per
// cut.  Pass 1 counts the items of every vertex, pass 2 writes them at the
// offsets of the tier-major, vertex-ordered layout; a stable descending sort
// by work inside each tier then gives exactly the host builder's lists.
//
// Waves take the vertices in the order t = i * m mod p (p the least prime
// >= n, m = 2654435761 mod p, t >= n skipped): a bijection of [0, p) that
// spreads every residue class of t.  A plain grid stride would hand one wave
// every t of one residue mod its wave count, a power of two, and R-MAT ids
// of one low-bit pattern share their degree class: on R-MAT scale 22 one
// wave drew 65 times the mean work.

__device__ __forceinline__ u64 wave_incl_scan_u64(u64 v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u64 t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ int tc_tier(int d, int max_small, int max_large) {
  return d <= max_small ? 0 : d <= max_large ? 1 : d <= 2048 ? 2 : 3;
}

// Vertex visited at step i of the spread order (t >= n: none).
__device__ __forceinline__ int64_t spread_vertex(int64_t i, int64_t p, u64 m) {
  return (int64_t)(((u64)i * m) % (u64)p);
}

// Greedy cut of the positions [0, P) of one vertex into items, resolved 64
// positions at a time: an item ends at the first position where its work
// reaches kTcSlice or its length reaches cap.  work_at(k) is the work of
// position k < P (called by each lane for its own position); every lane
// calls emit(k0, k1, work) with the same wave-uniform arguments, once per
// item in position order, items of zero work included.
template <class WorkAt, class Emit>
__device__ __forceinline__ void wave_greedy_cut(int P, int cap, int lane, WorkAt work_at, Emit emit,
                                                u64 slice = (u64)kTcSlice) {
  u64 A = 0;   // work of the open item before this chunk
  int k0 = 0;  // its first position
  for (int c0 = 0; c0 < P; c0 += 64) {
    const int k = c0 + lane;
    const u64 w = k < P ? work_at(k) : 0;
    const u64 S = wave_incl_scan_u64(w, lane);
    const int nvalid = min(64, P - c0);
    int ls = 0;
    u64 Sbase = 0;  // inclusive prefix at lane ls - 1
    for (;;) {
      const bool cond = lane >= ls && lane < nvalid &&
                        (A + (S - Sbase) >= slice || c0 + lane + 1 - k0 >= cap);
      const u64 m = __ballot(cond);
      if (!m) {
        A += __shfl(S, nvalid - 1, 64) - Sbase;
        break;
      }
      const int l = __builtin_ctzll(m);
      const u64 Sl = __shfl(S, l, 64);
      emit(k0, c0 + l + 1, A + (Sl - Sbase));
      k0 = c0 + l + 1;
      A = 0;
      Sbase = Sl;
      ls = l + 1;
    }
  }
  if (k0 < P) emit(k0, P, A);
}

template <bool WRITE>
__global__ __launch_bounds__(kPrepThreads) void k_tc_items(const int64_t *__restrict__ rowptr,
                                                           const int32_t *__restrict__ colidx,
                                                           const int64_t *__restrict__ rrowptr,
                                                           const int32_t *__restrict__ rcolidx, int64_t n,
                                                           int max_small, int max_large, int32_t *__restrict__ cnt,
                                                           const int32_t *__restrict__ off, int4 *__restrict__ items,
                                                           uint32_t *__restrict__ work, int64_t spread_p,
                                                           u64 spread_m) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (kPrepThreads / 64);
  const int64_t i0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kPrepThreads / 64) + (threadIdx.x >> 6)));
  for (int64_t i = i0; i < spread_p; i += nw) {
    const int64_t t = spread_vertex(i, spread_p, spread_m);
    if (t >= n) continue;
    const int64_t ob = rowptr[t];
    const int d = (int)(rowptr[t + 1] - ob);
    int count = 0;
    if (d >= 1) {
      const int tier = tc_tier(d, max_small, max_large);
      const int cap = tier == 0 ? 64 : tier == 1 ? 256 : INT32_MAX;
      const int64_t ib = rrowptr[t];
      const int P = d + (int)(rrowptr[t + 1] - ib);
      int64_t o = WRITE ? off[t] : 0;
      wave_greedy_cut(
          P, cap, lane,
          [&](int k) -> u64 {
            const int x = k < d ? colidx[ob + k] : rcolidx[ib + (k - d)];
            const int dx = (int)(rowptr[x + 1] - rowptr[x]);
            return (k < d ? dx < d : dx <= d) ? (u64)dx : 0;
          },
          [&](int k0, int k1, u64 wk) {
            if (!wk) return;  // items of zero work are dropped
            if (WRITE && lane == 0) {
              items[o] = make_int4((int)t, k0, k1, 0);
              work[o] = (uint32_t)wk;
            }
            ++o;
            ++count;
          });
    }
    if (!WRITE && lane == 0) cnt[t] = count;
  }
}

// masked[t] = cnt[t] if t is in tier `tier`, else 0
__global__ __launch_bounds__(kPrepThreads) void k_tc_tier_mask(const int64_t *__restrict__ rowptr,
                                                               const int32_t *__restrict__ cnt, int64_t n,
                                                               int max_small, int max_large, int tier,
                                                               int32_t *__restrict__ masked) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= n; t += stride) {
    if (t == n) {  // the scan's extra element (its output is the tier total)
      masked[t] = 0;
      continue;
    }
    const int d = (int)(rowptr[t + 1] - rowptr[t]);
    masked[t] = (d >= 1 && tc_tier(d, max_small, max_large) == tier) ? cnt[t] : 0;
  }
}

// off[t] = base + scan[t] for the vertices of tier `tier`
__global__ __launch_bounds__(kPrepThreads) void k_tc_tier_off(const int64_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ scan, int64_t n,
                                                              int max_small, int max_large, int tier, int32_t base,
                                                              int32_t *__restrict__ off) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += stride) {
    const int d = (int)(rowptr[t + 1] - rowptr[t]);
    if (d >= 1 && tc_tier(d, max_small, max_large) == tier) off[t] = base + scan[t];
  }
}

__global__ __launch_bounds__(kPrepThreads) void k_iota(uint32_t *__restrict__ v, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) v[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kPrepThreads) void k_gather_int4(const int4 *__restrict__ src,
                                                              const uint32_t *__restrict__ idx, int64_t n,
                                                              int4 *__restrict__ dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}

// ---- task lists of an undirected graph (xgm_api.hip, on the device) ----

__device__ __forceinline__ int dia_tier(int d) { return d <= 64 ? 0 : d <= 1024 ? 1 : d <= 8192 ? 2 : 3; }

// Diamond / motif edge items (xgm_api.hip diamond_lists).  Owners are the
// vertices of degree >= min_deg (3; 2 for the motif sums); the work of t is
// the sum of deg(p) over the neighbours p before t in (degree, id) order.
// Pass 1 writes key[t] = work + 2 for degree >= 3, 1 for an owner of degree
// 2, else 0 (a stable descending sort then gives the owner order: degree >=
// 3 by decreasing work, then degree 2 by id), and cnt[t] = the items of t:
// one for owners of degree <= 1024, the greedy slices above (work
// kDiaSlice or kDiaSlicePositions positions).  Pass 2 writes the items of
// owner t from off[t] on.
template <bool WRITE>
__global__ __launch_bounds__(kPrepThreads) void k_dia_items(const int64_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ colidx, int64_t n,
                                                            int min_deg, u64 *__restrict__ key,
                                                            int32_t *__restrict__ cnt, const int32_t *__restrict__ off,
                                                            int4 *__restrict__ items, int64_t spread_p, u64 spread_m) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (kPrepThreads / 64);
  const int64_t i0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kPrepThreads / 64) + (threadIdx.x >> 6)));
  for (int64_t i = i0; i < spread_p; i += nw) {
    const int64_t t = spread_vertex(i, spread_p, spread_m);
    if (t >= n) continue;
    const int64_t b = rowptr[t];
    const int d = (int)(rowptr[t + 1] - b);
    auto work_at = [&](int k) -> u64 {
      const int p = colidx[b + k];
      const int dp = (int)(rowptr[p + 1] - rowptr[p]);
      return (dp < d || (dp == d && p < t)) ? (u64)dp : 0;
    };
    const bool sliced = d >= min_deg && dia_tier(d) >= 2;
    if (!WRITE) {
      u64 w = 0;
      int count = 0;
      if (sliced) {
        wave_greedy_cut(d, kDiaSlicePositions, lane, work_at, [&](int, int, u64 wk) {
          w += wk;
          ++count;
        }, (u64)kDiaSlice);
      } else {
        for (int k = lane; k < d; k += 64) w += work_at(k);
        w = wave_sum_u64(w);
        count = d >= min_deg ? 1 : 0;
      }
      if (lane == 0) {
        key[t] = d >= 3 ? w + 2 : (d == 2 && min_deg <= 2) ? 1 : 0;
        cnt[t] = count;
      }
    } else if (d >= min_deg) {
      int64_t o = off[t];
      if (!sliced) {
        if (lane == 0) items[o] = make_int4((int)t, 0, d, 0);
      } else {
        wave_greedy_cut(d, kDiaSlicePositions, lane, work_at, [&](int k0, int k1, u64) {
          if (lane == 0) items[o] = make_int4((int)t, k0, k1, 0);
          ++o;
        }, (u64)kDiaSlice);
      }
    }
  }
}

// masked[j] = cnt[order[j]] when order[j] is an owner of tier `tier`, else
// 0; masked[n] = 0 (the scan's extra element)
__global__ __launch_bounds__(kPrepThreads) void k_dia_tier_mask(const int64_t *__restrict__ rowptr,
                                                                const int32_t *__restrict__ order,
                                                                const int32_t *__restrict__ cnt, int64_t n,
                                                                int min_deg, int tier, int32_t *__restrict__ masked) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j <= n; j += stride) {
    if (j == n) {
      masked[j] = 0;
      continue;
    }
    const int t = order[j];
    const int d = (int)(rowptr[t + 1] - rowptr[t]);
    masked[j] = (d >= min_deg && dia_tier(d) == tier) ? cnt[t] : 0;
  }
}

// off[order[j]] = base + scan[j] for the owners of tier `tier`
__global__ __launch_bounds__(kPrepThreads) void k_dia_tier_off(const int64_t *__restrict__ rowptr,
                                                               const int32_t *__restrict__ order,
                                                               const int32_t *__restrict__ scan, int64_t n,
                                                               int min_deg, int tier, int32_t base,
                                                               int32_t *__restrict__ off) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += stride) {
    const int t = order[j];
    const int d = (int)(rowptr[t + 1] - rowptr[t]);
    if (d >= min_deg && dia_tier(d) == tier) off[t] = base + scan[j];
  }
}

// key[v] = deg(v) if deg(v) >= 3, else 0 (vertex-induced motif tasks)
__global__ __launch_bounds__(kPrepThreads) void k_deg3_key(const int64_t *__restrict__ rowptr, int64_t n,
                                                           uint32_t *__restrict__ key) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride) {
    const int64_t d = rowptr[v + 1] - rowptr[v];
    key[v] = d >= 3 ? (uint32_t)d : 0u;
  }
}

// cut[i] = number of the descending keys above hi[i] (one thread)
__global__ void k_desc_cuts(const uint32_t *__restrict__ sorted, int64_t n, int4 hi, int64_t *__restrict__ cut) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int h[4] = {hi.x, hi.y, hi.z, hi.w};
  for (int i = 0; i < 4; ++i) {
    int64_t lo = 0, up = n;  // first position with key <= h[i]
    while (lo < up) {
      const int64_t mid = (lo + up) >> 1;
      if (sorted[mid] > (uint32_t)h[i]) lo = mid + 1;
      else up = mid;
    }
    cut[i] = lo;
  }
}

// sum over the vertices of each instance of d(d-1): every thread folds a
// contiguous run of vertices and flushes at each instance change
__global__ __launch_bounds__(kPrepThreads) void k_deg_pairs(const int64_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ slot, int64_t n,
                                                            int64_t per_thread, u64 *__restrict__ out) {
  const int64_t v0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * per_thread;
  const int64_t v1 = v0 + per_thread < n ? v0 + per_thread : n;
  u64 acc = 0;
  int cur = -1;
  for (int64_t v = v0; v < v1; ++v) {
    const int s = slot ? slot[v] : 0;
    if (s != cur) {
      if (acc) atomicAdd(out + cur, acc);
      acc = 0;
      cur = s;
    }
    const u64 d = (u64)(rowptr[v + 1] - rowptr[v]);
    acc += d * (d - 1);
  }
  if (acc) atomicAdd(out + cur, acc);
}

// ---- relabel by (degree, id) inside each instance (xgm_api.hip
// relabel_by_degree, on the device) ----

// key[v] = (instance << 32) | deg(v): an ascending stable sort gives the new
// order, instances kept in their id ranges
__global__ __launch_bounds__(kPrepThreads) void k_rl_key(const int64_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ slot, int64_t n,
                                                         u64 *__restrict__ key) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride)
    key[v] = ((u64)(uint32_t)(slot ? slot[v] : 0) << 32) | (u64)(rowptr[v + 1] - rowptr[v]);
}

// rank[byrank[r]] = r, len[r] = deg(byrank[r]); len[n] = 0
__global__ __launch_bounds__(kPrepThreads) void k_rl_rank(const int64_t *__restrict__ rowptr,
                                                          const int32_t *__restrict__ byrank, int64_t n,
                                                          int32_t *__restrict__ rank, int64_t *__restrict__ len) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= n; r += stride) {
    if (r == n) {
      len[r] = 0;
      continue;
    }
    const int v = byrank[r];
    rank[v] = (int32_t)r;
    len[r] = rowptr[v + 1] - rowptr[v];
  }
}

// ekey[rl_rowptr[r] + k] = (r << 32) | rank[N(byrank[r])[k]]: sorting the
// keys sorts every new row; one wave per row
__global__ __launch_bounds__(kPrepThreads) void k_rl_edges(const int64_t *__restrict__ rowptr,
                                                           const int32_t *__restrict__ colidx,
                                                           const int32_t *__restrict__ byrank,
                                                           const int32_t *__restrict__ rank,
                                                           const int64_t *__restrict__ rl_rowptr, int64_t n,
                                                           u64 *__restrict__ ekey) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (kPrepThreads / 64);
  for (int64_t r = (int64_t)blockIdx.x * (kPrepThreads / 64) + (threadIdx.x >> 6); r < n; r += nw) {
    const int v = byrank[r];
    const int64_t b = rowptr[v], d = rowptr[v + 1] - b, o = rl_rowptr[r];
    for (int64_t k = lane; k < d; k += 64) ekey[o + k] = ((u64)r << 32) | (u64)(uint32_t)rank[colidx[b + k]];
  }
}

__global__ __launch_bounds__(kPrepThreads) void k_rl_low(const u64 *__restrict__ ekey, int64_t nnz,
                                                         int32_t *__restrict__ col) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += stride)
    col[k] = (int32_t)(uint32_t)ekey[k];
}

// Wedge work of the relabelled graph: nlow[x] = |N(x) ∩ [0, x)|, work[x] =
// sum over those lower neighbours v of |N(v) ∩ [0, x)| (0 when nlow < 2).
// One wave per x (spread order); *npos += the vertices of positive work.
__global__ __launch_bounds__(kPrepThreads) void k_rl_work(const int64_t *__restrict__ rp,
                                                          const int32_t *__restrict__ ci, int64_t n,
                                                          int32_t *__restrict__ nlow, u64 *__restrict__ work,
                                                          unsigned long long *__restrict__ npos, int64_t spread_p,
                                                          u64 spread_m) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (kPrepThreads / 64);
  const int64_t i0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kPrepThreads / 64) + (threadIdx.x >> 6)));
  unsigned long long pos = 0;
  for (int64_t i = i0; i < spread_p; i += nw) {
    const int64_t x = spread_vertex(i, spread_p, spread_m);
    if (x >= n) continue;
    const int64_t b = rp[x];
    const int lo = lower_bound_dev(ci + b, (int)(rp[x + 1] - b), (int)x);
    u64 w = 0;
    if (lo >= 2) {
      for (int k = lane; k < lo; k += 64) {
        const int v = ci[b + k];
        const int64_t bv = rp[v];
        w += (u64)lower_bound_dev(ci + bv, (int)(rp[v + 1] - bv), (int)x);
      }
      w = wave_sum_u64(w);
    }
    if (lane == 0) {
      nlow[x] = lo;
      work[x] = w;
    }
    pos += w > 0;
  }
  if (lane == 0 && pos) atomicAdd(npos, pos);
}

// cut[e] = |N(v) ∩ [0, x)| for every entry e of row x with v = ci[e] < x
// (0 for the entries above x), nlow[x] = the number of entries below x.
// One wave per row (spread order); the counting kernels read these instead
// of binary-searching every lower list at every count.
__global__ __launch_bounds__(kPrepThreads) void k_rl_cuts(const int64_t *__restrict__ rp,
                                                          const int32_t *__restrict__ ci, int64_t n,
                                                          int32_t *__restrict__ cut, int32_t *__restrict__ nlow,
                                                          int64_t spread_p, u64 spread_m) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (kPrepThreads / 64);
  const int64_t i0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kPrepThreads / 64) + (threadIdx.x >> 6)));
  for (int64_t i = i0; i < spread_p; i += nw) {
    const int64_t x = spread_vertex(i, spread_p, spread_m);
    if (x >= n) continue;
    const int64_t b = rp[x];
    const int d = (int)(rp[x + 1] - b);
    int lo = 0;
    for (int k0 = 0; k0 < d; k0 += 64) {
      const int k = k0 + lane;
      int c = 0;
      bool below = false;
      if (k < d) {
        const int v = ci[b + k];
        below = v < x;
        if (below) {
          const int64_t bv = rp[v];
          c = lower_bound_dev(ci + bv, (int)(rp[v + 1] - bv), (int)x);
        }
        cut[b + k] = c;
      }
      lo += __popcll(__ballot(below));
    }
    if (lane == 0) nlow[x] = lo;
  }
}

// cnt[hoff[h] + k] = |N(v_k) ∩ [0, x)| for the lower neighbours v_k of hub
// x = hubs[h] (k < hoff[h + 1] - hoff[h]); one block per hub
__global__ __launch_bounds__(kPrepThreads) void k_rl_hub_counts(const int64_t *__restrict__ rp,
                                                                const int32_t *__restrict__ ci,
                                                                const int32_t *__restrict__ hubs,
                                                                const int64_t *__restrict__ hoff, int nh,
                                                                int32_t *__restrict__ cnt) {
  for (int h = blockIdx.x; h < nh; h += gridDim.x) {
    const int x = hubs[h];
    const int64_t b = rp[x], o = hoff[h], L = hoff[h + 1] - o;
    for (int64_t k = threadIdx.x; k < L; k += blockDim.x) {
      const int v = ci[b + k];
      const int64_t bv = rp[v];
      cnt[o + k] = lower_bound_dev(ci + bv, (int)(rp[v + 1] - bv), x);
    }
  }
}

// least prime >= max(n, 2) (trial division; n < 2^31)
int64_t spread_prime(int64_t n) {
  for (int64_t p = std::max<int64_t>(n, 2);; ++p) {
    bool prime = true;
    for (int64_t q = 2; q * q <= p; ++q)
      if (p % q == 0) {
        prime = false;
        break;
      }
    if (prime) return p;
  }
}

int prep_grid(int64_t items, int per_block, int num_cu) {
  const int64_t want = (items + per_block - 1) / per_block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)num_cu * 16));
}

}  // namespace

hipError_t launch_prep_degree(const int64_t *rowptr, int64_t n, int32_t *deg, int num_cu, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep_degree, dim3(prep_grid(n, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s, rowptr, n,
                     deg);
  return hipGetLastError();
}

hipError_t launch_prep_split(const int64_t *rowptr, const int32_t *colidx, const int32_t *deg, int64_t n, int64_t nnz,
                             int32_t *flag, int num_cu, hipStream_t s) {
  hipLaunchKernelGGL(k_prep_flags, dim3(prep_grid(nnz + 1, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s, rowptr,
                     colidx, deg, n, nnz, flag);
  return hipGetLastError();
}

hipError_t launch_prep_scatter(const int64_t *rowptr, const int32_t *colidx, const int32_t *flag, const int32_t *pos,
                               int64_t n, int64_t nnz, int64_t *orow, int32_t *ocol, int64_t *irow, int32_t *icol,
                               int num_cu, hipStream_t s) {
  if (nnz)
    hipLaunchKernelGGL(k_prep_scatter, dim3(prep_grid(nnz, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s, colidx,
                       flag, pos, nnz, ocol, icol);
  hipLaunchKernelGGL(k_prep_rows, dim3(prep_grid(n + 1, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s, rowptr, pos,
                     n, orow, irow);
  return hipGetLastError();
}

hipError_t launch_prep_work(const int64_t *orow, const int32_t *ocol, int64_t n, uint32_t *key, int32_t *ids,
                            int32_t *dplus, int num_cu, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep_work, dim3(prep_grid(n, kPrepThreads / 64, num_cu)), dim3(kPrepThreads), 0, s, orow,
                     ocol, n, key, ids, dplus);
  return hipGetLastError();
}

hipError_t launch_prep_mirror(const int64_t *rowptr, const int32_t *colidx, int64_t n, int64_t nnz, u64 *bad,
                              int num_cu, hipStream_t s) {
  if (n == 0 || nnz == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep_mirror, dim3(prep_grid(nnz, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s, rowptr,
                     colidx, n, nnz, bad);
  return hipGetLastError();
}

// Device buffers released at scope exit.
struct PrepBufs {
  std::vector<void *> p;
  ~PrepBufs() {
    for (void *x : p) (void)hipFree(x);
  }
  template <class T>
  hipError_t get(T **out, size_t count) {
    void *x = nullptr;
    hipError_t e = hipMalloc(&x, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) {
      p.push_back(x);
      *out = static_cast<T *>(x);
    }
    return e;
  }
};

hipError_t build_tc_items_device(const DevCSR &g, int max_small, int max_large, int num_cu, hipStream_t s,
                                 int32_t **out, int64_t counts[NUM_TIERS]) {
  *out = nullptr;
  const int64_t n = g.n;
  PrepBufs B;
  int32_t *cnt = nullptr, *masked = nullptr, *scan = nullptr, *off = nullptr;
  hipError_t e;
  if ((e = B.get(&cnt, n)) || (e = B.get(&masked, n + 1)) || (e = B.get(&scan, n + 1)) || (e = B.get(&off, n)))
    return e;
  const int wgrid = prep_grid(n, kPrepThreads / 64, num_cu), tgrid = prep_grid(n, kPrepThreads, num_cu);
  const int64_t sp = spread_prime(n);
  const u64 sm = 2654435761ull % (u64)sp;
  if (n)
    hipLaunchKernelGGL(k_tc_items<false>, dim3(wgrid), dim3(kPrepThreads), 0, s, g.rowptr, g.colidx, g.rrowptr,
                       g.rcolidx, n, max_small, max_large, cnt, (const int32_t *)nullptr, (int4 *)nullptr,
                       (uint32_t *)nullptr, sp, sm);
  if ((e = hipGetLastError())) return e;
  // tier-major offsets: one masked exclusive scan per tier
  size_t scan_bytes = 0;
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, masked, scan, (int)(n + 1), s))) return e;
  void *scan_tmp = nullptr;
  if ((e = B.get((char **)&scan_tmp, scan_bytes))) return e;
  int64_t total = 0;
  for (int tier = 0; tier < NUM_TIERS; ++tier) {
    hipLaunchKernelGGL(k_tc_tier_mask, dim3(prep_grid(n + 1, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s,
                       g.rowptr, (const int32_t *)cnt, n, max_small, max_large, tier, masked);
    size_t b = scan_bytes;
    if ((e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, b, masked, scan, (int)(n + 1), s))) return e;
    int32_t tier_items = 0;
    if ((e = hipMemcpyAsync(&tier_items, scan + n, sizeof(int32_t), hipMemcpyDeviceToHost, s))) return e;
    if ((e = hipStreamSynchronize(s))) return e;
    if (n)
      hipLaunchKernelGGL(k_tc_tier_off, dim3(tgrid), dim3(kPrepThreads), 0, s, g.rowptr, (const int32_t *)scan, n,
                         max_small, max_large, tier, (int32_t)total, off);
    counts[tier] = tier_items;
    total += tier_items;
  }
  if ((e = hipGetLastError())) return e;
  int4 *items = nullptr, *sorted = nullptr;
  uint32_t *work = nullptr, *work_out = nullptr, *idx = nullptr, *idx_out = nullptr;
  if ((e = B.get(&items, total)) || (e = B.get(&work, total)) || (e = B.get(&work_out, total)) ||
      (e = B.get(&idx, total)) || (e = B.get(&idx_out, total)))
    return e;
  if (n)
    hipLaunchKernelGGL(k_tc_items<true>, dim3(wgrid), dim3(kPrepThreads), 0, s, g.rowptr, g.colidx, g.rrowptr,
                       g.rcolidx, n, max_small, max_large, cnt, (const int32_t *)off, items, work, sp, sm);
  if (total)
    hipLaunchKernelGGL(k_iota, dim3(prep_grid(total, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s, idx, total);
  if ((e = hipGetLastError())) return e;
  // stable descending sort by work inside each tier (the host's stable_sort)
  size_t sort_bytes = 0;
  if ((e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sort_bytes, work, work_out, idx, idx_out,
                                                          (int)std::max<int64_t>(1, total), 0, 32, s)))
    return e;
  void *sort_tmp = nullptr;
  if ((e = B.get((char **)&sort_tmp, sort_bytes))) return e;
  int64_t base = 0;
  for (int tier = 0; tier < NUM_TIERS; ++tier) {
    if (counts[tier] > 1) {
      size_t b = sort_bytes;
      if ((e = hipcub::DeviceRadixSort::SortPairsDescending(sort_tmp, b, work + base, work_out + base, idx + base,
                                                              idx_out + base, (int)counts[tier], 0, 32, s)))
        return e;
    } else if (counts[tier] == 1) {
      if ((e = hipMemcpyAsync(idx_out + base, idx + base, sizeof(uint32_t), hipMemcpyDeviceToDevice, s))) return e;
    }
    base += counts[tier];
  }
  void *res = nullptr;
  if ((e = hipMalloc(&res, std::max<int64_t>(1, total) * sizeof(int4)))) return e;
  sorted = static_cast<int4 *>(res);
  if (total)
    hipLaunchKernelGGL(k_gather_int4, dim3(prep_grid(total, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s,
                       (const int4 *)items, (const uint32_t *)idx_out, total, sorted);
  if ((e = hipGetLastError()) || (e = hipStreamSynchronize(s))) {
    (void)hipFree(res);
    return e;
  }
  *out = reinterpret_cast<int32_t *>(sorted);
  return hipSuccess;
}

// Sort temporary storage of the builders below, grown on demand.
struct PrepTemp {
  PrepBufs &B;
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t need(size_t b) {
    if (b <= bytes) return hipSuccess;
    char *x = nullptr;
    hipError_t e = B.get(&x, b);
    if (e == hipSuccess) {
      p = x;
      bytes = b;
    }
    return e;
  }
};

hipError_t build_dia_items_device(const DevCSR &g, int min_deg, int num_cu, hipStream_t s, int32_t **out,
                                  int64_t counts[4]) {
  *out = nullptr;
  const int64_t n = g.n;
  PrepBufs B;
  PrepTemp T{B};
  u64 *key = nullptr, *key_out = nullptr;
  int32_t *ids = nullptr, *order = nullptr, *cnt = nullptr, *masked = nullptr, *scan = nullptr, *off = nullptr;
  hipError_t e;
  if ((e = B.get(&key, n)) || (e = B.get(&key_out, n)) || (e = B.get(&ids, n)) || (e = B.get(&order, n)) ||
      (e = B.get(&cnt, n)) || (e = B.get(&masked, n + 1)) || (e = B.get(&scan, n + 1)) || (e = B.get(&off, n)))
    return e;
  const int64_t sp = spread_prime(n);
  const u64 sm = 2654435761ull % (u64)sp;
  const int wgrid = prep_grid(n, kPrepThreads / 64, num_cu), tgrid = prep_grid(n + 1, kPrepThreads, num_cu);
  if (n) {
    hipLaunchKernelGGL(k_dia_items<false>, dim3(wgrid), dim3(kPrepThreads), 0, s, g.rowptr, g.colidx, n, min_deg,
                       key, cnt, (const int32_t *)nullptr, (int4 *)nullptr, sp, sm);
    hipLaunchKernelGGL(k_iota, dim3(prep_grid(n, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s, (uint32_t *)ids,
                       n);
    size_t b = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, key, key_out, ids, order, (int)n, 0, 64, s)) ||
        (e = T.need(b)))
      return e;
    if ((e = hipcub::DeviceRadixSort::SortPairsDescending(T.p, b, key, key_out, ids, order, (int)n, 0, 64, s)))
      return e;
  }
  if ((e = hipGetLastError())) return e;
  // tier-major offsets in owner order: one masked exclusive scan per tier
  size_t b = 0;
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, b, masked, scan, (int)(n + 1), s)) || (e = T.need(b))) return e;
  int64_t total = 0;
  for (int tier = 0; tier < 4; ++tier) {
    hipLaunchKernelGGL(k_dia_tier_mask, dim3(tgrid), dim3(kPrepThreads), 0, s, g.rowptr, (const int32_t *)order,
                       (const int32_t *)cnt, n, min_deg, tier, masked);
    size_t bb = T.bytes;
    if ((e = hipcub::DeviceScan::ExclusiveSum(T.p, bb, masked, scan, (int)(n + 1), s))) return e;
    int32_t tier_items = 0;
    if ((e = hipMemcpyAsync(&tier_items, scan + n, sizeof(int32_t), hipMemcpyDeviceToHost, s)) ||
        (e = hipStreamSynchronize(s)))
      return e;
    if (n)
      hipLaunchKernelGGL(k_dia_tier_off, dim3(tgrid), dim3(kPrepThreads), 0, s, g.rowptr, (const int32_t *)order,
                         (const int32_t *)scan, n, min_deg, tier, (int32_t)total, off);
    counts[tier] = tier_items;
    total += tier_items;
  }
  if ((e = hipGetLastError())) return e;
  void *res = nullptr;
  if ((e = hipMalloc(&res, std::max<int64_t>(1, total) * sizeof(int4)))) return e;
  if (n && total)
    hipLaunchKernelGGL(k_dia_items<true>, dim3(wgrid), dim3(kPrepThreads), 0, s, g.rowptr, g.colidx, n, min_deg,
                       (u64 *)nullptr, (int32_t *)nullptr, (const int32_t *)off, static_cast<int4 *>(res), sp, sm);
  if ((e = hipGetLastError()) || (e = hipStreamSynchronize(s))) {
    (void)hipFree(res);
    return e;
  }
  *out = static_cast<int32_t *>(res);
  return hipSuccess;
}

hipError_t build_deg3_order_device(const DevCSR &g, bool list_all, int num_cu, hipStream_t s, int32_t **out,
                                   int64_t counts[4], int64_t *max_deg) {
  *out = nullptr;
  *max_deg = 0;
  const int64_t n = g.n;
  PrepBufs B;
  PrepTemp T{B};
  uint32_t *key = nullptr, *key_out = nullptr;
  int32_t *ids = nullptr, *sorted = nullptr;
  int64_t *cut = nullptr;
  hipError_t e;
  if ((e = B.get(&key, n)) || (e = B.get(&key_out, n)) || (e = B.get(&ids, n)) || (e = B.get(&sorted, n)) ||
      (e = B.get(&cut, 4)))
    return e;
  int64_t c[4] = {0, 0, 0, 0};  // keys above 1024, 256, 64, 2
  uint32_t top = 0;
  if (n) {
    const int grid = prep_grid(n, kPrepThreads, num_cu);
    hipLaunchKernelGGL(k_deg3_key, dim3(grid), dim3(kPrepThreads), 0, s, g.rowptr, n, key);
    hipLaunchKernelGGL(k_iota, dim3(grid), dim3(kPrepThreads), 0, s, (uint32_t *)ids, n);
    size_t b = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, key, key_out, ids, sorted, (int)n, 0, 32, s)) ||
        (e = T.need(b)))
      return e;
    if ((e = hipcub::DeviceRadixSort::SortPairsDescending(T.p, b, key, key_out, ids, sorted, (int)n, 0, 32, s)))
      return e;
    hipLaunchKernelGGL(k_desc_cuts, dim3(1), dim3(64), 0, s, (const uint32_t *)key_out, n, make_int4(1024, 256, 64, 2),
                       cut);
    if ((e = hipGetLastError()) || (e = hipMemcpyAsync(c, cut, sizeof(c), hipMemcpyDeviceToHost, s)) ||
        (e = hipMemcpyAsync(&top, key_out, sizeof(top), hipMemcpyDeviceToHost, s)) || (e = hipStreamSynchronize(s)))
      return e;
  }
  // sorted = [deg > 1024 | > 256 | > 64 | >= 3 | the rest]; lists <= 64 first
  int64_t seg[4][2];
  if (list_all) {
    seg[0][0] = seg[0][1] = seg[1][0] = seg[1][1] = seg[2][0] = seg[2][1] = 0;
    seg[3][0] = 0;
    seg[3][1] = c[3];
  } else {
    seg[0][0] = c[2], seg[0][1] = c[3];
    seg[1][0] = c[1], seg[1][1] = c[2];
    seg[2][0] = c[0], seg[2][1] = c[1];
    seg[3][0] = 0, seg[3][1] = c[0];
  }
  void *res = nullptr;
  if ((e = hipMalloc(&res, std::max<int64_t>(1, c[3]) * sizeof(int32_t)))) return e;
  int64_t o = 0;
  for (int t = 0; t < 4; ++t) {
    counts[t] = seg[t][1] - seg[t][0];
    if (counts[t] &&
        (e = hipMemcpyAsync(static_cast<int32_t *>(res) + o, sorted + seg[t][0], counts[t] * sizeof(int32_t),
                            hipMemcpyDeviceToDevice, s))) {
      (void)hipFree(res);
      return e;
    }
    o += counts[t];
  }
  if ((e = hipStreamSynchronize(s))) {
    (void)hipFree(res);
    return e;
  }
  *max_deg = counts[3] ? (int64_t)top : 0;
  *out = static_cast<int32_t *>(res);
  return hipSuccess;
}

hipError_t degree_pair_sums_device(const DevCSR &g, const int32_t *slot, int n_inst, hipStream_t s, u64 *host_out) {
  PrepBufs B;
  u64 *acc = nullptr;
  hipError_t e;
  if ((e = B.get(&acc, std::max(1, n_inst))) || (e = launch_zero(acc, std::max(1, n_inst), s))) return e;
  const int64_t per = 1024, threads = (g.n + per - 1) / per;
  if (threads)
    hipLaunchKernelGGL(k_deg_pairs, dim3((unsigned)((threads + kPrepThreads - 1) / kPrepThreads)), dim3(kPrepThreads),
                       0, s, g.rowptr, slot, g.n, per, acc);
  if ((e = hipGetLastError()) ||
      (e = hipMemcpyAsync(host_out, acc, std::max(1, n_inst) * sizeof(u64), hipMemcpyDeviceToHost, s)))
    return e;
  return hipStreamSynchronize(s);
}

hipError_t relabel_by_degree_device(const DevCSR &g, const int32_t *slot, int num_cu, hipStream_t s,
                                    int64_t **rl_rowptr, int32_t **rl_colidx, std::vector<int32_t> &order,
                                    std::vector<int64_t> &work, std::vector<int32_t> &nlow) {
  *rl_rowptr = nullptr;
  *rl_colidx = nullptr;
  const int64_t n = g.n, nnz = g.nnz;
  PrepBufs B;
  PrepTemp T{B};
  u64 *key = nullptr, *key_out = nullptr, *ekey = nullptr, *ekey_out = nullptr, *wk = nullptr, *wk_out = nullptr;
  int32_t *ids = nullptr, *byrank = nullptr, *rank = nullptr, *nl = nullptr, *ord = nullptr;
  int64_t *len = nullptr;
  unsigned long long *npos = nullptr;
  hipError_t e;
  if ((e = B.get(&key, n)) || (e = B.get(&key_out, n)) || (e = B.get(&ids, n)) || (e = B.get(&byrank, n)) ||
      (e = B.get(&rank, n)) || (e = B.get(&len, n + 1)) || (e = B.get(&ekey, nnz)) || (e = B.get(&ekey_out, nnz)) ||
      (e = B.get(&nl, n)) || (e = B.get(&wk, n)) || (e = B.get(&wk_out, n)) || (e = B.get(&ord, n)) ||
      (e = B.get(&npos, 1)) || (e = launch_zero(npos, 1, s)))
    return e;
  int64_t *rp = nullptr;
  int32_t *ci = nullptr;
  if ((e = hipMalloc(&rp, (n + 1) * sizeof(int64_t)))) return e;
  if ((e = hipMalloc(&ci, std::max<int64_t>(1, nnz) * sizeof(int32_t)))) {
    (void)hipFree(rp);
    return e;
  }
  auto fail = [&](hipError_t x) {
    (void)hipFree(rp);
    (void)hipFree(ci);
    return x;
  };
  const int grid = prep_grid(n + 1, kPrepThreads, num_cu), wgrid = prep_grid(n, kPrepThreads / 64, num_cu);
  const int64_t sp = spread_prime(n);
  const u64 sm = 2654435761ull % (u64)sp;
  unsigned long long m = 0;
  if (n) {
    hipLaunchKernelGGL(k_rl_key, dim3(grid), dim3(kPrepThreads), 0, s, g.rowptr, slot, n, key);
    hipLaunchKernelGGL(k_iota, dim3(grid), dim3(kPrepThreads), 0, s, (uint32_t *)ids, n);
    size_t b = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, b, key, key_out, ids, byrank, (int)n, 0, 64, s)) ||
        (e = T.need(b)) || (e = hipcub::DeviceRadixSort::SortPairs(T.p, b, key, key_out, ids, byrank, (int)n, 0, 64, s)))
      return fail(e);
  }
  hipLaunchKernelGGL(k_rl_rank, dim3(grid), dim3(kPrepThreads), 0, s, g.rowptr, (const int32_t *)byrank, n, rank, len);
  size_t b = 0;
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, b, len, rp, (int)(n + 1), s)) || (e = T.need(b)) ||
      (e = hipcub::DeviceScan::ExclusiveSum(T.p, b, len, rp, (int)(n + 1), s)))
    return fail(e);
  if (n && nnz) {
    hipLaunchKernelGGL(k_rl_edges, dim3(wgrid), dim3(kPrepThreads), 0, s, g.rowptr, g.colidx,
                       (const int32_t *)byrank, (const int32_t *)rank, (const int64_t *)rp, n, ekey);
    const int rbits = n > 1 ? 64 - __builtin_clzll((unsigned long long)(n - 1)) : 1;
    b = 0;
    if ((e = hipcub::DeviceRadixSort::SortKeys(nullptr, b, ekey, ekey_out, (int)nnz, 0, 32 + rbits, s)) ||
        (e = T.need(b)) || (e = hipcub::DeviceRadixSort::SortKeys(T.p, b, ekey, ekey_out, (int)nnz, 0, 32 + rbits, s)))
      return fail(e);
    hipLaunchKernelGGL(k_rl_low, dim3(prep_grid(nnz, kPrepThreads, num_cu)), dim3(kPrepThreads), 0, s,
                       (const u64 *)ekey_out, nnz, ci);
  }
  if (n) {
    hipLaunchKernelGGL(k_rl_work, dim3(wgrid), dim3(kPrepThreads), 0, s, (const int64_t *)rp, (const int32_t *)ci, n,
                       nl, wk, npos, sp, sm);
    b = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, wk, wk_out, ids, ord, (int)n, 0, 64, s)) ||
        (e = T.need(b)) ||
        (e = hipcub::DeviceRadixSort::SortPairsDescending(T.p, b, wk, wk_out, ids, ord, (int)n, 0, 64, s)))
      return fail(e);
  }
  if ((e = hipGetLastError()) || (e = hipMemcpyAsync(&m, npos, sizeof(m), hipMemcpyDeviceToHost, s)) ||
      (e = hipStreamSynchronize(s)))
    return fail(e);
  try {
    order.resize(m);
    work.resize(n);
    nlow.resize(n);
  } catch (const std::bad_alloc &) {
    return fail(hipErrorOutOfMemory);
  }
  if ((m && (e = hipMemcpyAsync(order.data(), ord, m * sizeof(int32_t), hipMemcpyDeviceToHost, s))) ||
      (n && (e = hipMemcpyAsync(work.data(), wk, n * sizeof(int64_t), hipMemcpyDeviceToHost, s))) ||
      (n && (e = hipMemcpyAsync(nlow.data(), nl, n * sizeof(int32_t), hipMemcpyDeviceToHost, s))) ||
      (e = hipStreamSynchronize(s)))
    return fail(e);
  *rl_rowptr = rp;
  *rl_colidx = ci;
  return hipSuccess;
}

hipError_t build_rl_cuts_device(const int64_t *rl_rowptr, const int32_t *rl_colidx, int64_t n, int64_t nnz,
                                int num_cu, hipStream_t s, int32_t **cut, int32_t **nlow) {
  *cut = nullptr;
  *nlow = nullptr;
  void *c = nullptr, *l = nullptr;
  hipError_t e;
  if ((e = hipMalloc(&c, std::max<int64_t>(1, nnz) * sizeof(int32_t)))) return e;
  if ((e = hipMalloc(&l, std::max<int64_t>(1, n) * sizeof(int32_t)))) {
    (void)hipFree(c);
    return e;
  }
  const int64_t sp = spread_prime(n);
  const u64 sm = 2654435761ull % (u64)sp;
  if (n)
    hipLaunchKernelGGL(k_rl_cuts, dim3(prep_grid(n, kPrepThreads / 64, num_cu)), dim3(kPrepThreads), 0, s, rl_rowptr,
                       rl_colidx, n, static_cast<int32_t *>(c), static_cast<int32_t *>(l), sp, sm);
  if ((e = hipGetLastError()) || (e = hipStreamSynchronize(s))) {
    (void)hipFree(c);
    (void)hipFree(l);
    return e;
  }
  *cut = static_cast<int32_t *>(c);
  *nlow = static_cast<int32_t *>(l);
  return hipSuccess;
}

hipError_t cycle4_hub_counts_device(const int64_t *rl_rowptr, const int32_t *rl_colidx,
                                    const std::vector<int32_t> &hubs, const std::vector<int64_t> &hoff, int num_cu,
                                    hipStream_t s, std::vector<int32_t> &cnt) {
  const int nh = (int)hubs.size();
  const int64_t total = nh ? hoff[nh] : 0;
  cnt.assign(total, 0);
  if (!nh || !total) return hipSuccess;
  PrepBufs B;
  int32_t *dh = nullptr, *dc = nullptr;
  int64_t *doff = nullptr;
  hipError_t e;
  if ((e = B.get(&dh, nh)) || (e = B.get(&doff, nh + 1)) || (e = B.get(&dc, total)) ||
      (e = hipMemcpyAsync(dh, hubs.data(), nh * sizeof(int32_t), hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(doff, hoff.data(), (nh + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s)))
    return e;
  hipLaunchKernelGGL(k_rl_hub_counts, dim3(std::min(nh, num_cu * 4)), dim3(kPrepThreads), 0, s, rl_rowptr, rl_colidx,
                     (const int32_t *)dh, (const int64_t *)doff, nh, dc);
  if ((e = hipGetLastError()) || (e = hipMemcpyAsync(cnt.data(), dc, total * sizeof(int32_t), hipMemcpyDeviceToHost, s)))
    return e;
  return hipStreamSynchronize(s);
}

// Temporary storage (bytes) for prep_scan / prep_sort; called with
// temp == nullptr to size it.
hipError_t prep_scan(void *temp, size_t &bytes, const int32_t *in, int32_t *out, int64_t n, hipStream_t s) {
  return hipcub::DeviceScan::ExclusiveSum(temp, bytes, in, out, (int)n, s);
}

hipError_t prep_sort(void *temp, size_t &bytes, const uint32_t *key_in, uint32_t *key_out, const int32_t *ids_in,
                     int32_t *ids_out, int64_t n, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairsDescending(temp, bytes, key_in, key_out, ids_in, ids_out, (int)n, 0, 32,
                                                      s);
}

}  // namespace xgm
