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
constexpr int64_t kTcSlice = 1 << 16;

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
    const int64_t t = (int64_t)(((u64)i * spread_m) % (u64)spread_p);
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
      u64 A = 0;   // work of the open item before this chunk
      int k0 = 0;  // its first position
      for (int c0 = 0; c0 < P; c0 += 64) {
        const int k = c0 + lane;
        u64 w = 0;
        if (k < P) {
          const int x = k < d ? colidx[ob + k] : rcolidx[ib + (k - d)];
          const int dx = (int)(rowptr[x + 1] - rowptr[x]);
          if (k < d ? dx < d : dx <= d) w = (u64)dx;
        }
        const u64 S = wave_incl_scan_u64(w, lane);
        const int nvalid = min(64, P - c0);
        int ls = 0;
        u64 Sbase = 0;  // inclusive prefix at lane ls - 1
        for (;;) {
          const bool cond = lane >= ls && lane < nvalid &&
                            (A + (S - Sbase) >= (u64)kTcSlice || c0 + lane + 1 - k0 >= cap);
          const u64 m = __ballot(cond);
          if (!m) {
            A += __shfl(S, nvalid - 1, 64) - Sbase;
            break;
          }
          const int l = __builtin_ctzll(m);
          const u64 Sl = __shfl(S, l, 64);
          const u64 wk = A + (Sl - Sbase);
          if (wk) {
            if (WRITE && lane == 0) {
              items[o] = make_int4((int)t, k0, c0 + l + 1, 0);
              work[o] = (uint32_t)wk;
            }
            ++o;
            ++count;
          }
          k0 = c0 + l + 1;
          A = 0;
          Sbase = Sl;
          ls = l + 1;
        }
      }
      if (A) {
        if (WRITE && lane == 0) {
          items[o] = make_int4((int)t, k0, P, 0);
          work[o] = (uint32_t)A;
        }
        ++count;
      }
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
