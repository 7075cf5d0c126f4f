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
