// Bucket-bitmap triangle counting: X-GMiner's bucket prefilter for set
// intersection (xgminer/include/xgminer_bitmap.cuh:161-320 intersects two
// bucketed sets; :661-822 builds the buckets on the host).
//
// Every DAG out-list N+(u) is split into 64 buckets by the low six bits of
// the neighbour id, b = w & 63, each bucket kept sorted; a 64-bit mask marks
// the non-empty buckets.  |N+(u) ∩ N+(v)| is then the sum over the buckets
// set in mask(u) & mask(v) of the bucket-pair intersections: buckets empty
// on either side are never touched.  Sixty-four buckets make the mask one
// 64-bit word, so the prefilter is a single AND per edge (the reference
// takes 2^k buckets as a tuning parameter and keeps the bitmap as words).
//
// Layout (built here, in HBM): bcol[nnz] is the DAG's colidx with every
// list reordered bucket-major; boff[64 u + b] (uint16, relative to
// rowptr[u]) is where bucket b of u starts; the end of bucket 63 is d+(u).
// uint16 offsets hold because a degree-ordered DAG has d+(u) <= sqrt(2m)
// < 65536 for m < 2^31 (checked by the caller).
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace xgm {
namespace {

constexpr int kBucketThreads = 256;
constexpr int kBucketStage = 2048;  // staged bucketed list of the task vertex, ints per wave

// Lanes holding the same bucket id as this lane (6 ballots over its bits).
__device__ __forceinline__ u64 same_bucket_lanes(int b, bool valid) {
  u64 peers = __ballot(valid);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const u64 on = __ballot(valid && ((b >> i) & 1));
    peers &= ((b >> i) & 1) ? on : ~on;
  }
  return peers;
}

// One wave per vertex: bucket sizes (LDS counters), offsets by a wave scan,
// the non-empty mask by one ballot, then a stable scatter in list order
// (rank among same-bucket lanes of each 64-element window).
__global__ __launch_bounds__(kBucketThreads) void k_bucket_build(const int64_t *__restrict__ rowptr,
                                                                 const int32_t *__restrict__ colidx, int64_t n,
                                                                 int32_t *__restrict__ bcol,
                                                                 uint16_t *__restrict__ boff, u64 *__restrict__ bmap) {
  __shared__ int cnt[kBucketThreads / 64][64];
  __shared__ int fill[kBucketThreads / 64][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t u = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; u < n; u += nw) {
    const int64_t b0 = rowptr[u];
    const int d = (int)(rowptr[u + 1] - b0);
    cnt[wave][lane] = 0;
    wave_sync();
    for (int k = lane; k < d; k += 64) atomicAdd(&cnt[wave][colidx[b0 + k] & 63], 1);
    wave_sync();
    const int c = cnt[wave][lane];
    const int off = wave_incl_scan(c, lane) - c;
    fill[wave][lane] = off;
    boff[(size_t)u * 64 + lane] = (uint16_t)off;
    const u64 mask = __ballot(c > 0);
    if (lane == 0) bmap[u] = mask;
    wave_sync();
    for (int k0 = 0; k0 < d; k0 += 64) {
      const int k = k0 + lane;
      const bool valid = k < d;
      const int w = valid ? colidx[b0 + k] : 0;
      const int b = w & 63;
      const u64 peers = same_bucket_lanes(b, valid);
      const int rank = __popcll(peers & lanemask_lt(lane));
      if (valid) bcol[b0 + fill[wave][b] + rank] = w;
      wave_sync();
      if (valid && rank == 0) fill[wave][b] += __popcll(peers);
      wave_sync();
    }
  }
}

// Triangles: one wave per task vertex u (d+(u) >= 2, claimed in work
// order); the lanes take the edges (u, v), v in N+(u), 64 at a time, and
// add the bucket-pair intersections of N+(u) and N+(v) over mask(u) &
// mask(v), merging the two sorted buckets of each pair.  N+(u) in bucket
// order is staged in LDS when it fits.
__global__ __launch_bounds__(kBucketThreads) void k_tc_bucket(const int64_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ bcol,
                                                              const uint16_t *__restrict__ boff,
                                                              const u64 *__restrict__ bmap,
                                                              const int32_t *__restrict__ tasks, int ntasks,
                                                              int *__restrict__ queue, u64 *__restrict__ out,
                                                              const int32_t *__restrict__ slots) {
  __shared__ int sx[kBucketThreads / 64][kBucketStage];
  __shared__ int soff[kBucketThreads / 64][65];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u64 acc = 0;
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(queue, 1);
    t = __shfl(t, 0, 64);
    if (t >= ntasks) break;
    const int u = tasks[t];
    const int64_t b0 = rowptr[u];
    const int d = (int)(rowptr[u + 1] - b0);
    const u64 mu = bmap[u];
    const bool staged = d <= kBucketStage;
    wave_sync();  // the previous task's readers are done with sx / soff
    if (staged)
      for (int k = lane; k < d; k += 64) sx[wave][k] = bcol[b0 + k];
    soff[wave][lane] = boff[(size_t)u * 64 + lane];
    if (lane == 0) soff[wave][64] = d;
    wave_sync();
    const int32_t *xu = staged ? &sx[wave][0] : bcol + b0;
    u64 part = 0;
    for (int k = lane; k < d; k += 64) {
      const int v = staged ? xu[k] : bcol[b0 + k];
      const int64_t vb = rowptr[v];
      const int dv = (int)(rowptr[v + 1] - vb);
      const int32_t *xv = bcol + vb;
      const uint16_t *ov = boff + (size_t)v * 64;
      u64 m = mu & bmap[v];
      while (m) {
        const int b = __builtin_ctzll(m);
        m &= m - 1;
        // merge the two sorted buckets (the reference binary-searches the
        // smaller one in the larger, xgminer_bitmap.cuh:284-320; on R-MAT-22
        // that measured 275 ms against 181 ms for this merge)
        int i = soff[wave][b];
        const int ie = soff[wave][b + 1];
        int j = ov[b];
        const int je = b < 63 ? (int)ov[b + 1] : dv;
        while (i < ie && j < je) {
          const int x = xu[i], y = xv[j];
          part += (x == y);
          i += (x <= y);
          j += (y <= x);
        }
      }
    }
    task_flush(acc, part, slots, u, out, lane);
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Vertex-induced diamonds (GLUMIN P2) with X-GMiner's bucket prefilter
// (xgminer/src/gpu_kernels/xgminer/P2_kernel.cuh:110-230 builds the
// diagonal's common neighbourhood with XGMiner_BITMAP::intersection).  For
// the diagonal (v0, v1), v1 < v0, W = N(v0) ∩ N(v1) is gathered from the
// bucket pairs set in mask(v0) & mask(v1): lane b merges bucket b of both
// lists, once to count and once to write its hits at its scanned offset.
// The diamonds on the diagonal are the non-adjacent pairs of W: the sum over
// q of q - |W[0, q) ∩ N(W[q])|, every membership test first checking the
// element's bucket bit in mask(W[q]) and then searching that bucket only.
// W is in bucket order, which counts every pair once like any other order.
// A block per task v0, its waves take the v1 in turn; W lives in this
// wave's HBM slot of md ints (md >= every degree).
template <int NT>
__global__ __launch_bounds__(NT) void k_vmotif_diamond_bucket(const int64_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ bcol,
                                                              const uint16_t *__restrict__ boff,
                                                              const u64 *__restrict__ bmap,
                                                              const int32_t *__restrict__ tasks, int ntasks,
                                                              int32_t *__restrict__ scratch, int md,
                                                              int *__restrict__ queue, u64 *__restrict__ out,
                                                              const int32_t *__restrict__ slots) {
  __shared__ int sh_task, sh_next;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int32_t *W = scratch + ((size_t)blockIdx.x * (NT / 64) + wave) * (size_t)md;
  u64 acc = 0;
  for (;;) {
    if (tid == 0) {
      sh_task = atomicAdd(queue, 1);
      sh_next = 0;
    }
    __syncthreads();
    const int t = sh_task;
    if (t >= ntasks) break;
    const int v0 = tasks[t];
    const int64_t b0 = rowptr[v0];
    const int d0 = (int)(rowptr[v0 + 1] - b0);
    const int32_t *x0 = bcol + b0;
    const uint16_t *o0 = boff + (size_t)v0 * 64;
    const u64 m0 = bmap[v0];
    u64 part = 0;
    for (;;) {
      int i = 0;
      if (lane == 0) i = atomicAdd(&sh_next, 1);
      i = __shfl(i, 0, 64);
      if (i >= d0) break;
      const int v1 = x0[i];
      if (v1 >= v0) continue;
      const int64_t b1 = rowptr[v1];
      const int d1 = (int)(rowptr[v1 + 1] - b1);
      const int32_t *x1 = bcol + b1;
      const uint16_t *o1 = boff + (size_t)v1 * 64;
      const bool on = ((m0 & bmap[v1]) >> lane) & 1;
      int s0 = 0, e0 = 0, s1 = 0, e1 = 0, c = 0;
      if (on) {
        s0 = o0[lane];
        e0 = lane < 63 ? (int)o0[lane + 1] : d0;
        s1 = o1[lane];
        e1 = lane < 63 ? (int)o1[lane + 1] : d1;
        for (int p = s0, r = s1; p < e0 && r < e1;) {
          const int x = x0[p], y = x1[r];
          c += x == y;
          p += x <= y;
          r += y <= x;
        }
      }
      const int incl = wave_incl_scan(c, lane);
      const int n0 = __shfl(incl, 63, 64);
      if (c) {
        int o = incl - c;
        for (int p = s0, r = s1; p < e0 && r < e1;) {
          const int x = x0[p], y = x1[r];
          if (x == y) W[o++] = x;
          p += x <= y;
          r += y <= x;
        }
      }
      __threadfence_block();  // W written by other lanes of this wave
      wave_sync();
      uint32_t cnt = 0;
      for (int q = 1; q < n0; ++q) {
        const int v2 = W[q];
        const int64_t b2 = rowptr[v2];
        const int d2 = (int)(rowptr[v2 + 1] - b2);
        const uint16_t *o2 = boff + (size_t)v2 * 64;
        const u64 m2 = bmap[v2];
        uint32_t hit = 0;
        for (int k = lane; k < q; k += 64) {
          const int w = W[k];
          const int bb = w & 63;
          if ((m2 >> bb) & 1) {
            const int s = o2[bb], e = bb < 63 ? (int)o2[bb + 1] : d2;
            hit += sorted_contains(bcol + b2 + s, e - s, w) ? 1u : 0u;
          }
        }
        cnt += (uint32_t)q - (uint32_t)wave_sum_i32((int)hit);
      }
      if (lane == 0) part += cnt;
      wave_sync();  // W reads done before the next v1 overwrites it
    }
    task_flush(acc, part, slots, v0, out, lane);
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

}  // namespace

int vmotif_bucket_grid(int num_cu) { return num_cu * 2; }

hipError_t launch_vmotif_diamond_bucket(const DevCSR &g, const int32_t *bcol, const uint16_t *boff, const u64 *bmap,
                                        const TaskList &tasks, int32_t *scratch, int md, Launch &l, int *queue) {
  if (!tasks.count) return hipSuccess;
  const int grid = (int)std::min<int64_t>(tasks.count, vmotif_bucket_grid(l.num_cu));
  l.begin("vm_diamond_bucket", tasks.count);
  k_vmotif_diamond_bucket<256><<<grid, 256, 0, l.stream>>>(g.rowptr, bcol, boff, bmap, tasks.tasks,
                                                           (int)tasks.count, scratch, md, queue, l.d_out, l.slots);
  l.end();
  return hipGetLastError();
}

hipError_t launch_bucket_build(const int64_t *rowptr, const int32_t *colidx, int64_t n, int32_t *bcol, uint16_t *boff,
                               u64 *bmap, int num_cu, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int64_t want = (n + 3) / 4;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)num_cu * 16));
  k_bucket_build<<<grid, kBucketThreads, 0, s>>>(rowptr, colidx, n, bcol, boff, bmap);
  return hipGetLastError();
}

hipError_t launch_tc_bucket(const DevCSR &g, const int32_t *bcol, const uint16_t *boff, const u64 *bmap,
                            const TaskList &tasks, Launch &l, int *queue) {
  if (!tasks.count) return hipSuccess;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tc_bucket, kBucketThreads, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int64_t want = (tasks.count + 3) / 4;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)per_cu * l.num_cu));
  l.begin("tc_bucket", tasks.count);
  k_tc_bucket<<<grid, kBucketThreads, 0, l.stream>>>(g.rowptr, bcol, boff, bmap, tasks.tasks, (int)tasks.count,
                                                     queue, l.d_out, l.slots);
  l.end();
  return hipGetLastError();
}

}  // namespace xgm
