// k-clique counting (k = 4, 5) on a degree-oriented DAG, gfx950.
//
// Reference semantics: GraphMiner automine_4clique / automine_5clique
// (codes/GraphMiner/src/clique/cpu_kernels/automine_omp.h:67-83, 138-157);
// G²Miner plain kernels clique{4,5}_warp_edge
// (xgminer/src/gpu_kernels/glumin/gpu_GM_LUT_kernels/GM_kernels/clique4_warp_edge.cuh:1-31,
// clique5_warp_edge.cuh:1-40); GLUMIN LUT kernels clique4_warp_vertex_subgraph /
// clique5_warp_edge_subgraph (GM_LUT_kernels/*.cuh) over BinaryEncode bitmaps
// (xgminer/include/glumin/binary_encode.h:20-30, 88-95, 255-264).
//
// Two strategies, identical counts:
//  * LUT (vertex task v0, s = d+(v0) <= SMAX): the induced sub-DAG of
//    X = N+(v0) is built as s rows of s bits in LDS (bit (i,j) set iff
//    X[j] in N+(X[i])) from flattened probes of all N+(X[i]) -- the same
//    streaming pass triangle counting makes -- and then
//      K4(v0) = #triangles of the sub-DAG = sum_{i->j} popc(row_i & row_j)
//      K5(v0) = #4-cliques            = sum_{i->j} sum_{k in row_i&row_j} popc(row_i&row_j&row_k)
//    with 64-bit words and popcount; no global scratch at all.
//  * LIST (G²Miner-style nested intersections, any degree): per v1 in N+(v0),
//    W1 = N+(v0) ∩ N+(v1) is materialised in order (block ballot compaction);
//    K4 adds sum_{w in W1} |W1 ∩ N+(w)|, K5 nests once more through
//    W2 = W1 ∩ N+(w).  Sets live in LDS when small, else in a per-block HBM
//    slot (like the reference's per-warp vlists).
#include <climits>

#include "device_common.h"
#include "kernels.h"
#include "segments.h"

namespace xgm {
namespace {

// ------------------------------------------------------------- LUT engine
// Count cliques of `depth` vertices inside the sub-DAG encoded by rows
// (s rows of W 64-bit words; only the first weff words are non-zero).
template <int W>
__device__ __forceinline__ u64 lut_count(const u64 *rows, int s, int weff, int depth, int tid, int nthreads) {
  u64 acc = 0;
  const int items = s * weff;
  for (int t = tid; t < items; t += nthreads) {
    const int i = t / weff, q = t - i * weff;
    u64 bits = rows[i * W + q];
    if (depth == 2) {
      acc += __popcll(bits);
      continue;
    }
    while (bits) {
      const int j = q * 64 + __builtin_ctzll(bits);
      bits &= bits - 1;
      if (depth == 3) {
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (w < weff) acc += __popcll(rows[i * W + w] & rows[j * W + w]);
      } else {
        u64 tw[W];
#pragma unroll
        for (int w = 0; w < W; ++w) tw[w] = (w < weff) ? (rows[i * W + w] & rows[j * W + w]) : 0ull;
#pragma unroll
        for (int w2 = 0; w2 < W; ++w2) {
          u64 tb = tw[w2];
          while (tb) {
            const int k = w2 * 64 + __builtin_ctzll(tb);
            tb &= tb - 1;
#pragma unroll
            for (int w = 0; w < W; ++w)
              if (w < weff) acc += __popcll(tw[w] & rows[k * W + w]);
          }
        }
      }
    }
  }
  return acc;
}

template <int SMAX, int W>
struct LutStage {
  int x[SMAX];
  u64 rows[SMAX * W];
  Segs<SMAX> seg;
  int ns, L, task;
};

// Vertex-task LUT kernel: one block (NT threads) per vertex v0 with
// min_s <= d+(v0) <= SMAX.  depth = k - 1 (K4 -> triangles, K5 -> 4-cliques).
template <int SMAX, int W, int NT>
__global__ __launch_bounds__(NT) void k_clique_vlut(const int64_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx,
                                                    const int32_t *__restrict__ tasks, int ntasks, int min_s,
                                                    int depth, int *__restrict__ queue, u64 *__restrict__ out) {
  __shared__ LutStage<SMAX, W> S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = NT / 64;
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.task = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.task;
    if (t >= ntasks) break;
    const int v0 = tasks[t];
    const int64_t b = rowptr[v0];
    const int s = (int)(rowptr[v0 + 1] - b);
    if (s >= min_s) {
      const int weff = (s + 63) >> 6;
      for (int k = tid; k < s; k += NT) S.x[k] = colidx[b + k];
      for (int k = tid; k < s * W; k += NT) S.rows[k] = 0ull;
      __syncthreads();
      if (wave == 0) {
        int ns, L;
        wave_segments(S.seg, rowptr, colidx, S.x, 0, s, -1, KeepAll(), lane, ns, L);
        if (lane == 0) { S.ns = ns; S.L = L; }
      }
      __syncthreads();
      const int ns = S.ns, L = S.L;
      int lo, hi;
      wave_share(L, wave, NW, lo, hi);
      wave_walk<4>(S.seg, colidx, ns, lo, hi, lane, [&](const WindowLane &r) {
        if (r.valid) {
          const int pos = lower_bound_dev(S.x, s, r.w);
          if (pos < s && S.x[pos] == r.w)
            atomicOr(&S.rows[S.seg.idx[r.j] * W + (pos >> 6)], 1ull << (pos & 63));
        }
      });
      __syncthreads();
      acc += lut_count<W>(S.rows, s, weff, depth, tid, NT);
    }
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// ------------------------------------------------------------- LIST engine
constexpr int LIST_NT = 256;
constexpr int LIST_LCAP = 2048;  // sets up to this size live in LDS
constexpr int LIST_PCAP = 1024;  // segment piece

struct ListStage {
  int h0[2 * LIST_LCAP];  // hash of N+(v0) when s <= LIST_LCAP
  int w1[LIST_LCAP];
  int w2[LIST_LCAP];
  Segs<LIST_PCAP> seg;
  int cnt[LIST_NT / 64];
  int ns, L, task;
};

struct Member {
  const int *sorted;  // generic pointer (LDS or HBM), strictly increasing
  int n;
  const int *hash;    // LDS hash or null
  int hbits;
  __device__ __forceinline__ bool contains(int w) const {
    return hash ? hash_contains(hash, hbits, w) : sorted_contains(sorted, n, w);
  }
};

// out[] <- a[0..na) ∩ m, in order; returns the size.  All threads call.
__device__ int block_intersect_ordered(const int32_t *a, int na, const Member &m, int *out, ListStage &S) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int total = 0;
  for (int base = 0; base < na; base += LIST_NT) {
    const int i = base + tid;
    int w = 0;
    bool hit = false;
    if (i < na) {
      w = a[i];
      hit = m.contains(w);
    }
    const u64 bal = __ballot(hit);
    if (lane == 0) S.cnt[wave] = __popcll(bal);
    __syncthreads();
    int off = total, all = 0;
#pragma unroll
    for (int q = 0; q < LIST_NT / 64; ++q) {
      if (q < wave) off += S.cnt[q];
      all += S.cnt[q];
    }
    if (hit) out[off + __popcll(bal & lanemask_lt(lane))] = w;
    total += all;
    __syncthreads();
  }
  return total;
}

// sum over v in verts[0..nv) of |N+(v) ∩ m| with flattened probes.
__device__ u64 block_sum_intersections(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
                                       const int *verts, int nv, const Member &m, ListStage &S) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  u64 acc = 0;
  for (int k0 = 0; k0 < nv; k0 += LIST_PCAP) {
    const int k1 = min(nv, k0 + LIST_PCAP);
    __syncthreads();
    if (wave == 0) {
      int ns, L;
      wave_segments(S.seg, rowptr, colidx, verts, k0, k1, -1, KeepAll(), lane, ns, L);
      if (lane == 0) { S.ns = ns; S.L = L; }
    }
    __syncthreads();
    const int ns = S.ns, L = S.L;
    int lo, hi;
    wave_share(L, wave, LIST_NT / 64, lo, hi);
    wave_walk<4>(S.seg, colidx, ns, lo, hi, lane, [&](const WindowLane &r) {
      if (r.valid) acc += m.contains(r.w);
    });
  }
  __syncthreads();
  return acc;
}

// LIST kernel: one block per vertex v0 (any degree).  k = 4 or 5.
// scratch: per block 2*md ints in HBM for sets larger than LIST_LCAP.
__global__ __launch_bounds__(LIST_NT) void k_clique_list(const int64_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ colidx,
                                                         const int32_t *__restrict__ tasks, int ntasks, int k,
                                                         int *__restrict__ scratch, int md,
                                                         int *__restrict__ queue, u64 *__restrict__ out) {
  __shared__ ListStage S;
  const int tid = threadIdx.x, lane = tid & 63;
  int *g_w1 = scratch + (size_t)blockIdx.x * 2 * md;
  int *g_w2 = g_w1 + md;
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.task = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.task;
    if (t >= ntasks) break;
    const int v0 = tasks[t];
    const int64_t b0 = rowptr[v0];
    const int s = (int)(rowptr[v0 + 1] - b0);
    if (s >= k - 1) {
      Member m0{colidx + b0, s, nullptr, 0};
      if (s <= LIST_LCAP) {
        const int hb = hash_bits_for(s);
        for (int q = tid; q < (1 << hb); q += LIST_NT) S.h0[q] = -1;
        __syncthreads();
        for (int q = tid; q < s; q += LIST_NT) hash_insert(S.h0, hb, colidx[b0 + q]);
        __syncthreads();
        m0.hash = S.h0;
        m0.hbits = hb;
      }
      for (int a = 0; a < s; ++a) {
        const int v1 = colidx[b0 + a];
        const int64_t b1 = rowptr[v1];
        const int d1 = (int)(rowptr[v1 + 1] - b1);
        if (d1 < k - 2) continue;
        int *w1 = (min(s, d1) <= LIST_LCAP) ? S.w1 : g_w1;
        const int n1 = block_intersect_ordered(colidx + b1, d1, m0, w1, S);
        if (n1 < k - 2) continue;
        Member m1{w1, n1, nullptr, 0};
        if (k == 4) {
          acc += block_sum_intersections(rowptr, colidx, w1, n1, m1, S);
        } else {
          for (int c = 0; c < n1; ++c) {
            const int w = w1[c];
            const int64_t bw = rowptr[w];
            const int dw = (int)(rowptr[w + 1] - bw);
            if (dw < 2) continue;
            int *w2 = (min(n1, dw) <= LIST_LCAP) ? S.w2 : g_w2;
            const int n2 = block_intersect_ordered(colidx + bw, dw, m1, w2, S);
            if (n2 < 2) continue;
            Member m2{w2, n2, nullptr, 0};
            acc += block_sum_intersections(rowptr, colidx, w2, n2, m2, S);
          }
        }
      }
    }
    __syncthreads();
  }
  (void)lane;
  acc = wave_sum_u64(acc);
  if ((tid & 63) == 0 && acc) atomicAdd(out, acc);
}

template <class K>
int grid_for(K kernel, int nt, size_t shmem, int num_cu, int64_t tasks) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, shmem) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return (int)std::max<int64_t>(1, std::min<int64_t>(tasks, (int64_t)per_cu * num_cu));
}

template <int SMAX, int W, int NT>
hipError_t run_vlut(const char *name, const DevCSR &g, const TaskList &tl, int k, int *queue, Launch &l) {
  if (!tl.count) return hipSuccess;
  auto kern = k_clique_vlut<SMAX, W, NT>;
  int grid = grid_for(kern, NT, 0, l.num_cu, tl.count);
  l.begin(name, tl.count);
  kern<<<grid, NT, 0, l.stream>>>(g.rowptr, g.colidx, tl.tasks, (int)tl.count, k - 1, k - 1, queue, l.d_out);
  l.end();
  return hipGetLastError();
}

}  // namespace

int clique_list_grid(int num_cu, int64_t tasks) {
  return grid_for(k_clique_list, LIST_NT, 0, num_cu, tasks);
}

hipError_t launch_clique(const DevCSR &g, const CliqueTasks &ct, int k, int *scratch, int md, Launch &l,
                         int *d_queues) {
  hipError_t e;
  // list path first: it holds the heaviest single tasks
  if (ct.list.count) {
    int grid = clique_list_grid(l.num_cu, ct.list.count);
    l.begin(k == 4 ? "k4_list" : "k5_list", ct.list.count);
    k_clique_list<<<grid, LIST_NT, 0, l.stream>>>(g.rowptr, g.colidx, ct.list.tasks, (int)ct.list.count, k,
                                                  scratch, md, d_queues + 0, l.d_out);
    l.end();
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const bool k4 = (k == 4);
  if ((e = run_vlut<1024, 16, 1024>(k4 ? "k4_lut_1024" : "k5_lut_1024", g, ct.lut[3], k, d_queues + 1, l)) !=
      hipSuccess)
    return e;
  if ((e = run_vlut<512, 8, 256>(k4 ? "k4_lut_512" : "k5_lut_512", g, ct.lut[2], k, d_queues + 2, l)) != hipSuccess)
    return e;
  if ((e = run_vlut<256, 4, 256>(k4 ? "k4_lut_256" : "k5_lut_256", g, ct.lut[1], k, d_queues + 3, l)) != hipSuccess)
    return e;
  if ((e = run_vlut<64, 1, 64>(k4 ? "k4_lut_64" : "k5_lut_64", g, ct.lut[0], k, d_queues + 4, l)) != hipSuccess)
    return e;
  return hipSuccess;
}

}  // namespace xgm
