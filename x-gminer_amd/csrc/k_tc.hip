// Triangle counting on a degree-oriented DAG, gfx950.
//
// Reference semantics: GraphMiner TCSolver sums |N+(u) ∩ N+(v)| over every
// DAG edge u->v (codes/GraphMiner/src/triangle/omp_base.cc:15-21); the GPU
// reference gives one warp per edge and binary-searches the shorter list in
// the longer (codes/GraphMiner/src/triangle/gpu_kernels/bs_warp_edge.cuh:1-18,
// xgminer/include/glumin/set_intersect.cuh:440-465).
//
// Design here: one task per vertex u.  N+(u) is staged once in LDS (as an
// open-addressing hash, or as the sorted list for binary search) and every
// element w of every N+(v), v in N+(u), is probed against it.  The probes of
// all the v-segments are flattened into one index space [0, L(u)), so a
// wave64 keeps all 64 lanes busy even when the segments are short; lane ->
// segment is resolved per 64-wide window from a bitmask of segment ends.
// Global reads are the concatenated segments, coalesced within a segment.
#include <climits>

#include "device_common.h"
#include "kernels.h"
#include "segments.h"

namespace xgm {
namespace {

constexpr int STRAT_HASH = 1;
constexpr int STRAT_BSEARCH = 2;
constexpr int UNROLL = 4;  // windows mapped before their loads are issued

// LDS state of one staged task: N+(u) as a hash (load <= 1/4) or sorted
// copy, and the flattened segment table of the lists N+(v), v in N+(u).
template <int HCAP, int PCAP, int STRAT, int HSHIFT>
struct Stage {
  static constexpr int KEYS = (STRAT == STRAT_HASH) ? (HCAP << HSHIFT) : HCAP;
  int key[KEYS];
  Segs<PCAP, false> seg;
};

template <int STRAT>
__device__ __forceinline__ int member(const int *key, int hbits, const int32_t *nu_global, int du, bool lds, int w) {
  if (!lds) return sorted_contains(nu_global, du, w);
  if (STRAT == STRAT_HASH) return hash_contains(key, hbits, w);
  return sorted_contains(key, du, w);
}

// ---------------------------------------------------------------------------
// Wave-per-task kernel: 4 independent waves per block, each owning a Stage.
template <int CAP, int STRAT, int CLAIM, int HSHIFT>
__global__ __launch_bounds__(256) void k_tc_wave(const int64_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ colidx,
                                                 const int32_t *__restrict__ tasks, int ntasks,
                                                 int *__restrict__ queue, u64 *__restrict__ out) {
  __shared__ Stage<CAP, CAP, STRAT, HSHIFT> stage[4];
  __shared__ int flagrow[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  Stage<CAP, CAP, STRAT, HSHIFT> &s = stage[wave];
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    int t0 = 0;
    if (lane == 0) t0 = atomicAdd(queue, CLAIM);
    t0 = __shfl(t0, 0, 64);
    if (t0 >= ntasks) break;
    const int t1 = min(t0 + CLAIM, ntasks);
    for (int t = t0; t < t1; ++t) {
      const int u = tasks[t];
      const int64_t b = rowptr[u];
      const int du = (int)(rowptr[u + 1] - b);
      const int32_t *nu = colidx + b;
      int hbits = 0;
      if (STRAT == STRAT_HASH) {
        hbits = hash_bits_for(du, HSHIFT);
        for (int k = lane; k < (1 << hbits); k += 64) s.key[k] = -1;
        wave_sync();
        for (int k = lane; k < du; k += 64) hash_insert(s.key, hbits, nu[k]);
      } else {
        for (int k = lane; k < du; k += 64) s.key[k] = nu[k];
      }
      int ns, L;
      wave_segments(s.seg, rowptr, colidx, nu, 0, du, -1, KeepAll(), lane, ns, L);
      wave_sync();
      int cnt = 0;
      wave_walk<UNROLL>(s.seg, colidx, ns, 0, L, lane, fl, [&](const WindowLane &r) {
        if (r.valid) cnt += member<STRAT>(s.key, hbits, nu, du, true, r.w);
      });
      acc += (u64)cnt;
      wave_sync();  // stage is rewritten by the next task
    }
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// ---------------------------------------------------------------------------
// Block-per-task kernel: NT/64 waves share one Stage, build the segment
// table together and walk contiguous shares of it.  If d+(u) > CAP the
// membership test binary-searches N+(u) in HBM and the segments are
// processed in pieces of CAP.
template <int CAP, int PCAP, int STRAT, int NT, int HSHIFT>
__global__ __launch_bounds__(NT) void k_tc_block(const int64_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ colidx,
                                                 const int32_t *__restrict__ tasks, int ntasks,
                                                 int *__restrict__ queue, u64 *__restrict__ out) {
  __shared__ Stage<CAP, PCAP, STRAT, HSHIFT> s;
  __shared__ int tmp[PCAP];
  __shared__ int flagrow[NT / 64][64];
  __shared__ int sh_task, sh_ns, sh_L;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NW = NT / 64;
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    if (threadIdx.x == 0) sh_task = atomicAdd(queue, 1);
    __syncthreads();
    const int t = sh_task;
    if (t >= ntasks) break;
    const int u = tasks[t];
    const int64_t b = rowptr[u];
    const int du = (int)(rowptr[u + 1] - b);
    const int32_t *nu = colidx + b;
    const bool lds_member = du <= CAP;
    int hbits = 0;
    if (lds_member) {
      if (STRAT == STRAT_HASH) {
        hbits = hash_bits_for(du, HSHIFT);
        for (int k = threadIdx.x; k < (1 << hbits); k += NT) s.key[k] = -1;
        __syncthreads();
        for (int k = threadIdx.x; k < du; k += NT) hash_insert(s.key, hbits, nu[k]);
      } else {
        for (int k = threadIdx.x; k < du; k += NT) s.key[k] = nu[k];
      }
    }
    int cnt = 0;
    for (int k0 = 0; k0 < du; k0 += PCAP) {
      const int k1 = min(du, k0 + PCAP);
      __syncthreads();  // previous piece fully consumed
      block_segments<NT>(s.seg, tmp, rowptr, colidx, nu, k0, k1, -1, KeepAll(), &sh_ns, &sh_L);
      int lo, hi;
      wave_share(sh_L, wave, NW, lo, hi);
      wave_walk<UNROLL>(s.seg, colidx, sh_ns, lo, hi, lane, fl, [&](const WindowLane &r) {
        if (r.valid) cnt += member<STRAT>(s.key, hbits, nu, du, lds_member, r.w);
      });
    }
    acc += (u64)cnt;
    __syncthreads();  // sh_task and stage reused by the next claim
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

__global__ void k_zero_u64(u64 *p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <class K>
int persistent_grid(K kernel, int num_cu, int64_t tasks, int tasks_per_block, int nt = 256) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  int64_t want = (tasks + tasks_per_block - 1) / tasks_per_block;
  int64_t cap = (int64_t)per_cu * num_cu;
  return (int)std::max<int64_t>(1, std::min(want, cap));
}

template <int STRAT>
hipError_t launch_tc_strat(const DevCSR &g, const TaskList tiers[NUM_TIERS], Launch &l, int *d_queues) {
  hipError_t e = hipSuccess;
  // heavy tiers first so their long tasks overlap with nothing idle at the end
  if (tiers[TIER_BIG].count) {
    auto k = k_tc_block<2048, 512, STRAT, 512, 2>;
    int grid = persistent_grid(k, l.num_cu, tiers[TIER_BIG].count, 1, 512);
    l.begin("tc_block_big", tiers[TIER_BIG].count);
    k<<<grid, 512, 0, l.stream>>>(g.rowptr, g.colidx, tiers[TIER_BIG].tasks, (int)tiers[TIER_BIG].count,
                                  d_queues + TIER_BIG, l.d_out);
    l.end();
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (tiers[TIER_B2048].count) {
    auto k = k_tc_block<2048, 512, STRAT, 512, 2>;
    int grid = persistent_grid(k, l.num_cu, tiers[TIER_B2048].count, 1, 512);
    l.begin("tc_block_2048", tiers[TIER_B2048].count);
    k<<<grid, 512, 0, l.stream>>>(g.rowptr, g.colidx, tiers[TIER_B2048].tasks,
                                  (int)tiers[TIER_B2048].count, d_queues + TIER_B2048, l.d_out);
    l.end();
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (tiers[TIER_W256].count) {
    auto k = k_tc_wave<256, STRAT, 1, 2>;
    int grid = persistent_grid(k, l.num_cu, tiers[TIER_W256].count, 4);
    l.begin("tc_wave_256", tiers[TIER_W256].count);
    k<<<grid, 256, 0, l.stream>>>(g.rowptr, g.colidx, tiers[TIER_W256].tasks,
                                  (int)tiers[TIER_W256].count, d_queues + TIER_W256, l.d_out);
    l.end();
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (tiers[TIER_W64].count) {
    auto k = k_tc_wave<64, STRAT, 8, 2>;
    int grid = persistent_grid(k, l.num_cu, tiers[TIER_W64].count, 32);
    l.begin("tc_wave_64", tiers[TIER_W64].count);
    k<<<grid, 256, 0, l.stream>>>(g.rowptr, g.colidx, tiers[TIER_W64].tasks, (int)tiers[TIER_W64].count,
                                  d_queues + TIER_W64, l.d_out);
    l.end();
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return e;
}

}  // namespace

hipError_t launch_zero(void *ptr, size_t n_u64, hipStream_t s) {
  if (n_u64 == 0) return hipSuccess;
  unsigned blocks = (unsigned)((n_u64 + 255) / 256);
  k_zero_u64<<<blocks, 256, 0, s>>>((u64 *)ptr, n_u64);
  return hipGetLastError();
}

hipError_t launch_tc(const DevCSR &g, const TaskList tiers[NUM_TIERS], int strategy, Launch &l,
                     int *d_queues) {
  if (strategy == STRAT_BSEARCH) return launch_tc_strat<STRAT_BSEARCH>(g, tiers, l, d_queues);
  return launch_tc_strat<STRAT_HASH>(g, tiers, l, d_queues);
}

}  // namespace xgm
