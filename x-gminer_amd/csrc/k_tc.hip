// Triangle counting on a degree-oriented DAG, gfx950.
//
// Reference semantics: GraphMiner TCSolver sums |N+(u) ∩ N+(v)| over every
// DAG edge u->v (codes/GraphMiner/src/triangle/omp_base.cc:15-21); the GPU
// reference gives one warp per edge and binary-searches the shorter list in
// the longer (codes/GraphMiner/src/triangle/gpu_kernels/bs_warp_edge.cuh:1-18,
// xgminer/include/glumin/set_intersect.cuh:440-465).
//
// Design here: every DAG edge u->v is charged to the endpoint t whose
// out-list is longer (ties to v), and only the shorter list is streamed:
// per edge the work is min(d+(u), d+(v)) probes instead of d+(v).  A task
// item (t, k0, k1) stages N+(t) once in LDS (an open-addressing hash at load
// <= 1/4, or the sorted list for binary search; in HBM above the LDS cap)
// and probes every element of N+(x) for the partners x in positions
// [k0, k1) of the virtual list N+(t) ++ N-(t) (N- = in-neighbours, from the
// reverse DAG) that were charged to t.  The probes of all partner lists are
// flattened into one index space so a wave64 keeps all lanes busy; global
// reads are the concatenated partner lists, coalesced within a list.
#include <climits>

#include "device_common.h"
#include "kernels.h"
#include "merge_path.h"
#include "segments.h"

namespace xgm {
namespace {

constexpr int STRAT_HASH = 1;
constexpr int STRAT_BSEARCH = 2;
constexpr int UNROLL = 4;  // windows mapped before their loads are issued
// partners per segment-table piece of the block kernel: 480 keeps its LDS
// (hash 32 KB + table + flags) at 40,596 B, so four 512-thread blocks fit
// a CU's 160 KB and every SIMD runs 8 waves
constexpr int kBlockPieces = 480;

// LDS state of one staged task: N+(u) as a hash (load <= 1/4) or sorted
// copy, and the flattened segment table of the lists N+(v), v in N+(u).
template <int HCAP, int PCAP, int STRAT, int HSHIFT>
struct Stage {
  static constexpr int KEYS = (STRAT == STRAT_HASH) ? (HCAP << HSHIFT) : HCAP;
  alignas(16) int key[KEYS];  // hash: KEYS / 4 buckets of 4 slots
  Segs<PCAP, false> seg;
};

// Hash hits of a group of U windows whose invalid lanes carry kNoKey: the U
// home buckets are read together (one LDS round trip for the group), and
// hits and the full-bucket misses are found with vector integer ops only --
// no lane mask is formed per window, so the scalar unit waits on no vector
// compare until the one wave-uniform branch per group into the rare
// continuation (bhash_contains_after).  Returns this lane's hits.
template <int U>
__device__ __forceinline__ int group_hash_hits(const int *key, int hbits, const WindowLane (&r)[U]) {
  int4 q[U];
#pragma unroll
  for (int u = 0; u < U; ++u) q[u] = reinterpret_cast<const int4 *>(key)[hash_slot(r[u].w, hbits)];
  int c = 0;
  uint32_t open = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t d = bucket_dist(q[u], r[u].w);
    c += d == 0u ? 1 : 0;
    open |= q[u].w >= 0 ? d : 0u;  // nonzero: a miss in a full home bucket
  }
  if (__builtin_expect(__ballot(open != 0u) != 0, 0)) {
    // home buckets read again rather than kept in registers across the loop
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t h = hash_slot(r[u].w, hbits);
      const int4 q0 = reinterpret_cast<const int4 *>(key)[h];
      if (q0.w >= 0 && !bucket_has(q0, r[u].w)) c += bhash_contains_after(key, hbits, h, r[u].w) ? 1 : 0;
    }
  }
  return c;
}

// hbits: log2 of the hash table's buckets (4 slots each)
template <int STRAT>
__device__ __forceinline__ int member(const int *key, int hbits, const int32_t *nu_global, int du, bool lds, int w) {
  if (!lds) return sorted_contains(nu_global, du, w);
  if (STRAT == STRAT_HASH) return bhash_contains(key, hbits, w);
  return sorted_contains(key, du, w);
}

// Partner k of t in the virtual list N+(t) ++ N-(t), and whether the edge
// is charged to t: out-partner x iff d+(x) < d+(t), in-partner iff d+(x) <= d+(t).
struct TcVerts {
  const int32_t *out, *in;
  int dout;
  __device__ __forceinline__ int operator()(int k) const { return k < dout ? out[k] : in[k - dout]; }
};
struct TcKeep {
  int dout;
  __device__ __forceinline__ bool operator()(int k, int, int len) const { return k < dout ? len < dout : len <= dout; }
};

// ---------------------------------------------------------------------------
// Wave-per-item kernel: 4 independent waves per block, each owning a Stage.
// Items (t, k0, k1, -) with d+(t) <= CAP and k1 - k0 <= CAP.
template <int CAP, int STRAT, int HSHIFT>
__global__ __launch_bounds__(256) void k_tc_wave(const int64_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ colidx,
                                                 const int64_t *__restrict__ rrowptr,
                                                 const int32_t *__restrict__ rcolidx,
                                                 const int4 *__restrict__ items, int nitems,
                                                 int *__restrict__ queue, u64 *__restrict__ out,
                                                 const int32_t *__restrict__ slots) {
  __shared__ Stage<CAP, CAP, STRAT, HSHIFT> stage[4];
  __shared__ int flagrow[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  Stage<CAP, CAP, STRAT, HSHIFT> &s = stage[wave];
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(queue, 1);
    t = __shfl(t, 0, 64);
    if (t >= nitems) break;
    const int4 it = items[t];
    const int u = it.x;
    const int64_t b = rowptr[u];
    const int du = (int)(rowptr[u + 1] - b);
    const int32_t *nu = colidx + b;
    int hbits = 0;
    if (STRAT == STRAT_HASH) {
      // the whole table (CAP << HSHIFT slots, allocated anyway): load
      // du / slots, so missing probes run shorter than at the 2^-HSHIFT cap
      hbits = 31 - __clz(CAP << HSHIFT) - 2;  // buckets of 4 slots
      for (int k = lane; k < (4 << hbits); k += 64) s.key[k] = -1;
      wave_sync();
      for (int k = lane; k < du; k += 64) bhash_insert(s.key, hbits, nu[k]);
    } else {
      for (int k = lane; k < du; k += 64) s.key[k] = nu[k];
    }
    int ns = 0, L = 0;
    wave_segments_fn(s.seg, rowptr, it.y, it.z, TcVerts{nu, rcolidx + rrowptr[u], du}, TcKeep{du}, lane, ns, L);
    wave_sync();
    int cnt = 0;
    wave_walk_group<UNROLL, false, kNoKey>(s.seg, colidx, ns, 0, L, lane, fl, [&](const WindowLane (&r)[UNROLL]) {
      if (STRAT == STRAT_HASH) {
        cnt += group_hash_hits<UNROLL>(s.key, hbits, r);
      } else {
#pragma unroll
        for (int q = 0; q < UNROLL; ++q)
          if (r[q].valid) cnt += member<STRAT>(s.key, hbits, nu, du, true, r[q].w);
      }
    });
    task_flush(acc, (u64)cnt, slots, u, out, lane);
    wave_sync();  // stage is rewritten by the next item
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// ---------------------------------------------------------------------------
// Packed wave kernel for the smallest items (d+(t) <= 64, <= 64 partners):
// a wave claims G items with one atomic and runs them as one task -- the
// staged lists of all G items go into G hash regions through one flattened
// insert pass, and the partners of all G items form one segment table -- so
// the dependent loads of G items (item, rowptr, lists, partner rowptrs,
// partner lists) are in flight together instead of one item's chain at a
// time.
template <int G, int HSHIFT>
struct PackStage {
  static constexpr int SLOTS = 64 << HSHIFT;  // per item
  static constexpr int BBITS = 4 + HSHIFT;     // the whole region in buckets of 4 slots
  alignas(16) int key[G * SLOTS];
  Segs<64 * G, true> seg;                     // idx[] = item of the segment
  int u[G], k0[G], n[G], du[G], bits[G], cnt[G];
  int poff[G + 1], doff[G + 1];               // prefix sums of partners / staged lengths
  int64_t ob[G], ib[G];                        // rowptr[u], rrowptr[u]
};

template <int G, int HSHIFT>
__global__ __launch_bounds__(128) void k_tc_wave_packed(const int64_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ colidx,
                                                        const int64_t *__restrict__ rrowptr,
                                                        const int32_t *__restrict__ rcolidx,
                                                        const int4 *__restrict__ items, int nitems,
                                                        int *__restrict__ queue, u64 *__restrict__ out,
                                                        const int32_t *__restrict__ slots) {
  using PS = PackStage<G, HSHIFT>;
  __shared__ PS stage[2];
  __shared__ int flagrow[2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  PS &s = stage[wave];
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(queue, G);
    base = __shfl(base, 0, 64);
    if (base >= nitems) break;
    const int ng = min(G, nitems - base);
    // item metadata, one lane per item
    int n = 0, du = 0;
    if (lane < ng) {
      const int4 it = items[base + lane];
      const int64_t b = rowptr[it.x];
      du = (int)(rowptr[it.x + 1] - b);
      n = it.z - it.y;
      s.u[lane] = it.x;
      s.k0[lane] = it.y;
      s.n[lane] = n;
      s.du[lane] = du;
      s.ob[lane] = b;
      s.ib[lane] = rrowptr[it.x];
      s.cnt[lane] = 0;
    }
    const int pin = wave_incl_scan(n, lane), din = wave_incl_scan(du, lane);
    if (lane < ng) {
      s.poff[lane + 1] = pin;
      s.doff[lane + 1] = din;
    }
    if (lane == 0) { s.poff[0] = 0; s.doff[0] = 0; }
    for (int k = lane; k < ng * PS::SLOTS; k += 64) s.key[k] = -1;
    wave_sync();
    const int P = __shfl(pin, ng - 1, 64), D = __shfl(din, ng - 1, 64);
    auto item_of = [&](int q, const int *off) {
      int g = 0;
#pragma unroll
      for (int h = 1; h < G; ++h) g += (h < ng && q >= off[h]) ? 1 : 0;
      return g;
    };
    // staged lists N+(u_g) into hash region g
    for (int q = lane; q < D; q += 64) {
      const int g = item_of(q, s.doff);
      bhash_insert(s.key + g * PS::SLOTS, PS::BBITS, colidx[s.ob[g] + (q - s.doff[g])]);
    }
    // one segment table over the partners of all items
    int ns = 0, L = 0;
    wave_segments_fn(s.seg, rowptr, 0, P,
                     [&](int q) {  // partner k of u_g in N+(u_g) ++ N-(u_g)
                       const int g = item_of(q, s.poff);
                       const int k = s.k0[g] + q - s.poff[g];
                       return k < s.du[g] ? colidx[s.ob[g] + k] : rcolidx[s.ib[g] + (k - s.du[g])];
                     },
                     [&](int q, int, int len) {  // charged to u_g (the TcKeep rule)
                       const int g = item_of(q, s.poff);
                       const int k = s.k0[g] + q - s.poff[g];
                       return k < s.du[g] ? len < s.du[g] : len <= s.du[g];
                     },
                     lane, ns, L);
    if (lane == 0) s.seg.os[ns] = make_int2(0, L);
    wave_sync();
    for (int j = lane; j < ns; j += 64) s.seg.idx[j] = item_of(s.seg.idx[j], s.poff);
    wave_sync();
    int cnt = 0;
    wave_walk<UNROLL, true>(s.seg, colidx, ns, 0, L, lane, fl, [&](const WindowLane &r) {
      bool hit = false;
      if (r.valid) {
        const int g = s.seg.idx[r.j];
        hit = bhash_contains(s.key + g * PS::SLOTS, PS::BBITS, r.w);
      }
      if (!slots) {
        cnt += hit;
      } else {
        const int h = segment_run_hits(__ballot(hit), r, lane);
        if (h) atomicAdd(&s.cnt[s.seg.idx[r.j]], h);
      }
    });
    if (!slots) {
      acc += (u64)cnt;
    } else {
      wave_sync();
      if (lane < ng && s.cnt[lane]) atomicAdd(out + slots[s.u[lane]], (u64)s.cnt[lane]);
    }
    wave_sync();  // stage is rewritten by the next claim
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// ---------------------------------------------------------------------------
// Block-per-item kernel: NT/64 waves share one Stage, build the segment
// table together and walk contiguous shares of it.  If d+(t) > CAP the
// membership test binary-searches N+(t) in HBM.  Partners are taken in
// pieces of PCAP.
template <int CAP, int PCAP, int STRAT, int NT, int HSHIFT>
__global__ __launch_bounds__(NT, 8) void k_tc_block(const int64_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ colidx,
                                                 const int64_t *__restrict__ rrowptr,
                                                 const int32_t *__restrict__ rcolidx,
                                                 const int4 *__restrict__ items, int nitems,
                                                 int *__restrict__ queue, u64 *__restrict__ out,
                                                 const int32_t *__restrict__ slots) {
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
    const int t = __builtin_amdgcn_readfirstlane(sh_task);  // uniform: the item is read with scalar loads
    if (t >= nitems) break;
    const int4 it = items[t];
    const int u = it.x;
    const int64_t b = rowptr[u];
    const int du = (int)(rowptr[u + 1] - b);
    const int32_t *nu = colidx + b;
    const bool lds_member = du <= CAP;
    int hbits = 0;
    if (lds_member) {
      if (STRAT == STRAT_HASH) {
        hbits = 31 - __clz(CAP << HSHIFT) - 2;  // the whole table in buckets of 4, as in k_tc_wave
        for (int k = threadIdx.x; k < (4 << hbits); k += NT) s.key[k] = -1;
        __syncthreads();
        for (int k = threadIdx.x; k < du; k += NT) bhash_insert(s.key, hbits, nu[k]);
      } else {
        for (int k = threadIdx.x; k < du; k += NT) s.key[k] = nu[k];
      }
    }
    const TcVerts vert{nu, rcolidx + rrowptr[u], du};
    int cnt = 0;
    for (int k0 = it.y; k0 < it.z; k0 += PCAP) {
      const int k1 = min(it.z, k0 + PCAP);
      __syncthreads();  // previous piece fully consumed
      block_segments_fn<NT>(s.seg, tmp, rowptr, k0, k1, vert, TcKeep{du}, &sh_ns, &sh_L);
      int lo, hi;
      wave_share(sh_L, wave, NW, lo, hi);
      wave_walk_group<UNROLL, true, kNoKey>(s.seg, colidx, sh_ns, lo, hi, lane, fl, [&](const WindowLane (&r)[UNROLL]) {
        if (STRAT == STRAT_HASH && lds_member) {
          cnt += group_hash_hits<UNROLL>(s.key, hbits, r);
        } else {
#pragma unroll
          for (int q = 0; q < UNROLL; ++q)
            if (r[q].valid) cnt += member<STRAT>(s.key, hbits, nu, du, lds_member, r[q].w);
        }
      });
    }
    task_flush(acc, (u64)cnt, slots, u, out, lane);
    __syncthreads();  // sh_task and stage reused by the next claim
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Merge strategy: a wave per task vertex u (d+(u) >= 2), N+(u) staged in
// LDS up to kMergeStage elements; each lane takes one partner v in N+(u) and
// merges N+(u) with N+(v) element by element (GraphMiner's merge count,
// codes/GraphMiner/include/VertexSet.h:65-76, one lane per DAG edge); a pair
// with both lists above kHybridMin is merged by the whole wave along the
// merge path instead (the reference's hybrid rule, operations.cuh:208-300).
constexpr int kMergeStage = 1024;
constexpr int kMergeWaves = 4;
struct MergeStage {
  int nu[kMergeStage];
  int As[kMergeTile];
  int Bs[kMergeTile + 1];
  int split[64];
};

__global__ __launch_bounds__(64 * kMergeWaves) void k_tc_merge(const int64_t *__restrict__ rowptr,
                                                                const int32_t *__restrict__ colidx,
                                                                const int32_t *__restrict__ tasks, int ntasks,
                                                                int *__restrict__ queue, u64 *__restrict__ out,
                                                                const int32_t *__restrict__ slots) {
  __shared__ MergeStage stage[kMergeWaves];
  const int lane = threadIdx.x & 63;
  MergeStage &S = stage[threadIdx.x >> 6];
  u64 acc = 0;
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(queue, 1);
    t = __shfl(t, 0, 64);
    if (t >= ntasks) break;
    const int u = tasks[t];
    const int64_t ub = rowptr[u];
    const int du = (int)(rowptr[u + 1] - ub);
    const bool lds = du <= kMergeStage;
    wave_sync();  // the previous task's reads of S.nu are done
    if (lds)
      for (int i = lane; i < du; i += 64) S.nu[i] = colidx[ub + i];
    wave_sync();
    const int *Nu = lds ? S.nu : colidx + ub;
    uint32_t cnt = 0;
    for (int k0 = 0; k0 < du; k0 += 64) {
      const int k = k0 + lane;
      int v = 0;
      bool wave_pair = false;
      if (k < du) {
        v = Nu[k];
        const int64_t vb = rowptr[v];
        const int dv = (int)(rowptr[v + 1] - vb);
        wave_pair = du > kHybridMin && dv > kHybridMin;
        if (!wave_pair) {
          const int32_t *Nv = colidx + vb;
          int i = 0, j = 0;
          while (i < du && j < dv) {
            const int x = Nu[i], y = Nv[j];
            cnt += (x == y);
            i += (x <= y);
            j += (y <= x);
          }
        }
      }
      u64 m = __ballot(wave_pair);
      while (m) {
        const int src = __builtin_ctzll(m);
        m &= m - 1;
        const int vv = __shfl(v, src, 64);
        const int64_t vb = rowptr[vv];
        const uint32_t c = merge_walk(Nu, du, colidx + vb, (int)(rowptr[vv + 1] - vb), S.As, S.Bs, S.split, lane,
                                      [](int, bool found) { return found; }, nullptr);
        if (lane == 0) cnt += c;
      }
    }
    task_flush(acc, (u64)cnt, slots, u, out, lane);
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

template <class K>
hipError_t run_tc(const char *name, K kernel, int nt, const DevCSR &g, const TaskList &items, int *queue, Launch &l) {
  if (!items.count) return hipSuccess;
  int grid = persistent_grid(kernel, l.num_cu, items.count, nt / 64 > 1 && nt == 256 ? 4 : 1, nt);
  l.begin(name, items.count);
  kernel<<<grid, nt, 0, l.stream>>>(g.rowptr, g.colidx, g.rrowptr, g.rcolidx, (const int4 *)items.tasks,
                                    (int)items.count, queue, l.d_out, l.slots);
  l.end();
  return hipGetLastError();
}

template <int STRAT, int HS>
hipError_t launch_tc_strat(const DevCSR &g, const TaskList tiers[NUM_TIERS], Launch &l, int *d_queues) {
  hipError_t e;
  // heavy tiers first so their long items overlap with nothing idle at the end
  if ((e = run_tc("tc_block_big", k_tc_block<2048, kBlockPieces, STRAT, 512, HS>, 512, g, tiers[TIER_BIG], d_queues + TIER_BIG,
                  l)))
    return e;
  if ((e = run_tc("tc_block_2048", k_tc_block<2048, kBlockPieces, STRAT, 512, HS>, 512, g, tiers[TIER_B2048],
                  d_queues + TIER_B2048, l)))
    return e;
  if ((e = run_tc("tc_wave_256", k_tc_wave<256, STRAT, HS>, 256, g, tiers[TIER_W256], d_queues + TIER_W256, l)))
    return e;
  if (STRAT == STRAT_HASH) {
    // smallest items: 8 per wave, two waves per block
    constexpr int kPack = 8;
    const TaskList &it = tiers[TIER_W64];
    if (it.count) {
      auto k = k_tc_wave_packed<kPack, HS>;
      const int grid = persistent_grid(k, l.num_cu, it.count, 2 * kPack, 128);
      l.begin("tc_wave_64", it.count);
      k<<<grid, 128, 0, l.stream>>>(g.rowptr, g.colidx, g.rrowptr, g.rcolidx, (const int4 *)it.tasks,
                                    (int)it.count, d_queues + TIER_W64, l.d_out, l.slots);
      l.end();
      if ((e = hipGetLastError())) return e;
    }
    return hipSuccess;
  }
  if ((e = run_tc("tc_wave_64", k_tc_wave<64, STRAT, HS>, 256, g, tiers[TIER_W64], d_queues + TIER_W64, l)))
    return e;
  return hipSuccess;
}

}  // namespace

hipError_t launch_zero(void *ptr, size_t n_u64, hipStream_t s) {
  if (n_u64 == 0) return hipSuccess;
  unsigned blocks = (unsigned)((n_u64 + 255) / 256);
  k_zero_u64<<<blocks, 256, 0, s>>>((u64 *)ptr, n_u64);
  return hipGetLastError();
}

// Plain streaming copy of n 16-byte words: the reference point for measured
// HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters.  Each
// block copies one contiguous 1 MiB chunk, 8 words per lane in flight (all
// loads, then all stores).  On MI355X this reaches 5.4-5.6 TB/s over 1-4
// GiB, against 4.6-4.8 TB/s for a grid-stride loop of one word per lane and
// 5.2-5.4 TB/s for hipMemcpy device to device.
constexpr int kCopyThreads = 512, kCopyUnroll = 8;
constexpr size_t kCopyChunk = (size_t)1 << 16;  // words per block
__global__ __launch_bounds__(kCopyThreads) void k_copy_u4(int4 *__restrict__ dst, const int4 *__restrict__ src,
                                                          size_t n) {
  const size_t b0 = (size_t)blockIdx.x * kCopyChunk;
  const size_t b1 = b0 + kCopyChunk < n ? b0 + kCopyChunk : n;
  size_t i = b0 + threadIdx.x;
  for (; i + (kCopyUnroll - 1) * kCopyThreads < b1; i += kCopyUnroll * kCopyThreads) {
    int4 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) v[u] = src[i + u * kCopyThreads];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) dst[i + u * kCopyThreads] = v[u];
  }
  for (; i < b1; i += kCopyThreads) dst[i] = src[i];
}

hipError_t launch_copy(void *dst, const void *src, size_t n16, int num_cu, hipStream_t s) {
  (void)num_cu;
  if (n16 == 0) return hipSuccess;
  const unsigned grid = (unsigned)((n16 + kCopyChunk - 1) / kCopyChunk);
  k_copy_u4<<<grid, kCopyThreads, 0, s>>>((int4 *)dst, (const int4 *)src, n16);
  return hipGetLastError();
}

hipError_t launch_tc_merge(const DevCSR &g, const TaskList &tasks, Launch &l, int *queue) {
  if (!tasks.count) return hipSuccess;
  const int grid = persistent_grid(k_tc_merge, l.num_cu, tasks.count, kMergeWaves, 64 * kMergeWaves);
  l.begin("tc_merge", tasks.count);
  k_tc_merge<<<grid, 64 * kMergeWaves, 0, l.stream>>>(g.rowptr, g.colidx, tasks.tasks, (int)tasks.count, queue,
                                                      l.d_out, l.slots);
  l.end();
  return hipGetLastError();
}

hipError_t launch_tc(const DevCSR &g, const TaskList tiers[NUM_TIERS], int strategy, int hash_shift, Launch &l,
                     int *d_queues) {
  if (strategy == STRAT_BSEARCH) return launch_tc_strat<STRAT_BSEARCH, 2>(g, tiers, l, d_queues);
  if (hash_shift == 1) return launch_tc_strat<STRAT_HASH, 1>(g, tiers, l, d_queues);
  return launch_tc_strat<STRAT_HASH, 2>(g, tiers, l, d_queues);
}

}  // namespace xgm
