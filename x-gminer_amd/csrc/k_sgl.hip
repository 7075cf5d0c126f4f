// Edge-induced diamond and 4-cycle counting on the undirected graph, gfx950.
//
// Diamond -- reference: sum over edges v1 < v0 of C(|N(v0) ∩ N(v1)|, 2)
// (codes/GraphMiner/src/sgl/cpu_kernels/diamond.h:1-14; GPU
// codes/GraphMiner/src/sgl/gpu_kernels/diamond_count.cuh:1-20).  Here every
// undirected edge is owned by its endpoint t that is larger in (degree, id)
// order; N(t) is staged once (LDS hash, or HBM binary search for huge
// lists) and the lists of all partners p < t are probed against it, with a
// per-partner hit count reduced per window and turned into C(c, 2).
//
// 4-cycle -- reference: for v1 < v0 and v2 < v1 in N(v0), count
// |N(v1) ∩ N(v2) ∩ [0, v0)| (codes/GraphMiner/src/sgl/cpu_kernels/rectangle.h:1-11;
// GPU rectangle_nested.cuh:1-24).  Two strategies with identical counts:
//  * PAIR: that bounded-intersection loop itself, block per v0.
//  * WEDGE (default): every 4-cycle has one highest vertex x in any total
//    order and sum_x sum_{w < x} C(c_x(w), 2) counts each once, where
//    c_x(w) = |{v in N(x): v < x, w in N(v)}|.  On a copy of the graph
//    relabelled by (degree, id) rank, the lists N(v) ∩ [0, x) are prefixes,
//    and c_x(w) is aggregated in an LDS hash (or per-block HBM counters for
//    the few hub tasks); each increment from c to c+1 adds c to the count.
#include <climits>

#include "device_common.h"
#include "kernels.h"
#include "segments.h"

namespace xgm {
namespace {

// ------------------------------------------------------------------ diamond
template <int HCAP, int PCAP>
struct DiaStage {
  static_assert((HCAP & (HCAP - 1)) == 0, "power-of-two hash table");
  alignas(16) int key[2 * HCAP];  // buckets of 4 slots
  Segs<PCAP, false> seg;
  int cnt[PCAP];
  int ns, L, task;
};

// Items (t, k0, k1, -): owner t and the slice [k0, k1) of N(t) whose
// partners this item probes (hub owners are split over many items).
// MOTIF: instead of sum C(c, 2), the four per-edge sums of the reference's
// 4-motif formula (src/motif/cpu_kernels/automine_formula.h:27-36) with
// c = |N(t) ∩ N(p)|, su = d(t) - c - 1, sv = d(p) - c - 1, added into
// out[0..3] (out[4 * slot ..] with instances): sum su(su-1) + sv(sv-1),
// sum su*sv, sum c*(su + sv), sum c*(c-1).
template <int HCAP, int PCAP, int NT, bool LDS_SET, bool MOTIF>
__global__ __launch_bounds__(NT) void k_diamond(const int64_t *__restrict__ rowptr,
                                                const int32_t *__restrict__ colidx,
                                                const int4 *__restrict__ items, int ntasks,
                                                int *__restrict__ queue, u64 *__restrict__ out,
                                                const int32_t *__restrict__ slots) {
  __shared__ DiaStage<HCAP, PCAP> S;
  __shared__ int flagrow[NT / 64][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0, macc[4] = {0, 0, 0, 0};
  for (;;) {
    if (tid == 0) S.task = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.task;
    if (t >= ntasks) break;
    const int4 it = items[t];
    const int v = it.x;
    u64 part = 0, mpart[4] = {0, 0, 0, 0};
    const int64_t bt = rowptr[v];
    const int dt = (int)(rowptr[v + 1] - bt);
    const int32_t *nt = colidx + bt;
    int hb = 0;
    if (LDS_SET) {
      // the whole table (2 * HCAP slots, allocated anyway): the smaller the
      // owner, the lower the load and the shorter a missing probe's run
      hb = 31 - __clz(2 * HCAP) - 2;  // 2^hb buckets of 4 slots
      for (int q = tid; q < (4 << hb); q += NT) S.key[q] = -1;
      __syncthreads();
      for (int q = tid; q < dt; q += NT) bhash_insert(S.key, hb, nt[q]);
    }
    // partner p owned by v iff (deg p, p) < (deg v, v)
    auto keep = [dt, v](int p, int lp) { return lp < dt || (lp == dt && p < v); };
    for (int k0 = it.y; k0 < it.z; k0 += PCAP) {
      const int k1 = min(it.z, k0 + PCAP);
      __syncthreads();  // previous piece consumed; S.cnt is the build's scratch
      block_segments<NT>(S.seg, S.cnt, rowptr, colidx, nt, k0, k1, -1, keep, &S.ns, &S.L);
      const int ns = S.ns, L = S.L;
      for (int q = tid; q < ns; q += NT) S.cnt[q] = 0;
      __syncthreads();
      int lo, hi;
      wave_share(L, wave, NT / 64, lo, hi);
      wave_walk_group<4, true>(S.seg, colidx, ns, lo, hi, lane, fl, [&](const WindowLane (&r)[4]) {
        bool hit[4];
        if (LDS_SET) {
          // the four home buckets read together, then the rare full-bucket walks
          uint32_t h[4];
          int4 q[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            h[u] = hash_slot(r[u].w, hb);
            q[u] = reinterpret_cast<const int4 *>(S.key)[h[u]];
          }
          bool more[4];
          u64 any = 0;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            hit[u] = bucket_has(q[u], r[u].w);
            more[u] = r[u].valid && !hit[u] && q[u].w >= 0;
            any |= __ballot(more[u]);
          }
          // one wave-uniform branch per group for the rare full home buckets
          if (__builtin_expect(any != 0, 0)) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (more[u]) hit[u] = bhash_contains_after(S.key, hb, h[u], r[u].w);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) hit[u] = hit[u] && r[u].valid;
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) hit[u] = r[u].valid && sorted_contains(nt, dt, r[u].w);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const u64 hm = __ballot(hit[u]);
          if (r[u].heads == 1ull) {
            // the window lies in one list (wave-uniform): its hits are one
            // scalar popcount, added by one lane
            if (lane == 0 && hm) atomicAdd(&S.cnt[r[u].j], __popcll(hm));
          } else {
            const int c = segment_run_hits(hm, r[u], lane);
            if (c) atomicAdd(&S.cnt[r[u].j], c);
          }
        }
      });
      __syncthreads();
      for (int q = tid; q < ns; q += NT) {
        const u64 c = (u64)S.cnt[q];
        if (MOTIF) {
          const u64 dp = (u64)(S.seg.os[q + 1].y - S.seg.os[q].y);  // partner list, probed whole
          const u64 su = (u64)dt - c - 1, sv = dp - c - 1;
          mpart[0] += su * (su - 1) + sv * (sv - 1);
          mpart[1] += su * sv;
          mpart[2] += c * (su + sv);
          mpart[3] += c * (c - 1);
        } else {
          part += c * (c - 1) / 2;  // c = 0 gives 0 in modular arithmetic
        }
      }
    }
    if (MOTIF) {
      if (!slots) {
#pragma unroll
        for (int j = 0; j < 4; ++j) macc[j] += mpart[j];
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const u64 w = wave_sum_u64(mpart[j]);
          if (lane == 0 && w) atomicAdd(out + 4 * (size_t)slots[v] + j, w);
        }
      }
    } else {
      task_flush(acc, part, slots, v, out, lane);
    }
    __syncthreads();
  }
  if (MOTIF) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const u64 w = wave_sum_u64(macc[j]);
      if (lane == 0 && w) atomicAdd(out + j, w);
    }
  } else {
    acc = wave_sum_u64(acc);
    if (lane == 0 && acc) atomicAdd(out, acc);
  }
}

// Owners of degree > 8192 (hub slices).  N(t) is staged in LDS as a binned
// sorted set when ids fit 16 + kDiaBinBits bits and d <= kDiaBinCap: the id
// space is cut into 2^kDiaBinBits bins of 2^shift ids, first[b] is the
// position of the first element of bin b, and an element keeps only its low
// shift bits (uint16), so a lookup reads first[b], first[b + 1] and searches
// the few elements of its bin.  Otherwise N(t) is staged sorted (d <= LCAP)
// and binary-searched, or, above that, an LDS index of 1024 samples narrows
// each search in HBM to one d/1024-element stretch.
constexpr int kDiaBinBits = 14, kDiaBins = 1 << kDiaBinBits, kDiaBinCap = 56000;
template <int LCAP, int PCAP, int NT>
struct DiaBigStage {
  union {
    int list[LCAP];
    int sample[1024];
    struct {
      uint16_t first[kDiaBins + 2];
      uint16_t low[kDiaBinCap];
    } bin;
  };
  Segs<PCAP, false> seg;
  int cnt[PCAP];
  int ns, L, task;
};
static_assert(sizeof(((DiaBigStage<36352, 1000, 1024> *)nullptr)->bin) <= 36352 * sizeof(int),
              "the binned set shares the sorted list's LDS");

template <int LCAP, int PCAP, int NT, bool MOTIF>
__global__ __launch_bounds__(NT) void k_diamond_big(const int64_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx,
                                                    const int4 *__restrict__ items, int ntasks,
                                                    int *__restrict__ queue, u64 *__restrict__ out,
                                                    const int32_t *__restrict__ slots, int shift) {
  __shared__ DiaBigStage<LCAP, PCAP, NT> S;
  __shared__ int flagrow[NT / 64][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc[4] = {0, 0, 0, 0};
  for (;;) {
    if (tid == 0) S.task = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.task;
    if (t >= ntasks) break;
    const int4 it = items[t];
    const int v = it.x;
    u64 part[4] = {0, 0, 0, 0};
    const int64_t bt = rowptr[v];
    const int dt = (int)(rowptr[v + 1] - bt);
    const int32_t *nt = colidx + bt;
    const bool binned = shift <= 16 && dt <= kDiaBinCap;
    const bool in_lds = !binned && dt <= LCAP;
    const int lowmask = (1 << shift) - 1;
    if (binned) {
      // bins before the first element start at 0, bins after the last at dt
      const int b0 = dt ? nt[0] >> shift : kDiaBins;
      const int bl = dt ? nt[dt - 1] >> shift : -1;
      for (int b = tid; b <= b0; b += NT) S.bin.first[b] = 0;
      for (int b = bl + 1 + tid; b <= kDiaBins; b += NT) S.bin.first[b] = (uint16_t)dt;
      for (int q = tid; q < dt; q += NT) {
        const int w = nt[q];
        S.bin.low[q] = (uint16_t)(w & lowmask);
        if (q > 0)  // element q opens the bins (bin of q-1, bin of q]
          for (int b = (nt[q - 1] >> shift) + 1; b <= (w >> shift); ++b) S.bin.first[b] = (uint16_t)q;
      }
    } else if (in_lds) {
      for (int q = tid; q < dt; q += NT) S.list[q] = nt[q];
    } else {
      for (int q = tid; q < 1024; q += NT) S.sample[q] = nt[(int64_t)q * dt / 1024];
    }
    auto contains = [&](int w) {
      if (binned) {
        const int b = w >> shift;
        int lo = S.bin.first[b];
        const int hi = S.bin.first[b + 1];
        const int key = w & lowmask;
        int n = hi - lo;  // lower bound of key in low[lo, hi)
        while (n > 0) {
          const int h = n >> 1;
          if ((int)S.bin.low[lo + h] < key) {
            lo += h + 1;
            n -= h + 1;
          } else {
            n = h;
          }
        }
        return lo < hi && (int)S.bin.low[lo] == key;
      }
      if (in_lds) return sorted_contains(S.list, dt, w);
      // last sample <= w, then its stretch [j d / 1024, (j + 1) d / 1024) in HBM
      const int j = upper_bound_dev(S.sample, 1024, w) - 1;
      if (j < 0) return false;
      const int a = (int)((int64_t)j * dt / 1024);
      const int b = j == 1023 ? dt : (int)((int64_t)(j + 1) * dt / 1024);
      return sorted_contains(nt + a, b - a, w);
    };
    auto keep = [dt, v](int p, int lp) { return lp < dt || (lp == dt && p < v); };
    for (int k0 = it.y; k0 < it.z; k0 += PCAP) {
      const int k1 = min(it.z, k0 + PCAP);
      __syncthreads();  // staged set written / previous piece consumed; S.cnt is scratch
      block_segments<NT>(S.seg, S.cnt, rowptr, colidx, nt, k0, k1, -1, keep, &S.ns, &S.L);
      const int ns = S.ns, L = S.L;
      for (int q = tid; q < ns; q += NT) S.cnt[q] = 0;
      __syncthreads();
      int lo, hi;
      wave_share(L, wave, NT / 64, lo, hi);
      wave_walk<4, true>(S.seg, colidx, ns, lo, hi, lane, fl, [&](const WindowLane &r) {
        const bool hit = r.valid && contains(r.w);
        const u64 hm = __ballot(hit);
        if (r.heads == 1ull) {  // one list (wave-uniform): one lane adds the popcount
          if (lane == 0 && hm) atomicAdd(&S.cnt[r.j], __popcll(hm));
        } else {
          const int h = segment_run_hits(hm, r, lane);
          if (h) atomicAdd(&S.cnt[r.j], h);
        }
      });
      __syncthreads();
      for (int q = tid; q < ns; q += NT) {
        const u64 c = (u64)S.cnt[q];
        if (MOTIF) {
          const u64 dp = (u64)(S.seg.os[q + 1].y - S.seg.os[q].y);
          const u64 su = (u64)dt - c - 1, sv = dp - c - 1;
          part[0] += su * (su - 1) + sv * (sv - 1);
          part[1] += su * sv;
          part[2] += c * (su + sv);
          part[3] += c * (c - 1);
        } else {
          part[0] += c * (c - 1) / 2;
        }
      }
    }
    if (!slots) {
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += part[j];
    } else {
#pragma unroll
      for (int j = 0; j < (MOTIF ? 4 : 1); ++j) {
        const u64 w = wave_sum_u64(part[j]);
        if (lane == 0 && w) atomicAdd(out + (MOTIF ? 4 * (size_t)slots[v] + j : (size_t)slots[v]), w);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < (MOTIF ? 4 : 1); ++j) {
    const u64 w = wave_sum_u64(acc[j]);
    if (lane == 0 && w) atomicAdd(out + j, w);
  }
}

// Packed variant for owners of degree <= 64: a wave claims G items with one
// atomic; the staged lists of all G owners go into G hash regions through
// one flattened insert pass and the partners of all G owners form one
// segment table, so the dependent loads of G items are in flight together.
template <int G>
struct DiaPack {
  static constexpr int SLOTS = 128;  // 64 keys at load <= 1/2
  static constexpr int BBITS = 5;     // the whole region: 32 buckets of 4 slots
  alignas(16) int key[G * SLOTS];
  Segs<64 * G, true> seg;            // idx[] = item of the segment
  int cnt[64 * G];
  int u[G], k0[G], d[G], bits[G];
  int poff[G + 1], doff[G + 1];
  int64_t ob[G];
  u64 sum[G][4];                      // per-item sums (instance slots only)
};

template <int G, bool MOTIF>
__global__ __launch_bounds__(128) void k_diamond_packed(const int64_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ colidx,
                                                        const int4 *__restrict__ items, int ntasks,
                                                        int *__restrict__ queue, u64 *__restrict__ out,
                                                        const int32_t *__restrict__ slots) {
  using PS = DiaPack<G>;
  __shared__ PS stage[2];
  __shared__ int flagrow[2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  PS &s = stage[wave];
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc[4] = {0, 0, 0, 0};
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(queue, G);
    base = __shfl(base, 0, 64);
    if (base >= ntasks) break;
    const int ng = min(G, ntasks - base);
    int n = 0, dt = 0;
    if (lane < ng) {
      const int4 it = items[base + lane];
      const int64_t b = rowptr[it.x];
      dt = (int)(rowptr[it.x + 1] - b);
      n = it.z - it.y;
      s.u[lane] = it.x;
      s.k0[lane] = it.y;
      s.d[lane] = dt;
      s.ob[lane] = b;
#pragma unroll
      for (int j = 0; j < 4; ++j) s.sum[lane][j] = 0;
    }
    const int pin = wave_incl_scan(n, lane), din = wave_incl_scan(dt, lane);
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
    for (int q = lane; q < D; q += 64) {
      const int g = item_of(q, s.doff);
      bhash_insert(s.key + g * PS::SLOTS, PS::BBITS, colidx[s.ob[g] + (q - s.doff[g])]);
    }
    int ns = 0, L = 0;
    wave_segments_fn(s.seg, rowptr, 0, P,
                     [&](int q) {
                       const int g = item_of(q, s.poff);
                       return colidx[s.ob[g] + s.k0[g] + (q - s.poff[g])];
                     },
                     [&](int q, int p, int lp) {  // p owned by u_g: (deg p, p) < (deg u_g, u_g)
                       const int g = item_of(q, s.poff);
                       return lp < s.d[g] || (lp == s.d[g] && p < s.u[g]);
                     },
                     lane, ns, L);
    if (lane == 0) s.seg.os[ns] = make_int2(0, L);
    wave_sync();
    for (int j = lane; j < ns; j += 64) {
      s.seg.idx[j] = item_of(s.seg.idx[j], s.poff);
      s.cnt[j] = 0;
    }
    wave_sync();
    wave_walk<4, true>(s.seg, colidx, ns, 0, L, lane, fl, [&](const WindowLane &r) {
      bool hit = false;
      if (r.valid) {
        const int g = s.seg.idx[r.j];
        hit = bhash_contains(s.key + g * PS::SLOTS, PS::BBITS, r.w);
      }
      const int h = segment_run_hits(__ballot(hit), r, lane);
      if (h) atomicAdd(&s.cnt[r.j], h);
    });
    wave_sync();
    for (int q = lane; q < ns; q += 64) {
      const u64 c = (u64)s.cnt[q];
      const int g = s.seg.idx[q];
      u64 v[4] = {0, 0, 0, 0};
      if (MOTIF) {
        const u64 dp = (u64)(s.seg.os[q + 1].y - s.seg.os[q].y);
        const u64 su = (u64)s.d[g] - c - 1, sv = dp - c - 1;
        v[0] = su * (su - 1) + sv * (sv - 1);
        v[1] = su * sv;
        v[2] = c * (su + sv);
        v[3] = c * (c - 1);
      } else {
        v[0] = c * (c - 1) / 2;
      }
      if (!slots) {
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] += v[j];
      } else {
#pragma unroll
        for (int j = 0; j < (MOTIF ? 4 : 1); ++j)
          if (v[j]) atomicAdd(&s.sum[g][j], v[j]);
      }
    }
    if (slots) {
      wave_sync();
      if (lane < ng) {
        const int32_t sl = slots[s.u[lane]];
        if (MOTIF) {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (s.sum[lane][j]) atomicAdd(out + 4 * (size_t)sl + j, s.sum[lane][j]);
        } else if (s.sum[lane][0]) {
          atomicAdd(out + sl, s.sum[lane][0]);
        }
      }
    }
    wave_sync();  // stage is rewritten by the next claim
  }
#pragma unroll
  for (int j = 0; j < (MOTIF ? 4 : 1); ++j) {
    const u64 w = wave_sum_u64(acc[j]);
    if (lane == 0 && w) atomicAdd(out + j, w);
  }
}

// ------------------------------------------------------------ 4-cycle WEDGE
template <int HSLOTS, int PCAP>
struct WedgeStage {
  int key[HSLOTS];
  int val[HSLOTS];
  Segs<PCAP, false> seg;
  int tmp[PCAP];  // scratch of the block-cooperative segment build
  int ns, L, task;
};

// insert-or-increment; returns the count before this increment
__device__ __forceinline__ int hash_bump(int *key, int *val, int bits, int w) {
  const uint32_t mask = (1u << bits) - 1u;
  uint32_t h = hash_slot(w, bits);
  for (;;) {
    int k = key[h];
    if (k == -1) {
      int prev = atomicCAS(&key[h], -1, w);
      if (prev == -1 || prev == w) return atomicAdd(&val[h], 1);
      k = prev;
    }
    if (k == w) return atomicAdd(&val[h], 1);
    h = (h + 1) & mask;
  }
}

// Block per task x of the relabelled graph.  DENSE: counters in the row
// dense + blockIdx.x * stride (HBM), indexed by w - lo where lo is the first
// id of x's instance (inst_lo[slots[x]]; 0 without instances), reset after
// the task; else LDS hash sized for wedge work <= HSLOTS / 2.
template <int HSLOTS, int PCAP, int NT, bool DENSE>
__global__ __launch_bounds__(NT) void k_cycle4_wedge(const int64_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colidx,
                                                     const int32_t *__restrict__ cut,
                                                     const int32_t *__restrict__ nlowv,
                                                     const int32_t *__restrict__ tasks, int ntasks,
                                                     int *__restrict__ dense, int64_t stride,
                                                     const int64_t *__restrict__ inst_lo,
                                                     int *__restrict__ queue, u64 *__restrict__ out,
                                                     const int32_t *__restrict__ slots) {
  __shared__ WedgeStage<DENSE ? 1 : HSLOTS, PCAP> S;
  __shared__ int flagrow[NT / 64][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  int *const row = DENSE ? dense + (size_t)blockIdx.x * (size_t)stride : nullptr;
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.task = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.task;
    if (t >= ntasks) break;
    const int x = tasks[t];
    u64 part = 0;
    const int64_t bx = rowptr[x];
    const int32_t *nx = colidx + bx;
    const int nlow = nlowv[x];  // |N(x) ∩ [0, x)|
    int *cnt = DENSE ? row - (inst_lo ? inst_lo[slots[x]] : 0) : nullptr;
    int hb = 0;
    if (!DENSE) {
      hb = 1;
      while ((1 << hb) < HSLOTS) ++hb;
      for (int q = tid; q < HSLOTS; q += NT) { S.key[q] = -1; S.val[q] = 0; }
    }
    for (int k0 = 0; k0 < nlow; k0 += PCAP) {
      const int k1 = min(nlow, k0 + PCAP);
      __syncthreads();
      block_segments_cut<NT>(S.seg, S.tmp, rowptr, nx, cut + bx, k0, k1, &S.ns, &S.L);
      const int ns = S.ns, L = S.L;
      int lo, hi;
      wave_share(L, wave, NT / 64, lo, hi);
      wave_walk<4, DENSE>(S.seg, colidx, ns, lo, hi, lane, fl, [&](const WindowLane &r) {
        if (r.valid) {
          const int old = DENSE ? atomicAdd(&cnt[r.w], 1) : hash_bump(S.key, S.val, hb, r.w);
          part += (u64)old;
        }
      });
      (void)ns;
    }
    if (DENSE) {
      // reset the counters this task touched (its atomics have all returned)
      for (int k0 = 0; k0 < nlow; k0 += PCAP) {
        const int k1 = min(nlow, k0 + PCAP);
        __syncthreads();
        block_segments_cut<NT>(S.seg, S.tmp, rowptr, nx, cut + bx, k0, k1, &S.ns, &S.L);
        const int ns = S.ns, L = S.L;
        int lo, hi;
        wave_share(L, wave, NT / 64, lo, hi);
        wave_walk<4>(S.seg, colidx, ns, lo, hi, lane, fl, [&](const WindowLane &r) {
          if (r.valid) cnt[r.w] = 0;
        });
      }
      __threadfence();  // the zeros reach L2 before the next task's atomics
    }
    task_flush(acc, part, slots, x, out, lane);
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Packed wedge tasks (<= 128 wedges, <= 64 lower neighbours): a wave claims G
// tasks with one atomic; their lower lists N(v) ∩ [0, x_g) form one segment
// table (each cut at its own task's x) and c_x(w) goes into G LDS hash
// regions of 256 slots, so the dependent loads of G tasks overlap.
template <int G>
struct WedgePack {
  static constexpr int SLOTS = 256;  // <= 128 wedges at load <= 1/2
  int key[G * SLOTS];
  int val[G * SLOTS];
  Segs<64 * G, true> seg;            // idx[] = task of the segment
  int x[G], nl[G], off[G + 1];
  int64_t bx[G];
  u64 sum[G];
};

template <int G>
__global__ __launch_bounds__(128) void k_cycle4_wedge_packed(const int64_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ colidx,
                                                             const int32_t *__restrict__ cut,
                                                             const int32_t *__restrict__ nlowv,
                                                             const int32_t *__restrict__ tasks, int ntasks,
                                                             int *__restrict__ queue, u64 *__restrict__ out,
                                                             const int32_t *__restrict__ slots) {
  using PS = WedgePack<G>;
  __shared__ PS stage[2];
  __shared__ int flagrow[2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  PS &S = stage[wave];
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(queue, G);
    base = __shfl(base, 0, 64);
    if (base >= ntasks) break;
    const int ng = min(G, ntasks - base);
    int nl = 0;
    if (lane < ng) {
      const int x = tasks[base + lane];
      const int64_t b = rowptr[x];
      nl = nlowv[x];  // |N(x) ∩ [0, x)| <= 64
      S.x[lane] = x;
      S.bx[lane] = b;
      S.nl[lane] = nl;
      S.sum[lane] = 0;
    }
    const int inc = wave_incl_scan(nl, lane);
    if (lane < ng) S.off[lane + 1] = inc;
    if (lane == 0) S.off[0] = 0;
    for (int q = lane; q < ng * PS::SLOTS; q += 64) {
      S.key[q] = -1;
      S.val[q] = 0;
    }
    wave_sync();
    const int P = __shfl(inc, ng - 1, 64);
    auto item_of = [&](int q) {
      int g = 0;
#pragma unroll
      for (int h = 1; h < G; ++h) g += (h < ng && q >= S.off[h]) ? 1 : 0;
      return g;
    };
    // one segment per lower neighbour v of x_g: N(v) ∩ [0, x_g)
    int ns = 0, L = 0;
    for (int k = 0; k < P; k += 64) {
      const int q = k + lane;
      int len = 0, g = 0;
      int64_t vb = 0;
      if (q < P) {
        g = item_of(q);
        const int64_t e = S.bx[g] + (q - S.off[g]);
        vb = rowptr[colidx[e]];
        len = cut[e];  // |N(v) ∩ [0, x_g)|
      }
      const int incl = wave_incl_scan(len, lane) + L;
      const u64 m = __ballot(len > 0);
      if (len > 0) {
        const int pos = ns + __popcll(m & lanemask_lt(lane));
        S.seg.os[pos] = make_int2((int)vb, incl - len);
        S.seg.idx[pos] = g;
      }
      ns += __popcll(m);
      L = __shfl(incl, 63, 64);
    }
    if (lane == 0) S.seg.os[ns] = make_int2(0, L);
    wave_sync();
    u64 part = 0;
    wave_walk<4, true>(S.seg, colidx, ns, 0, L, lane, fl, [&](const WindowLane &r) {
      if (r.valid) {
        const int g = S.seg.idx[r.j];
        const int old = hash_bump(S.key + g * PS::SLOTS, S.val + g * PS::SLOTS, 8, r.w);
        if (old) {
          if (slots) atomicAdd(&S.sum[g], (u64)old);
          else part += (u64)old;
        }
      }
    });
    if (!slots) {
      acc += part;
    } else {
      wave_sync();
      if (lane < ng && S.sum[lane]) atomicAdd(out + slots[S.x[lane]], S.sum[lane]);
    }
    wave_sync();  // stage is rewritten by the next claim
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Range items of the wedge strategy: (x, r0, r1, half) counts the wedges of
// x whose far end w lies in [r0, r1) (r1 <= x).  [r0, r1) is one or more
// consecutive spans of `span` ids, taken one after the other, each with
// c_x(w) in LDS counters indexed by w - s0 (s0 = the span's first id):
// span = R with R 32-bit counters, or with half = 1, span = 2R with 16-bit
// counters packed two per word, which the host picks only when x has fewer
// than 65536 lower neighbours: c_x(w) never exceeds that number, so a
// half-word never carries into its neighbour.  The spans of one task
// partition the ids below x, so every w is counted whole in exactly one
// span and the previous values returned by the increments still sum to
// C(c_x(w), 2); no HBM counters or global atomics are involved.
//
// Cutting every lower neighbour's list to a span (scattered probes into the
// lists) dominates this kernel's time.  The first span of an item finds its
// start in each list by binary search; every span's end is found by a
// galloping search from its start and kept in the block's cursor row
// (cursor + blockIdx.x * cursor_stride; the host sizes cursor_stride to the
// largest lower-neighbour count) as the next span's start, next to the
// list's end, both as colidx positions, and the id stored there: later
// spans read one cursor instead of the neighbour id and its rowptr entries,
// and a list whose next id lies beyond the span is skipped without a
// probe.  Each cursor is written and read back by the same thread.
// Positions are int32, as in the segment tables.
template <int R, int PCAP, int NT>
struct RangeStage {
  uint32_t cnt[R];
  Segs<PCAP, false> seg;
  int wcnt[NT / 64], wlen[NT / 64];  // per-wave totals of the segment compaction
  int flagrow[NT / 64][64];
  int ns, L, task;
};

// First index i in [lo, n) with a[i] >= key, where every a[j], j < lo, is
// below key: probes lo, lo + 1, lo + 3, lo + 7, ... until one reaches key,
// then binary-searches the last gap (about 2 log2 of the distance moved).
__device__ __forceinline__ int gallop_lower_bound(const int32_t *__restrict__ a, int lo, int n, int key) {
  int hi = lo, step = 1;
  while (hi < n && a[hi] < key) {
    lo = hi + 1;
    hi = (n - hi > step) ? hi + step : n;
    step <<= 1;
  }
  return lo + lower_bound_dev(a + lo, hi - lo, key);
}

template <int R, int PCAP, int NT>
__global__ __launch_bounds__(NT) void k_cycle4_range(const int64_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colidx,
                                                     const int32_t *__restrict__ nlowv,
                                                     const int4 *__restrict__ items, int nitems,
                                                     int4 *__restrict__ cursor, int64_t cursor_stride,
                                                     int *__restrict__ queue, u64 *__restrict__ out,
                                                     const int32_t *__restrict__ slots) {
  static_assert(PCAP % NT == 0, "whole lists per thread");
  __shared__ RangeStage<R, PCAP, NT> S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(S.flagrow[wave], lane);
  int4 *const cur = cursor + (size_t)blockIdx.x * (size_t)cursor_stride;
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.task = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.task;
    if (t >= nitems) break;
    const int4 it = items[t];
    const int x = it.x, r0 = it.y, r1 = it.z;
    const bool half = it.w != 0;  // 16-bit counters: w - s0 in word (w - s0) / 2
    const int span = half ? 2 * R : R;
    const int64_t bx = rowptr[x];
    const int32_t *nx = colidx + bx;
    const int nlow = nlowv[x];
    u64 part = 0;
    for (int s0 = r0, s1 = r0; s0 < r1; s0 = s1) {  // s1 never passes r1: no int overflow near 2^31
      s1 = r1 - s0 > span ? s0 + span : r1;
      const bool first = s0 == r0, last = s1 == r1;
      const int words = half ? (s1 - s0 + 1) >> 1 : s1 - s0;
      __syncthreads();  // the previous span's walk is done with the counters
      for (int q = tid; q < words; q += NT) S.cnt[q] = 0u;
      for (int k0 = 0; k0 < nlow; k0 += PCAP) {
        const int n = min(nlow - k0, PCAP);
        __syncthreads();  // counters zeroed / previous piece walked
        // each thread cuts lists tid + h * NT, h < PCAP / NT; all of its
        // cursors (or list bounds) are loaded before the first is used, so
        // their latencies overlap
        // (span start, list end) as colidx positions, and the id at the
        // span start (INT32_MAX past the end; unknown = -1 on a first span)
        int4 c[PCAP / NT];
#pragma unroll
        for (int h = 0; h < PCAP / NT; ++h) {
          const int q = tid + h * NT;
          c[h] = make_int4(0, 0, INT32_MAX, 0);
          if (q < n) {
            if (first) {
              const int v = nx[k0 + q];
              const int64_t vb = rowptr[v];
              c[h] = make_int4((int)vb, (int)rowptr[v + 1], -1, 0);
            } else {
              c[h] = cur[k0 + q];
            }
          }
        }
#pragma unroll
        for (int h = 0; h < PCAP / NT; ++h) {
          const int q = tid + h * NT;
          if (q >= n) continue;
          if (first) {
            const int32_t *nv = colidx + c[h].x;
            const int dv = c[h].y - c[h].x;
            // an item that starts at the task's first id starts every list at 0
            c[h].x += (dv == 0 || nv[0] >= s0) ? 0 : lower_bound_dev(nv, dv, s0);
          }
          int len = 0;
          if (c[h].z < s1) {  // else the list has nothing in this span: no probe, cursor unchanged
            // a known id at the span start is below s1: its probe is skipped
            len = c[h].z >= 0 ? 1 + gallop_lower_bound(colidx + c[h].x + 1, 0, c[h].y - c[h].x - 1, s1)
                              : gallop_lower_bound(colidx + c[h].x, 0, c[h].y - c[h].x, s1);
            const int p = c[h].x + len;
            if (!last) cur[k0 + q] = make_int4(p, c[h].y, p < c[h].y ? colidx[p] : INT32_MAX, 0);
          }
          c[h].w = len;  // the piece's length in this span (0: none)
        }
        // compact the non-empty pieces with a block-wide scan of (pieces,
        // elements): every thread writes its own at the scanned positions
        int pc = 0, pl = 0;
#pragma unroll
        for (int h = 0; h < PCAP / NT; ++h) {
          const bool here = tid + h * NT < n && c[h].w > 0;
          pc += here ? 1 : 0;
          pl += here ? c[h].w : 0;
        }
        const int ic = wave_incl_scan(pc, lane), il = wave_incl_scan(pl, lane);
        if (lane == 63) {
          S.wcnt[wave] = ic;
          S.wlen[wave] = il;
        }
        __syncthreads();
        int pos = ic - pc, start = il - pl;
        for (int w = 0; w < wave; ++w) {
          pos += S.wcnt[w];
          start += S.wlen[w];
        }
#pragma unroll
        for (int h = 0; h < PCAP / NT; ++h) {
          if (tid + h * NT < n && c[h].w > 0) {
            S.seg.os[pos++] = make_int2(c[h].x, start);
            start += c[h].w;
          }
        }
        if (tid == NT - 1) {  // pos and start are the totals here
          S.seg.os[pos] = make_int2(0, start);
          S.ns = pos;
          S.L = start;
        }
        __syncthreads();
        int lo, hi;
        wave_share(S.L, wave, NT / 64, lo, hi);
        wave_walk<4>(S.seg, colidx, S.ns, lo, hi, lane, fl, [&](const WindowLane &r) {
          if (r.valid) {
            const int i = r.w - s0;
            if (half) {
              const int sh = (i & 1) << 4;
              part += (u64)((atomicAdd(&S.cnt[i >> 1], 1u << sh) >> sh) & 0xFFFFu);
            } else {
              part += (u64)atomicAdd(&S.cnt[i], 1u);
            }
          }
        });
      }
    }
    task_flush(acc, part, slots, x, out, lane);
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Hub tasks of the wedge strategy, split into items (x, k0, k1, slot): the
// slice [k0, k1) of x's lower neighbours.  All items of one hub add into
// the same HBM counter row hub + slot * stride (indexed like the dense rows)
// with device-scope atomics, so the returned previous values still sum to
// sum_w C(c_x(w), 2).  The counter region is zeroed before the launch.
template <int PCAP, int NT>
__global__ __launch_bounds__(NT) void k_cycle4_hub(const int64_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ colidx,
                                                   const int32_t *__restrict__ cut,
                                                   const int4 *__restrict__ items, int nitems,
                                                   int *__restrict__ hub, int64_t stride,
                                                   const int64_t *__restrict__ inst_lo,
                                                   int *__restrict__ queue, u64 *__restrict__ out,
                                                   const int32_t *__restrict__ slots) {
  __shared__ Segs<PCAP, false> seg;
  __shared__ int flagrow[NT / 64][64];
  __shared__ int sh_item, sh_ns, sh_L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    if (tid == 0) sh_item = atomicAdd(queue, 1);
    __syncthreads();
    const int t = sh_item;
    if (t >= nitems) break;
    const int4 it = items[t];
    const int x = it.x;
    u64 part = 0;
    const int64_t bx = rowptr[x];
    const int32_t *nx = colidx + bx;
    int *cnt = hub + (size_t)it.w * (size_t)stride - (inst_lo ? inst_lo[slots[x]] : 0);
    for (int k0 = it.y; k0 < it.z; k0 += PCAP) {
      const int k1 = min(it.z, k0 + PCAP);
      __syncthreads();
      if (wave == 0) {
        int ns = 0, L = 0;
        wave_segments_range_cut(seg, rowptr, nx, cut + bx, k0, k1, lane, ns, L);
        if (lane == 0) {
          seg.os[ns] = make_int2(0, L);
          sh_ns = ns;
          sh_L = L;
        }
      }
      __syncthreads();
      int lo, hi;
      wave_share(sh_L, wave, NT / 64, lo, hi);
      wave_walk<4>(seg, colidx, sh_ns, lo, hi, lane, fl, [&](const WindowLane &r) {
        if (r.valid) part += (u64)atomicAdd(&cnt[r.w], 1);
      });
    }
    task_flush(acc, part, slots, x, out, lane);
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// ------------------------------------------------------------- 4-cycle PAIR
constexpr int PAIR_HCAP = 2048;
struct PairStage {
  int key[2 * PAIR_HCAP];
  Segs<1024, false> seg;
  int ns, L, task;
};

__global__ __launch_bounds__(256) void k_cycle4_pair(const int64_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colidx,
                                                     const int32_t *__restrict__ tasks, int ntasks,
                                                     int *__restrict__ queue, u64 *__restrict__ out,
                                                     const int32_t *__restrict__ slots) {
  __shared__ PairStage S;
  __shared__ int flagrow[4][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.task = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.task;
    if (t >= ntasks) break;
    const int v0 = tasks[t];
    u64 part = 0;
    const int64_t b0 = rowptr[v0];
    const int32_t *n0 = colidx + b0;
    const int nlow = lower_bound_dev(n0, (int)(rowptr[v0 + 1] - b0), v0);  // v1, v2 < v0
    for (int a = 1; a < nlow; ++a) {
      const int v1 = n0[a];
      const int64_t b1 = rowptr[v1];
      const int32_t *n1 = colidx + b1;
      const int s1 = lower_bound_dev(n1, (int)(rowptr[v1 + 1] - b1), v0);  // N(v1) ∩ [0, v0)
      if (s1 == 0) continue;
      const bool lds = s1 <= PAIR_HCAP;
      int hb = 0;
      __syncthreads();
      if (lds) {
        hb = hash_bits_for(s1);
        for (int q = tid; q < (1 << hb); q += 256) S.key[q] = -1;
        __syncthreads();
        for (int q = tid; q < s1; q += 256) hash_insert(S.key, hb, n1[q]);
      }
      // v2 = n0[0..a) (all < v1), lists cut at v0
      for (int k0 = 0; k0 < a; k0 += 1024) {
        const int k1 = min(a, k0 + 1024);
        __syncthreads();
        if (wave == 0) {
          int ns, L;
          wave_segments(S.seg, rowptr, colidx, n0, k0, k1, v0, KeepAll(), lane, ns, L);
          if (lane == 0) { S.ns = ns; S.L = L; }
        }
        __syncthreads();
        const int ns = S.ns, L = S.L;
        int lo, hi;
        wave_share(L, wave, 4, lo, hi);
        wave_walk<4>(S.seg, colidx, ns, lo, hi, lane, fl, [&](const WindowLane &r) {
          if (r.valid) part += lds ? hash_contains(S.key, hb, r.w) : sorted_contains(n1, s1, r.w);
        });
      }
    }
    task_flush(acc, part, slots, v0, out, lane);
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

template <class K>
int grid_of(K kernel, int nt, int num_cu, int64_t tasks, int cap_per_cu = 1 << 20) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  per_cu = std::min(per_cu, cap_per_cu);
  return (int)std::max<int64_t>(1, std::min<int64_t>(tasks, (int64_t)per_cu * num_cu));
}

template <int HCAP, int PCAP, int NT, bool LDS_SET, bool MOTIF>
hipError_t run_diamond(const char *name, const DevCSR &g, const TaskList &tl, int *queue, Launch &l) {
  if (!tl.count) return hipSuccess;
  auto k = k_diamond<HCAP, PCAP, NT, LDS_SET, MOTIF>;
  int grid = grid_of(k, NT, l.num_cu, tl.count);
  l.begin(name, tl.count);
  k<<<grid, NT, 0, l.stream>>>(g.rowptr, g.colidx, (const int4 *)tl.tasks, (int)tl.count, queue, l.d_out, l.slots);
  l.end();
  return hipGetLastError();
}

template <int HSLOTS, int PCAP, int NT>
hipError_t run_wedge_lds(const char *name, const DevCSR &g, const TaskList &tl, int *queue, Launch &l) {
  if (!tl.count) return hipSuccess;
  auto k = k_cycle4_wedge<HSLOTS, PCAP, NT, false>;
  int grid = grid_of(k, NT, l.num_cu, tl.count);
  l.begin(name, tl.count);
  k<<<grid, NT, 0, l.stream>>>(g.rowptr, g.colidx, g.cut, g.nlow, tl.tasks, (int)tl.count, nullptr, 0, nullptr,
                               queue, l.d_out, l.slots);
  l.end();
  return hipGetLastError();
}

}  // namespace

template <bool MOTIF>
hipError_t launch_diamond_mode(const DevCSR &g, const TaskList tl[4], Launch &l, int *d_queues) {
  hipError_t e;
  if (tl[3].count) {  // owners above 8192: binned set in LDS, sorted set (36,352 ints), or a sample index
    auto k = k_diamond_big<36352, 1000, 1024, MOTIF>;
    const int grid = grid_of(k, 1024, l.num_cu, tl[3].count);
    l.begin(MOTIF ? "motif_edges_hbm_set" : "diamond_hbm_set", tl[3].count);
    // bins of 2^shift ids: ids < 2^idbits in 2^kDiaBinBits bins
    const int idbits = g.n > 1 ? 64 - __builtin_clzll((unsigned long long)(g.n - 1)) : 1;
    const int shift = std::max(0, idbits - kDiaBinBits);
    k<<<grid, 1024, 0, l.stream>>>(g.rowptr, g.colidx, (const int4 *)tl[3].tasks, (int)tl[3].count, d_queues + 3,
                                   l.d_out, l.slots, shift);
    l.end();
    if ((e = hipGetLastError())) return e;
  }
  // 64 KB of LDS hash (+ a table of 1000 partners) per block leaves room for
  // two blocks per CU; 1024 threads each keep ~28 waves resident
  if ((e = run_diamond<8192, 1000, 1024, true, MOTIF>(MOTIF ? "motif_edges_block_8192" : "diamond_block_8192", g,
                                                      tl[2], d_queues + 2, l)))
    return e;
  if ((e = run_diamond<1024, 1024, 256, true, MOTIF>(MOTIF ? "motif_edges_block_1024" : "diamond_block_1024", g,
                                                     tl[1], d_queues + 1, l)))
    return e;
  if (tl[0].count) {  // owners of degree <= 64: 8 per wave, two waves per block
    constexpr int kPack = 8;
    auto k = k_diamond_packed<kPack, MOTIF>;
    const int grid = grid_of(k, 128, l.num_cu, (tl[0].count + 2 * kPack - 1) / (2 * kPack));
    l.begin(MOTIF ? "motif_edges_wave_64" : "diamond_wave_64", tl[0].count);
    k<<<grid, 128, 0, l.stream>>>(g.rowptr, g.colidx, (const int4 *)tl[0].tasks, (int)tl[0].count, d_queues + 0,
                                  l.d_out, l.slots);
    l.end();
    if ((e = hipGetLastError())) return e;
  }
  return hipSuccess;
}

hipError_t launch_diamond(const DevCSR &g, const TaskList tl[4] /* int4 items */, Launch &l, int *d_queues) {
  return launch_diamond_mode<false>(g, tl, l, d_queues);
}

hipError_t launch_motif_edges(const DevCSR &g, const TaskList tl[4], Launch &l, int *d_queues) {
  return launch_diamond_mode<true>(g, tl, l, d_queues);
}

int cycle4_dense_grid(int num_cu, int64_t tasks) {
  return grid_of(k_cycle4_wedge<1, 1024, 256, true>, 256, num_cu, tasks, 1);
}

int cycle4_range_grid(int num_cu, int64_t items) {
  return grid_of(k_cycle4_range<kRangeCounters, 2048, 1024>, 1024, num_cu, items);
}

hipError_t launch_cycle4_wedge(const DevCSR &g, const TaskList tl[4], int *dense, const TaskList *hub_rounds,
                               int n_rounds, int *hub, size_t hub_round_u64, const TaskList &range_items,
                               const CounterRows &rows, Launch &l, int *d_queues) {
  if (!g.cut || !g.nlow) return hipErrorInvalidValue;  // the relabelled graph's cut positions
  hipError_t e;
  if (range_items.count) {
    // 2048-neighbour pieces: 128 KB of counters + 24 KB of segment tables
    // still fit one CU's LDS, with half the barriers of 1024-pieces
    auto k = k_cycle4_range<kRangeCounters, 2048, 1024>;
    const int grid = cycle4_range_grid(l.num_cu, range_items.count);
    l.begin("cycle4_wedge_range", range_items.count);
    k<<<grid, 1024, 0, l.stream>>>(g.rowptr, g.colidx, g.nlow, (const int4 *)range_items.tasks, (int)range_items.count,
                                   rows.cursor, rows.cursor_stride, d_queues + 6, l.d_out, l.slots);
    l.end();
    if ((e = hipGetLastError())) return e;
  }
  int64_t hub_items = 0;
  for (int r = 0; r < n_rounds; ++r) hub_items += hub_rounds[r].count;
  if (hub_items) {
    auto k = k_cycle4_hub<1024, 256>;
    l.begin("cycle4_wedge_hub", hub_items);
    for (int r = 0; r < n_rounds; ++r) {
      const TaskList &it = hub_rounds[r];
      if (!it.count) continue;
      if ((e = launch_zero(hub, hub_round_u64, l.stream))) return e;
      if ((e = launch_zero(d_queues + 4, 1, l.stream))) return e;
      int grid = grid_of(k, 256, l.num_cu, it.count);
      k<<<grid, 256, 0, l.stream>>>(g.rowptr, g.colidx, g.cut, (const int4 *)it.tasks, (int)it.count, hub, rows.stride,
                                    rows.inst_lo, d_queues + 4, l.d_out, l.slots);
      if ((e = hipGetLastError())) return e;
    }
    l.end();
  }
  if (tl[2].count) {
    int grid = cycle4_dense_grid(l.num_cu, tl[2].count);
    l.begin("cycle4_wedge_hbm", tl[2].count);
    k_cycle4_wedge<1, 1024, 256, true><<<grid, 256, 0, l.stream>>>(g.rowptr, g.colidx, g.cut, g.nlow, tl[2].tasks,
                                                                    (int)tl[2].count, dense, rows.stride, rows.inst_lo,
                                                                    d_queues + 2, l.d_out, l.slots);
    l.end();
    if ((e = hipGetLastError())) return e;
  }
  // 1000-partner pieces keep the 8192-slot stage (with its build scratch)
  // at two blocks per CU
  if ((e = run_wedge_lds<8192, 1000, 1024>("cycle4_wedge_8192", g, tl[1], d_queues + 1, l))) return e;
  if ((e = run_wedge_lds<1024, 256, 64>("cycle4_wedge_1024", g, tl[0], d_queues + 0, l))) return e;
  if (tl[3].count) {  // <= 128 wedges and <= 64 lower neighbours: 8 tasks per wave
    constexpr int kPack = 8;
    auto k = k_cycle4_wedge_packed<kPack>;
    const int grid = grid_of(k, 128, l.num_cu, (tl[3].count + 2 * kPack - 1) / (2 * kPack));
    l.begin("cycle4_wedge_packed", tl[3].count);
    k<<<grid, 128, 0, l.stream>>>(g.rowptr, g.colidx, g.cut, g.nlow, tl[3].tasks, (int)tl[3].count, d_queues + 7,
                                  l.d_out, l.slots);
    l.end();
    if ((e = hipGetLastError())) return e;
  }
  return hipSuccess;
}

hipError_t launch_cycle4_pair(const DevCSR &g, const TaskList &tl, Launch &l, int *d_queues) {
  if (!tl.count) return hipSuccess;
  int grid = grid_of(k_cycle4_pair, 256, l.num_cu, tl.count);
  l.begin("cycle4_pair", tl.count);
  k_cycle4_pair<<<grid, 256, 0, l.stream>>>(g.rowptr, g.colidx, tl.tasks, (int)tl.count, d_queues, l.d_out, l.slots);
  l.end();
  return hipGetLastError();
}

}  // namespace xgm
