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
#include "merge_path.h"
#include "segments.h"

namespace xgm {
namespace {

// ------------------------------------------------------------- LUT engine
// LDS row layout: row r starts at r * W + r / 8 (W >= 16), so eight rows
// share one padding word.  The counting loops have every lane of a wave
// read the same word index w of a different row; with rows of 8W bytes
// back to back those reads all land in one LDS bank.  The per-row shift
// spreads them over the banks and keeps w a constant offset from the row
// base.  lut_words(s) is the footprint of s rows.
template <int W>
__device__ __forceinline__ int lut_word(int r, int w) {
  return r * W + (W >= 16 ? (r >> 3) : 0) + w;
}
template <int W>
__host__ __device__ constexpr int lut_words(int s) {
  return s * W + (W >= 16 ? (s + 7) >> 3 : 0);
}

template <int W>
__device__ __forceinline__ u64 lut_count(const u64 *rows, int s, int weff, int depth, int tid, int nthreads) {
  u64 acc = 0;
  const int items = s * weff;
  for (int t = tid; t < items; t += nthreads) {
    const int i = t / weff, q = t - i * weff;
    u64 bits = rows[lut_word<W>(i, q)];
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
          if (w < weff) acc += __popcll(rows[lut_word<W>(i, w)] & rows[lut_word<W>(j, w)]);
      } else {
        u64 tw[W];
#pragma unroll
        for (int w = 0; w < W; ++w) tw[w] = (w < weff) ? (rows[lut_word<W>(i, w)] & rows[lut_word<W>(j, w)]) : 0ull;
#pragma unroll
        for (int w2 = 0; w2 < W; ++w2) {
          u64 tb = tw[w2];
          while (tb) {
            const int k = w2 * 64 + __builtin_ctzll(tb);
            tb &= tb - 1;
#pragma unroll
            for (int w = 0; w < W; ++w)
              if (w < weff) acc += __popcll(tw[w] & rows[lut_word<W>(k, w)]);
          }
        }
      }
    }
  }
  return acc;
}

// Row access for the wave-cooperative counters: LDS rows in the padded
// lut_word layout, or HBM rows at a runtime stride of weff words.
template <int W>
struct LdsRows {
  const u64 *rows;
  __device__ __forceinline__ u64 operator()(int r, int w) const { return rows[lut_word<W>(r, w)]; }
};
struct HbmRows {
  const u64 *rows;
  int stride;
  __device__ __forceinline__ u64 operator()(int r, int w) const { return rows[(size_t)r * stride + w]; }
};

// Wave-cooperative LUT count for depth 3 and 4 (W = the register bound on
// the words per row, weff <= W): waves claim rows a through an LDS counter
// (lowest first: in the triangular layout those rows hold the most bits);
// the set bits j of row a are compacted into a per-wave queue of 128 ids
// (ballot rank) and handed out one per lane, so each lane owns one pair
// (a, j) instead of a 64-bit word of row a with 0..64 pairs in it, and the
// row-a words are the same address across the wave.  TRI: rows hold only
// later ranks (rank-ordered LUT), so words below (j+1)/64 are skipped.
// jq: 128 ints of LDS for this wave; rowq: the block's row counter, zero
// on entry.  Only rows a0, a0 + astep, a0 + 2 astep, ... are counted (one
// part of a task split over several blocks; a0 = 0, astep = 1 counts all).
// All lanes of every wave call.
template <int W, bool TRI, class Rows>
__device__ __forceinline__ u64 lut_count_wq(Rows row, int s, int weff, int depth, int *jq, int *rowq, int lane,
                                            int a0 = 0, int astep = 1) {
  u64 acc = 0;
  for (;;) {
    int a = 0;
    if (lane == 0) a = atomicAdd(rowq, 1);
    a = a0 + astep * __shfl(a, 0, 64);
    if (a >= s) break;
    int pending = 0;
    for (int q = TRI ? (a >> 6) : 0; q <= weff; ++q) {
      if (q < weff) {
        const u64 bits = row(a, q);
        if ((bits >> lane) & 1) jq[pending + __popcll(bits & lanemask_lt(lane))] = q * 64 + lane;
        pending += __popcll(bits);
        wave_sync();
      }
      while (pending >= 64 || (q == weff && pending > 0)) {
        const int take = min(pending, 64);
        if (lane < take) {
          const int j = jq[lane];
          const int ws = TRI ? (j + 1) >> 6 : 0;
          // a pair's sum stays below 2^22 (at most 2048 k times 2048 bits):
          // 32-bit accumulation lets each popcount add into its counter
          uint32_t c = 0;
          if (depth == 3) {
#pragma unroll
            for (int w = 0; w < W; ++w)
              if (w >= ws && w < weff) c += (uint32_t)__popcll(row(a, w) & row(j, w));
          } else {
            u64 tw[W];
#pragma unroll
            for (int w = 0; w < W; ++w) tw[w] = (w >= ws && w < weff) ? (row(a, w) & row(j, w)) : 0ull;
#pragma unroll
            for (int w2 = 0; w2 < W; ++w2) {
              u64 tb = tw[w2];
              while (tb) {
                const int k = w2 * 64 + __builtin_ctzll(tb);
                tb &= tb - 1;
                if (TRI) {
                  // row k holds ranks above k only: its words start at
                  // (k + 1) / 64, which is w2, or w2 + 1 for a word's last bit
                  const bool last_bit = (k & 63) == 63;
#pragma unroll
                  for (int w = w2; w < W; ++w)
                    if ((w > w2 || !last_bit) && w < weff) c += (uint32_t)__popcll(tw[w] & row(k, w));
                } else {
#pragma unroll
                  for (int w = 0; w < W; ++w)
                    if (w < weff) c += (uint32_t)__popcll(tw[w] & row(k, w));
                }
              }
            }
          }
          acc += c;
        }
        wave_sync();
        if (lane + 64 < pending) jq[lane] = jq[lane + 64];
        wave_sync();
        pending -= take;
      }
    }
  }
  return acc;
}

// lut_count_wq with the row width fixed at compile time when weff lies in
// [WMIN, WMAX], so the unrolled word loops spend no issue slots on words
// past the task's last one (wave-uniform branch per task); other widths take
// the WMAX-wide loop with per-word predicates.
template <int WMIN, int WMAX, bool TRI, class Rows>
__device__ __forceinline__ u64 lut_count_wq_fit(Rows row, int s, int weff, int depth, int *jq, int *rowq, int lane,
                                                int a0 = 0, int astep = 1) {
  if constexpr (WMIN < WMAX) {
    if (weff == WMIN) return lut_count_wq<WMIN, TRI>(row, s, WMIN, depth, jq, rowq, lane, a0, astep);
    return lut_count_wq_fit<WMIN + 1, WMAX, TRI>(row, s, weff, depth, jq, rowq, lane, a0, astep);
  } else {
    return lut_count_wq<WMAX, TRI>(row, s, weff, depth, jq, rowq, lane, a0, astep);
  }
}

// Stage shared by the clique kernels: a sorted vertex set x (N+(v0) for a
// vertex task, W1 for an edge task), a second set x2, the LUT rows over x,
// the segment table and block-reduction slots.
template <int SMAX, int W, int NT>
struct CStage {
  u64 rows[W ? lut_words<W>(SMAX) : 1];  // first: row words at immediate offsets from the row base
  int x[SMAX];
  int x2[SMAX];
  Segs<SMAX> seg;
  int cnt[NT / 64];
  int flagrow[NT / 64][64];
  int ns, L, item, rowq;
};

struct Member {
  const int *sorted;  // generic pointer (LDS or HBM), strictly increasing
  int n;
  __device__ __forceinline__ bool contains(int w) const { return sorted_contains(sorted, n, w); }
};

// out[] <- a[0..na) ∩ m, in order; returns the size.  All threads call.
template <int NT, class St>
__device__ int block_intersect_ordered(const int32_t *a, int na, const Member &m, int *out, St &S) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int total = 0;
  for (int base = 0; base < na; base += NT) {
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
    for (int q = 0; q < NT / 64; ++q) {
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
template <int NT, int PCAP, class St>
__device__ u64 block_sum_intersections(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
                                       const int *verts, int nv, const Member &m, St &S, WaveFlags &fl) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  u64 acc = 0;
  for (int k0 = 0; k0 < nv; k0 += PCAP) {
    const int k1 = min(nv, k0 + PCAP);
    __syncthreads();
    if (wave == 0) {
      int ns, L;
      wave_segments(S.seg, rowptr, colidx, verts, k0, k1, -1, KeepAll(), lane, ns, L);
      if (lane == 0) { S.ns = ns; S.L = L; }
    }
    __syncthreads();
    int lo, hi;
    wave_share(S.L, wave, NT / 64, lo, hi);
    wave_walk<4, true>(S.seg, colidx, S.ns, lo, hi, lane, fl, [&](const WindowLane &r) {
      if (r.valid) acc += m.contains(r.w);
    });
  }
  __syncthreads();
  return acc;
}

// Build the LUT of the induced sub-DAG on S.x[0..s) (s <= SMAX) and count
// its cliques of `depth` vertices.  All threads call; returns this thread's
// partial count.
// Build the LUT of the subgraph induced on S.x[0..s) (s <= SMAX): row i
// gets bit j iff S.x[j] is in the list of S.x[i] (the sub-DAG for a DAG,
// the symmetric adjacency for an undirected graph).  S.rowq is zero on
// exit.  All threads call.
template <int SMAX, int W, int NT>
__device__ void block_lut_build(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
                                CStage<SMAX, W, NT> &S, int s, WaveFlags &fl) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int q = tid; q < lut_words<W>(s); q += NT) S.rows[q] = 0ull;
  __syncthreads();
  // every thread loads its members' list bounds; x2 (free here) is scratch
  block_segments<NT>(S.seg, S.x2, rowptr, colidx, S.x, 0, s, -1, KeepAll(), &S.ns, &S.L);
  int lo, hi;
  wave_share(S.L, wave, NT / 64, lo, hi);
  wave_walk<4, true>(S.seg, colidx, S.ns, lo, hi, lane, fl, [&](const WindowLane &r) {
    if (r.valid) {
      const int pos = lower_bound_dev(S.x, s, r.w);
      if (pos < s && S.x[pos] == r.w) atomicOr(&S.rows[lut_word<W>(S.seg.idx[r.j], pos >> 6)], 1ull << (pos & 63));
    }
  });
  if (tid == 0) S.rowq = 0;
  __syncthreads();
}

template <int SMAX, int W, int NT, int WMIN = 1>
__device__ u64 block_lut_count(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
                               CStage<SMAX, W, NT> &S, int s, int depth, WaveFlags &fl) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int weff = (s + 63) >> 6;
  block_lut_build<SMAX, W, NT>(rowptr, colidx, S, s, fl);
  u64 c;
  if (W >= 4 && (depth == 4 || (depth == 3 && W >= 16))) {
    // the segment table (dead after the build) holds the per-wave queues
    // (depth 3 on rows of <= 8 words measured faster per thread-word)
    static_assert(sizeof(S.seg) >= (NT / 64) * 128 * sizeof(int), "queue space");
    c = lut_count_wq_fit<WMIN, W, false>(LdsRows<W>{S.rows}, s, weff, depth,
                                         reinterpret_cast<int *>(&S.seg) + wave * 128, &S.rowq, lane);
  } else {
    c = lut_count<W>(S.rows, s, weff, depth, tid, NT);
  }
  __syncthreads();
  return c;
}

// Triangular variant of lut_count: rows are indexed in DAG (degree, id)
// order, so row a only holds bits b > a and every AND below starts at the
// word holding bit b+1 (words below it are never read, so the packed
// upper-triangular rows serve as well).  One thread per row word; only
// rows a0, a0 + astep, ... are counted.
template <int W, class Rows>
__device__ __forceinline__ u64 lut_count_tri_rows(Rows row, int s, int weff, int depth, int tid, int nthreads,
                                                   int a0 = 0, int astep = 1) {
  u64 acc = 0;
  // items (a, q) with q >= a / 64 only; enumerate the triangle row by row
  const int na = s > a0 ? (s - a0 + astep - 1) / astep : 0;
  const int items = na * weff;
  for (int t = tid; t < items; t += nthreads) {
    const int i = t / weff, q = t - i * weff, a = a0 + astep * i;
    if (q < (a >> 6)) continue;
    u64 bits = row(a, q);
    if (depth == 2) {
      acc += __popcll(bits);
      continue;
    }
    while (bits) {
      const int j = q * 64 + __builtin_ctzll(bits);
      bits &= bits - 1;
      const int ws = (j + 1) >> 6;
      if (depth == 3) {
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (w >= ws && w < weff) acc += __popcll(row(a, w) & row(j, w));
      } else {
        u64 tw[W];
#pragma unroll
        for (int w = 0; w < W; ++w) tw[w] = (w >= ws && w < weff) ? (row(a, w) & row(j, w)) : 0ull;
#pragma unroll
        for (int w2 = 0; w2 < W; ++w2) {
          u64 tb = tw[w2];
          while (tb) {
            const int k = w2 * 64 + __builtin_ctzll(tb);
            tb &= tb - 1;
            const int ks = (k + 1) >> 6;
#pragma unroll
            for (int w = 0; w < W; ++w)
              if (w >= ks && w < weff) acc += __popcll(tw[w] & row(k, w));
          }
        }
      }
    }
  }
  return acc;
}

template <int W>
__device__ __forceinline__ u64 lut_count_tri(const u64 *rows, int s, int weff, int depth, int tid, int nthreads) {
  return lut_count_tri_rows<W>(LdsRows<W>{rows}, s, weff, depth, tid, nthreads);
}

// In-LDS bitonic sort of n <= SMAX 64-bit keys with an int payload (padded
// to a power of two with ~0); all threads call.
template <int SMAX, int NT>
__device__ __forceinline__ void block_bitonic_sort(u64 *key, int *val, int n) {
  int m = 1;
  while (m < n) m <<= 1;
  for (int i = n + threadIdx.x; i < m; i += NT) key[i] = ~0ull;
  __syncthreads();
  for (int size = 2; size <= m; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < m; i += NT) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const u64 a = key[i], b = key[j];
          if ((a > b) == up) {
            key[i] = b;
            key[j] = a;
            const int t = val[i];
            val[i] = val[j];
            val[j] = t;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Member ranks of a rank-LUT task in an LDS hash: 2^bbits buckets of four
// member ids (16 bytes, empty -1, filled in order, so a bucket whose last
// slot is empty ends a lookup) and the rank of the id in slot i at rank[i].
// A probe's home bucket settles it except when that bucket is full, so a
// lookup is one 16-byte read plus one rank read on a hit, where a binary
// search over the sorted members took log2(s) dependent reads.
__device__ __forceinline__ void rank_hash_insert(int *ids, uint16_t *rank, int bbits, int key, int r) {
  const uint32_t mask = (1u << bbits) - 1u;
  uint32_t h = hash_slot(key, bbits);
  for (;;) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (atomicCAS(&ids[4 * h + q], -1, key) == -1) {
        rank[4 * h + q] = (uint16_t)r;
        return;
      }
    h = (h + 1) & mask;
  }
}

// Slot of key in bucket q (0..3), or -1.
__device__ __forceinline__ int bucket_slot(const int4 &q, int key) {
  int si = q.w == key ? 3 : -1;
  si = q.z == key ? 2 : si;
  si = q.y == key ? 1 : si;
  return q.x == key ? 0 : si;
}

// Rest of a lookup whose home bucket h was full and lacked key: the slot
// index 4 * bucket + i of key, or -1.
__device__ __noinline__ int rank_hash_after(const int *ids, int bbits, uint32_t h, int key) {
  const uint32_t mask = (1u << bbits) - 1u;
  for (;;) {
    h = (h + 1) & mask;
    const int4 q = reinterpret_cast<const int4 *>(ids)[h];
    const int si = bucket_slot(q, key);
    if (si >= 0) return 4 * (int)h + si;
    if (q.w < 0) return -1;
  }
}

// LUT build of a rank-LUT task: walks the flattened lists of the members
// (segment j = the list of the member of rank seg.idx[j]) and calls
// set_bit(slot, j) for every probed id the rank hash holds.  Invalid lanes
// probe kNoKey, which no bucket holds; the four home buckets of a window
// group are read together, and one wave-uniform branch per group covers
// the misses in full buckets.
template <class Seg, class SetBit>
__device__ __forceinline__ void rank_hash_walk(const Seg &seg, const int32_t *__restrict__ colidx, int ns, int lo,
                                               int hi, int lane, WaveFlags &fl, const int *ids, int bbits,
                                               SetBit &&set_bit) {
  wave_walk_group<4, true, kNoKey>(seg, colidx, ns, lo, hi, lane, fl, [&](const WindowLane (&r)[4]) {
    int4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = reinterpret_cast<const int4 *>(ids)[hash_slot(r[u].w, bbits)];
    uint32_t open = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int si = bucket_slot(q[u], r[u].w);
      if (si >= 0) set_bit(4 * (int)hash_slot(r[u].w, bbits) + si, r[u].j);
      open |= (si < 0 && q[u].w >= 0) ? 1u : 0u;
    }
    if (__builtin_expect(__ballot(open != 0u) != 0, 0)) {
      __asm__ volatile("" ::: "memory");  // home buckets read again, not kept live
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t h = hash_slot(r[u].w, bbits);
        const int4 q0 = reinterpret_cast<const int4 *>(ids)[h];
        if (q0.w >= 0 && bucket_slot(q0, r[u].w) < 0) {
          const int at = rank_hash_after(ids, bbits, h, r[u].w);
          if (at >= 0) set_bit(at, r[u].j);
        }
      }
    }
  });
}

// Rank-ordered vertex LUT (strategy LUT_RANK): the members of X = N+(v0)
// are ranked in (undirected degree, id) order, which is the DAG
// orientation order, so the LUT is upper-triangular.  A probed id's rank
// comes from the rank hash: 4 SMAX id slots (load <= 1/4) up to 256
// members, 2 SMAX above.  perm (rank of member q) and the segment-build
// scratch live in the rows' space until the rows are zeroed.
template <int SMAX, int W, int NT>
struct RStage {
  static constexpr int kSlots = SMAX <= 256 ? 4 * SMAX : 2 * SMAX;
  static constexpr int kHashBits = 31 - __builtin_clz(kSlots / 4);  // SMAX is a power of two here
  // the rows first: at LDS offset 0 a row word's address is the row base
  // plus an immediate offset, so the count loop adds nothing per read
  union {
    u64 rows[lut_words<W>(SMAX)];
    struct {
      int tmp[SMAX];  // scratch of the block-cooperative segment build
      uint16_t perm[SMAX];
    } pre;
  };
  alignas(16) int hid[kSlots];
  uint16_t hrank[kSlots];
  union {  // the sort keys are dead once the hash is built, before the segments
    struct {
      u64 key[SMAX];
      int val[SMAX];
    };
    Segs<SMAX> seg;
  };
  int flagrow[NT / 64][64];
  int ns, L, item, rowq;
};
static_assert(sizeof(RStage<64, 1, 64>::rows) >= sizeof(RStage<64, 1, 64>::pre), "pre fits the rows");

template <int SMAX, int W, int NT>
__global__ __launch_bounds__(NT, (SMAX <= 256 ? 7 : 1)) void k_clique_vlut_rank(const int64_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ colidx,
                                                         const int64_t *__restrict__ rrowptr,
                                                         const int32_t *__restrict__ tasks, int ntasks, int min_s,
                                                         int depth, int *__restrict__ queue, u64 *__restrict__ out,
                                                         u64 *__restrict__ err, u64 *__restrict__ tlog,
                                                         const int32_t *__restrict__ slots) {
  __shared__ RStage<SMAX, W, NT> S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(S.flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.item = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.item;
    if (t >= ntasks) break;
    const u64 t_start = tlog ? (u64)wall_clock64() : 0;
    u64 part = 0;
    const int v0 = tasks[t];
    const int64_t b = rowptr[v0];
    const int s = (int)(rowptr[v0 + 1] - b);
    if (s >= min_s) {
      const int weff = (s + 63) >> 6;
      const int32_t *nv = colidx + b;  // the members, id-sorted
      for (int q = tid; q < s; q += NT) {
        const int x = nv[q];
        const u64 deg = (u64)((rowptr[x + 1] - rowptr[x]) + (rrowptr[x + 1] - rrowptr[x]));
        S.key[q] = (deg << 32) | (u64)(uint32_t)x;  // (undirected degree, id): the orientation order
        S.val[q] = q;
      }
      using RS = RStage<SMAX, W, NT>;
      for (int q = tid; q < RS::kSlots; q += NT) S.hid[q] = -1;
      __syncthreads();
      block_bitonic_sort<SMAX, NT>(S.key, S.val, s);
      for (int r = tid; r < s; r += NT) {
        const int q = S.val[r];
        S.pre.perm[q] = (uint16_t)r;
        rank_hash_insert(S.hid, S.hrank, RS::kHashBits, nv[q], r);
      }
      __syncthreads();  // keys are overwritten by the segment table
      block_segments<NT>(S.seg, S.pre.tmp, rowptr, colidx, nv, 0, s, -1, KeepAll(), &S.ns, &S.L);
      // segment j's member index becomes its rank (the LUT row)
      for (int j = tid; j < S.ns; j += NT) S.seg.idx[j] = S.pre.perm[S.seg.idx[j]];
      __syncthreads();  // perm and tmp are dead: the rows take their space
      for (int q = tid; q < lut_words<W>(s); q += NT) S.rows[q] = 0ull;
      __syncthreads();
      int lo, hi;
      wave_share(S.L, wave, NT / 64, lo, hi);
      // a bit at or below the diagonal means the DAG is not in (degree, id)
      // orientation: flagged once per wave after the walk (the count is
      // then discarded), so the bit writes need no branch
      bool bad = false;
      rank_hash_walk(S.seg, colidx, S.ns, lo, hi, lane, fl, S.hid, RS::kHashBits, [&](int at, int j) {
        const int col = S.hrank[at], row = S.seg.idx[j];
        bad |= col <= row;
        atomicOr(&S.rows[lut_word<W>(row, col >> 6)], 1ull << (col & 63));
      });
      if (__ballot(bad) && lane == 0) atomicOr(err, 1ull);
      if (tid == 0) S.rowq = 0;
      __syncthreads();
      if (W >= 4 && (depth == 4 || (depth == 3 && W >= 16))) {
        // the sort keys (dead after the LUT build) hold the per-wave queues
        // (depth 3 on rows of <= 8 words measured faster per thread-word)
        static_assert(sizeof(S.key) + sizeof(S.val) >= (NT / 64) * 128 * sizeof(int), "queue space");
        // tasks one bin up also land here: widths from half the previous bin
        part += lut_count_wq_fit<(W >= 8 ? W / 4 + 1 : W), W, true>(
            LdsRows<W>{S.rows}, s, weff, depth, reinterpret_cast<int *>(S.key) + wave * 128, &S.rowq, lane);
      } else {
        part += lut_count_tri<W>(S.rows, s, weff, depth, tid, NT);
      }
    }
    task_flush(acc, part, slots, v0, out, tid & 63);
    __syncthreads();
    if (tlog && tid == 0) tlog[t] = (u64)wall_clock64() - t_start;
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Packed upper-triangular rows for the rank LUT: row a stores only its
// words w >= a/64 (the rank order puts every bit of row a above a), so s
// rows of weff words take about half of s * weff.  tri_base(a) + w is the
// address of word w of row a.
__host__ __device__ constexpr int tri_words(int s, int weff) {
  return 64 * ((s >> 6) * weff - (s >> 6) * ((s >> 6) - 1) / 2) + (s & 63) * (weff - (s >> 6));
}
// tri_words(a, weff) - a / 64 with 24-bit multiplies (every factor is below
// 2^11), which issue at the full rate where 32-bit ones take four slots:
// the count loops evaluate it for every row they touch
__device__ __forceinline__ int tri_base(int a, int weff) {
  const uint32_t q = (uint32_t)a >> 6, r = (uint32_t)a & 63u, e = (uint32_t)weff;
  const uint32_t full = __umul24(q, e) - (__umul24(q, q - 1u) >> 1);  // words of the q full 64-row blocks / 64
  return (int)((full << 6) + __umul24(r, e - q) - q);
}

struct TriRows {
  const u64 *rows;
  int weff;
  __device__ __forceinline__ u64 operator()(int r, int w) const { return rows[tri_base(r, weff) + w]; }
};

// Rank-ordered LUT with the rows packed upper-triangular in LDS: for the
// smaller hubs (1024 < d+(v0) <= 1344, 21 words per row, 118 KB instead of
// rows in HBM) and for the 512-row bin (18 KB instead of 33 KB, so more
// blocks fit a CU).  The segment-build scratch lives
// in the rows' space before they are zeroed; the sort keys hold the count
// queues afterwards.  Tasks outside [min_s, max_s] are left to other kernels.
constexpr int pow2_ceil(int n) { return n <= 1 ? 1 : 2 * pow2_ceil((n + 1) / 2); }

template <int SMAX, int W, int NT>
struct TStage {
  static constexpr int KP = pow2_ceil(SMAX);  // the bitonic sort pads to a power of two
  // up to 512 members: ranks from a rank hash of 2 KP id slots (load <= 1/2);
  // the hub bin (1344 members, whose rows fill the LDS) binary-searches x
  static constexpr bool kHash = SMAX <= 512;
  static constexpr int kSlots = kHash ? 2 * KP : 4;
  static constexpr int kHashBits = 31 - __builtin_clz(kSlots / 4);
  union {  // first: row words at immediate offsets from the row base
    u64 rows[tri_words(SMAX, W)];
    struct {
      int tmp[SMAX];  // scratch of the block-cooperative segment build
      uint16_t perm[kHash ? SMAX : 1];  // rank of member q (hash builds)
    } pre;
  };
  alignas(16) int hid[kSlots];
  uint16_t hrank[kSlots];
  int x[kHash ? 1 : SMAX];
  int perm[kHash ? 1 : SMAX];
  union {
    struct {
      u64 key[KP];
      int val[KP];
    };
    Segs<SMAX> seg;
  };
  int flagrow[NT / 64][64];
  int ns, L, item, rowq;
};

template <int SMAX, int W, int NT>
__global__ __launch_bounds__(NT) void k_clique_vlut_rank_tri(const int64_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ colidx,
                                                             const int64_t *__restrict__ rrowptr,
                                                             const int32_t *__restrict__ tasks, int ntasks, int min_s,
                                                             int max_s, int depth, int nparts, int *__restrict__ queue,
                                                             u64 *__restrict__ out, u64 *__restrict__ err,
                                                             u64 *__restrict__ tlog, const int32_t *__restrict__ slots) {
  static_assert(W * 64 >= SMAX, "stage bounds");
  __shared__ TStage<SMAX, W, NT> S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(S.flagrow[wave], lane);
  u64 acc = 0;
  // item = task * nparts + piece: every piece builds the task's whole LUT
  // and counts the rows a = piece (mod nparts), so a few large tasks spread
  // over many blocks instead of one block each
  const int nitems = ntasks * nparts;
  for (;;) {
    if (tid == 0) S.item = atomicAdd(queue, 1);
    __syncthreads();
    const int item = S.item;
    if (item >= nitems) break;
    const int t = item / nparts, piece = item - t * nparts;
    const u64 t_start = tlog ? (u64)wall_clock64() : 0;
    u64 part = 0;
    const int v0 = tasks[t];
    const int64_t b = rowptr[v0];
    const int s = (int)(rowptr[v0 + 1] - b);
    const bool mine = s >= min_s && s <= max_s;
    if (mine) {
      using TS = TStage<SMAX, W, NT>;
      const int weff = (s + 63) >> 6;
      const int32_t *nv = colidx + b;  // the members, id-sorted
      for (int q = tid; q < s; q += NT) {
        const int x = nv[q];
        if constexpr (!TS::kHash) S.x[q] = x;
        const u64 deg = (u64)((rowptr[x + 1] - rowptr[x]) + (rrowptr[x + 1] - rrowptr[x]));
        S.key[q] = (deg << 32) | (u64)(uint32_t)x;  // (undirected degree, id): the orientation order
        S.val[q] = q;
      }
      if constexpr (TS::kHash)
        for (int q = tid; q < TS::kSlots; q += NT) S.hid[q] = -1;
      __syncthreads();
      block_bitonic_sort<TS::KP, NT>(S.key, S.val, s);
      for (int r = tid; r < s; r += NT) {
        if constexpr (TS::kHash) {
          const int q = S.val[r];
          S.pre.perm[q] = (uint16_t)r;
          rank_hash_insert(S.hid, S.hrank, TS::kHashBits, nv[q], r);
        } else {
          S.perm[S.val[r]] = r;
        }
      }
      __syncthreads();  // keys are overwritten by the segment table
      block_segments<NT>(S.seg, S.pre.tmp, rowptr, colidx, nv, 0, s, -1, KeepAll(), &S.ns, &S.L);
      if constexpr (TS::kHash) {
        // segment j's member index becomes its rank (the LUT row) and the
        // row's base word: rank | tri_base(rank) << 10
        static_assert(SMAX <= 1024, "ranks fit 10 bits");
        for (int j = tid; j < S.ns; j += NT) {
          const int r = S.pre.perm[S.seg.idx[j]];
          S.seg.idx[j] = r | (tri_base(r, weff) << 10);
        }
        __syncthreads();  // perm is dead: the rows take its space
      }
      for (int q = tid; q < tri_words(s, weff); q += NT) S.rows[q] = 0ull;  // tmp is dead
      __syncthreads();
      int lo, hi;
      wave_share(S.L, wave, NT / 64, lo, hi);
      auto set_bit = [&](int col, int row) {
        if (col <= row)
          atomicOr(err, 1ull);  // DAG not in (degree, id) orientation
        else
          atomicOr(&S.rows[tri_base(row, weff) + (col >> 6)], 1ull << (col & 63));
      };
      if constexpr (TS::kHash) {
        // as in k_clique_vlut_rank: a bit at or below the diagonal is
        // flagged once per wave; its write stays inside the rows' space
        // (tri_base(row) + col / 64 with col <= row is this row's first
        // stored word or a word before it)
        bool bad = false;
        rank_hash_walk(S.seg, colidx, S.ns, lo, hi, lane, fl, S.hid, TS::kHashBits, [&](int at, int j) {
          const int v = S.seg.idx[j], row = v & 1023, col = S.hrank[at];
          bad |= col <= row;
          atomicOr(&S.rows[(v >> 10) + (col >> 6)], 1ull << (col & 63));
        });
        if (__ballot(bad) && lane == 0) atomicOr(err, 1ull);
      } else {
        wave_walk<4, true>(S.seg, colidx, S.ns, lo, hi, lane, fl, [&](const WindowLane &r) {
          if (r.valid) {
            const int pos = lower_bound_dev(S.x, s, r.w);
            if (pos < s && S.x[pos] == r.w) set_bit(S.perm[pos], S.perm[S.seg.idx[r.j]]);
          }
        });
      }
      if (depth == 3 && W < 16) {
        // K4 on rows of <= 8 words: one thread per row word (the rows'
        // few bits leave most lanes of the pair queues idle)
        __syncthreads();
        part += lut_count_tri_rows<W>(TriRows{S.rows, weff}, s, weff, depth, tid, NT, piece, nparts);
      } else {
        if (tid == 0) S.rowq = 0;
        __syncthreads();
        static_assert(sizeof(S.key) >= (NT / 64) * 128 * sizeof(int), "queue space");
        part += lut_count_wq_fit<(W > 16 ? 17 : W / 4 + 1), W, true>(
            TriRows{S.rows, weff}, s, weff, depth, reinterpret_cast<int *>(S.key) + wave * 128, &S.rowq, lane, piece,
            nparts);
      }
    }
    task_flush(acc, part, slots, v0, out, tid & 63);
    __syncthreads();
    if (mine && tlog && tid == 0) tlog[t] = (u64)wall_clock64() - t_start;
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Triangular counting over rows in HBM with a runtime word count (no
// per-thread row copies): K5 terms recompute row_a & row_j per word.
__device__ __forceinline__ u64 lut_count_tri_wide(const u64 *__restrict__ rows, int s, int weff, int depth, int tid,
                                                  int nthreads) {
  u64 acc = 0;
  const int items = s * weff;
  for (int t = tid; t < items; t += nthreads) {
    const int a = t / weff, q = t - a * weff;
    if (q < (a >> 6)) continue;
    const u64 *ra = rows + (size_t)a * weff;
    u64 bits = ra[q];
    if (depth == 2) {
      acc += __popcll(bits);
      continue;
    }
    while (bits) {
      const int j = q * 64 + __builtin_ctzll(bits);
      bits &= bits - 1;
      const u64 *rj = rows + (size_t)j * weff;
      const int ws = (j + 1) >> 6;
      if (depth == 3) {
        for (int w = ws; w < weff; ++w) acc += __popcll(ra[w] & rj[w]);
      } else {
        for (int w2 = ws; w2 < weff; ++w2) {
          u64 tb = ra[w2] & rj[w2];
          while (tb) {
            const int k = w2 * 64 + __builtin_ctzll(tb);
            tb &= tb - 1;
            const u64 *rk = rows + (size_t)k * weff;
            for (int w = (k + 1) >> 6; w < weff; ++w) acc += __popcll(ra[w] & rj[w] & rk[w]);
          }
        }
      }
    }
  }
  return acc;
}

// Rank-ordered vertex LUT for hubs (1024 < d+(v0) <= GMAX): the staging
// (ids, ranks, sort keys, segment table) stays in LDS, the s x s-bit rows
// live in this block's HBM slot (GMAX * GMAX / 8 bytes).
template <int GMAX, int NT>
struct GStage {
  int x[GMAX];
  int perm[GMAX];
  union {
    struct {
      u64 key[GMAX];
      int val[GMAX];
    };
    Segs<GMAX> seg;
  };
  int flagrow[NT / 64][64];
  int ns, L, item, rowq;
};

template <int GMAX, int NT>
__global__ __launch_bounds__(NT) void k_clique_vlut_rank_hbm(const int64_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ colidx,
                                                             const int64_t *__restrict__ rrowptr,
                                                             const int32_t *__restrict__ tasks, int ntasks, int min_s,
                                                             int depth, u64 *__restrict__ lut_slots,
                                                             int *__restrict__ queue, u64 *__restrict__ out,
                                                             u64 *__restrict__ err, u64 *__restrict__ tlog,
                                                             const int32_t *__restrict__ slots) {
  __shared__ GStage<GMAX, NT> S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveFlags fl;
  fl.init(S.flagrow[wave], lane);
  u64 *rows = lut_slots + (size_t)blockIdx.x * ((size_t)GMAX * GMAX / 64);
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.item = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.item;
    if (t >= ntasks) break;
    const u64 t_start = tlog ? (u64)wall_clock64() : 0;
    u64 part = 0;
    const int v0 = tasks[t];
    const int64_t b = rowptr[v0];
    const int s = (int)(rowptr[v0 + 1] - b);
    const bool mine = s >= min_s && s <= GMAX;
    if (mine) {
      const int weff = (s + 63) >> 6;
      for (int q = tid; q < s; q += NT) {
        const int x = colidx[b + q];
        S.x[q] = x;
        const u64 deg = (u64)((rowptr[x + 1] - rowptr[x]) + (rrowptr[x + 1] - rrowptr[x]));
        S.key[q] = (deg << 32) | (u64)(uint32_t)x;
        S.val[q] = q;
      }
      for (int q = tid; q < s * weff; q += NT) rows[q] = 0ull;
      __threadfence();  // zero fill reaches L2 before the atomics below
      __syncthreads();
      block_bitonic_sort<GMAX, NT>(S.key, S.val, s);
      for (int r = tid; r < s; r += NT) S.perm[S.val[r]] = r;
      __syncthreads();
      if (wave == 0) {
        int ns, L;
        wave_segments(S.seg, rowptr, colidx, S.x, 0, s, -1, KeepAll(), lane, ns, L);
        if (lane == 0) { S.ns = ns; S.L = L; }
      }
      __syncthreads();
      int lo, hi;
      wave_share(S.L, wave, NT / 64, lo, hi);
      wave_walk<4, true>(S.seg, colidx, S.ns, lo, hi, lane, fl, [&](const WindowLane &r) {
        if (r.valid) {
          const int pos = lower_bound_dev(S.x, s, r.w);
          if (pos < s && S.x[pos] == r.w) {
            const int col = S.perm[pos], row = S.perm[S.seg.idx[r.j]];
            if (col <= row) atomicOr(err, 1ull);
            atomicOr(&rows[(size_t)row * weff + (col >> 6)], 1ull << (col & 63));
          }
        }
      });
      // The bits were set by atomics performed in L2; an agent-scope fence
      // (acquire invalidates this CU's L1) is needed before plain loads, or
      // lines cached by the zero fill or an earlier task read stale.
      __threadfence();
      if (tid == 0) S.rowq = 0;
      __syncthreads();
      if (depth >= 3 && weff <= 32) {
        // the sort keys (dead after the LUT build) hold the per-wave queues
        static_assert(sizeof(S.key) >= (NT / 64) * 128 * sizeof(int), "queue space");
        part += lut_count_wq<32, true>(HbmRows{rows, weff}, s, weff, depth, reinterpret_cast<int *>(S.key) + wave * 128,
                                       &S.rowq, lane);
      } else {
        part += lut_count_tri_wide(rows, s, weff, depth, tid, NT);
      }
    }
    task_flush(acc, part, slots, v0, out, tid & 63);
    __syncthreads();
    if (mine && tlog && tid == 0) tlog[t] = (u64)wall_clock64() - t_start;
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Vertex-task LUT kernel: one block (NT threads) per vertex v0 with
// min_s <= d+(v0) <= SMAX.  depth = k - 1 (K4 -> triangles, K5 -> 4-cliques).
template <int SMAX, int W, int NT>
__global__ __launch_bounds__(NT) void k_clique_vlut(const int64_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx,
                                                    const int32_t *__restrict__ tasks, int ntasks, int min_s,
                                                    int depth, int *__restrict__ queue, u64 *__restrict__ out,
                                                    u64 *__restrict__ tlog, const int32_t *__restrict__ slots) {
  __shared__ CStage<SMAX, W, NT> S;
  const int tid = threadIdx.x;
  WaveFlags fl;
  fl.init(S.flagrow[tid >> 6], tid & 63);
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.item = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.item;
    if (t >= ntasks) break;
    const u64 t_start = tlog ? (u64)wall_clock64() : 0;
    u64 part = 0;
    const int v0 = tasks[t];
    const int64_t b = rowptr[v0];
    const int s = (int)(rowptr[v0 + 1] - b);
    if (s >= min_s) {
      for (int q = tid; q < s; q += NT) S.x[q] = colidx[b + q];
      part += block_lut_count<SMAX, W, NT, (W >= 8 ? SMAX / 128 + 1 : 1)>(rowptr, colidx, S, s, depth, fl);
    }
    task_flush(acc, part, slots, v0, out, tid & 63);
    __syncthreads();
    if (tlog && tid == 0) tlog[t] = (u64)wall_clock64() - t_start;
  }
  acc = wave_sum_u64(acc);
  if ((tid & 63) == 0 && acc) atomicAdd(out, acc);
}

// Packed vertex-LUT kernel for s = d+(v0) <= 64 (one 64-bit word per row):
// a wave claims G tasks with one atomic; the G sets N+(v0) are staged in
// one flattened pass and the lists of all their members form one segment
// table whose index is the LUT row (g * 64 + i), so the dependent loads of
// G tasks are in flight together.  Task g's members are found through its
// own id hash (kPackBuckets buckets of four slots, load <= 1/2; hpos = the
// member's position in N+(v0), its LUT column).  Counting runs lut_count
// per task.
constexpr int kPackBucketBits = 5, kPackSlots = 4 << kPackBucketBits;
template <int G>
struct LutPack {
  alignas(16) int hid[G * kPackSlots];
  uint8_t hpos[G * kPackSlots];
  u64 rows[G * 64];
  Segs<64 * G, true> seg;
  int v0[G], s[G], off[G + 1];
  uint32_t tcnt[G];  // per-task counts (instances)
  int64_t b[G];
};

// Insert key with value v into a hash of 4-slot id buckets (as
// rank_hash_insert, byte values).
__device__ __forceinline__ void id_hash_insert(int *ids, uint8_t *val, int bbits, int key, int v) {
  const uint32_t mask = (1u << bbits) - 1u;
  uint32_t h = hash_slot(key, bbits);
  for (;;) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (atomicCAS(&ids[4 * h + q], -1, key) == -1) {
        val[4 * h + q] = (uint8_t)v;
        return;
      }
    h = (h + 1) & mask;
  }
}

template <int G>
__global__ __launch_bounds__(128) void k_clique_vlut_packed(const int64_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ colidx,
                                                            const int32_t *__restrict__ tasks, int ntasks, int min_s,
                                                            int depth, int *__restrict__ queue, u64 *__restrict__ out,
                                                            const int32_t *__restrict__ slots) {
  __shared__ LutPack<G> stage[2];
  __shared__ int flagrow[2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  LutPack<G> &S = stage[wave];
  WaveFlags fl;
  fl.init(flagrow[wave], lane);
  u64 acc = 0;
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(queue, G);
    base = __shfl(base, 0, 64);
    if (base >= ntasks) break;
    const int ng = min(G, ntasks - base);
    int sz = 0;
    if (lane < ng) {
      const int v = tasks[base + lane];
      const int64_t b = rowptr[v];
      sz = (int)(rowptr[v + 1] - b);
      if (sz < min_s) sz = 0;  // no clique of this size through v
      S.v0[lane] = v;
      S.b[lane] = b;
      S.s[lane] = sz;
      S.tcnt[lane] = 0;
    }
    const int inc = wave_incl_scan(sz, lane);
    if (lane < ng) S.off[lane + 1] = inc;
    if (lane == 0) S.off[0] = 0;
    for (int q = lane; q < G * 64; q += 64) S.rows[q] = 0ull;
    for (int q = lane; q < ng * kPackSlots; q += 64) S.hid[q] = -1;
    wave_sync();
    const int tot = __shfl(inc, ng - 1, 64);
    auto item_of = [&](int q) {
      int g = 0;
#pragma unroll
      for (int h = 1; h < G; ++h) g += (h < ng && q >= S.off[h]) ? 1 : 0;
      return g;
    };
    for (int q = lane; q < tot; q += 64) {
      const int g = item_of(q), i = q - S.off[g];
      id_hash_insert(S.hid + g * kPackSlots, S.hpos + g * kPackSlots, kPackBucketBits, colidx[S.b[g] + i], i);
    }
    wave_sync();
    int ns = 0, L = 0;
    wave_segments_fn(S.seg, rowptr, 0, tot,
                     [&](int q) {
                       const int g = item_of(q);
                       return colidx[S.b[g] + (q - S.off[g])];
                     },
                     [](int, int, int) { return true; }, lane, ns, L);
    if (lane == 0) S.seg.os[ns] = make_int2(0, L);
    wave_sync();
    for (int j = lane; j < ns; j += 64) {
      const int q = S.seg.idx[j], g = item_of(q);
      S.seg.idx[j] = g * 64 + (q - S.off[g]);  // LUT row of the segment's member
    }
    wave_sync();
    // invalid lanes probe kNoKey in table 0, which holds no such id; the
    // four home buckets of a window group are read together, and one
    // wave-uniform branch per group covers the misses in full buckets
    wave_walk_group<4, true, kNoKey>(S.seg, colidx, ns, 0, L, lane, fl, [&](const WindowLane (&r)[4]) {
      int row[4];
      int4 q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        row[u] = r[u].valid ? S.seg.idx[r[u].j] : 0;
        q[u] = reinterpret_cast<const int4 *>(S.hid + (row[u] >> 6) * kPackSlots)[hash_slot(r[u].w, kPackBucketBits)];
      }
      uint32_t open = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int si = bucket_slot(q[u], r[u].w);
        if (si >= 0) {
          const int at = (row[u] >> 6) * kPackSlots + 4 * (int)hash_slot(r[u].w, kPackBucketBits) + si;
          atomicOr(&S.rows[row[u]], 1ull << S.hpos[at]);
        }
        open |= (si < 0 && q[u].w >= 0) ? 1u : 0u;
      }
      if (__builtin_expect(__ballot(open != 0u) != 0, 0)) {
        __asm__ volatile("" ::: "memory");  // home buckets read again, not kept live
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int *tab = S.hid + (row[u] >> 6) * kPackSlots;
          const uint32_t h = hash_slot(r[u].w, kPackBucketBits);
          const int4 q0 = reinterpret_cast<const int4 *>(tab)[h];
          if (q0.w >= 0 && bucket_slot(q0, r[u].w) < 0) {
            const int at = rank_hash_after(tab, kPackBucketBits, h, r[u].w);
            if (at >= 0) atomicOr(&S.rows[row[u]], 1ull << S.hpos[(row[u] >> 6) * kPackSlots + at]);
          }
        }
      }
    });
    wave_sync();
    // count over the rows of all ng tasks at once, one row per lane (row i
    // of task g at flattened position off[g] + i), so a round of 64 rows is
    // not spent on one small task; as lut_count with one word per row
    for (int t = lane; t < tot; t += 64) {
      const int g = item_of(t);
      const u64 *rg = S.rows + g * 64;
      const u64 ri = rg[t - S.off[g]];
      uint32_t c = 0;  // below C(64, 4)
      if (depth == 2) {
        c = (uint32_t)__popcll(ri);
      } else {
        u64 bits = ri;
        while (bits) {
          const int j = __builtin_ctzll(bits);
          bits &= bits - 1;
          const u64 m = ri & rg[j];
          if (depth == 3) {
            c += (uint32_t)__popcll(m);
          } else {
            u64 tb = m;
            while (tb) {
              const int k = __builtin_ctzll(tb);
              tb &= tb - 1;
              c += (uint32_t)__popcll(m & rg[k]);
            }
          }
        }
      }
      if (!slots) {
        acc += c;
      } else if (c) {
        atomicAdd(&S.tcnt[g], c);
      }
    }
    if (slots) {
      wave_sync();
      if (lane < ng && S.tcnt[lane]) atomicAdd(out + slots[S.v0[lane]], (u64)S.tcnt[lane]);
    }
    wave_sync();  // stage is rewritten by the next claim
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Edge-task kernel (G²Miner edge parallelism): one block per DAG edge
// (v0, v1).  W1 = N+(v0) ∩ N+(v1) is materialised in order; with LUT it is
// encoded as a bitmap sub-DAG and K4 adds its edges, K5 its triangles
// (GLUMIN's edge-subgraph scheme); without LUT (or |W1| > SMAX) the nested
// list intersections of automine_4clique / automine_5clique run instead.
// W1 follows the reference's hybrid rule (intersect_num_hybrid,
// xgminer/include/glumin/operations.cuh:208-300): when both lists exceed
// kHybridMin, wave 0 merges them along the wave64 merge path; otherwise
// every thread binary-searches one element of N+(v1) in N+(v0).
// scratch: per block 2*md ints in HBM for sets larger than SMAX.
template <bool ON>
struct EdgeMergeLds {  // wave 0's merge-path tiles (the 1024-row edge kernels only)
  int As[kMergeTile];
  int Bs[kMergeTile + 1];
  int split[64];
  int n;
};
template <>
struct EdgeMergeLds<false> {
  int n;
};

template <int SMAX, int W, int NT>
__global__ __launch_bounds__(NT) void k_clique_edge(const int64_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx,
                                                    const int2 *__restrict__ items, int nitems, int k,
                                                    int *__restrict__ scratch, int md,
                                                    int *__restrict__ queue, u64 *__restrict__ out,
                                                    u64 *__restrict__ tlog, const int32_t *__restrict__ slots) {
  __shared__ CStage<SMAX, W, NT> S;
  __shared__ EdgeMergeLds<(SMAX >= 1024)> mg;
  const int tid = threadIdx.x;
  WaveFlags fl;
  fl.init(S.flagrow[tid >> 6], tid & 63);
  int *g_w1 = scratch + (size_t)blockIdx.x * 2 * md;
  int *g_w2 = g_w1 + md;
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.item = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.item;
    if (t >= nitems) break;
    const u64 t_start = tlog ? (u64)wall_clock64() : 0;
    u64 part = 0;
    const int2 it = items[t];
    const int v0 = it.x;
    const int64_t b0 = rowptr[it.x], b1 = rowptr[it.y];
    const int d0 = (int)(rowptr[it.x + 1] - b0), d1 = (int)(rowptr[it.y + 1] - b1);
    if (d0 >= k - 2 && d1 >= k - 2) {
      int *w1 = (min(d0, d1) <= SMAX) ? S.x : g_w1;
      int n1 = 0;
      bool merged = false;
      if constexpr (SMAX >= 1024) {
        if (d0 > kHybridMin && d1 > kHybridMin) {
          if (tid < 64) {
            const uint32_t c = merge_walk(colidx + b1, d1, colidx + b0, d0, mg.As, mg.Bs, mg.split, tid,
                                          [](int, bool found) { return found; }, w1);
            if (tid == 0) mg.n = (int)c;
          }
          __syncthreads();  // W1 written by wave 0
          n1 = mg.n;
          merged = true;
        }
      }
      if (!merged) n1 = block_intersect_ordered<NT>(colidx + b1, d1, Member{colidx + b0, d0}, w1, S);
      if (n1 >= k - 2) {
        bool done = false;
        if constexpr (W > 0) {
          if (w1 == S.x) {
            part += block_lut_count<SMAX, W, NT>(rowptr, colidx, S, n1, k - 2, fl);
            done = true;
          }
        }
        if (done) {
        } else if (k == 4) {
          part += block_sum_intersections<NT, SMAX>(rowptr, colidx, w1, n1, Member{w1, n1}, S, fl);
        } else {
          for (int c = 0; c < n1; ++c) {
            const int w = w1[c];
            const int64_t bw = rowptr[w];
            const int dw = (int)(rowptr[w + 1] - bw);
            if (dw < 2) continue;
            int *w2 = (min(n1, dw) <= SMAX) ? S.x2 : g_w2;
            const int n2 = block_intersect_ordered<NT>(colidx + bw, dw, Member{w1, n1}, w2, S);
            if (n2 < 2) continue;
            part += block_sum_intersections<NT, SMAX>(rowptr, colidx, w2, n2, Member{w2, n2}, S, fl);
          }
        }
      }
    }
    task_flush(acc, part, slots, v0, out, tid & 63);
    __syncthreads();
    if (tlog && tid == 0) tlog[t] = (u64)wall_clock64() - t_start;
  }
  acc = wave_sum_u64(acc);
  if ((tid & 63) == 0 && acc) atomicAdd(out, acc);
}

// ------------------------------------------------- vertex-induced 4-motifs
// GLUMIN P1 (3-star), P2 (diamond), P3 (tailed triangle) on an undirected
// graph, each induced embedding counted once (GraphMiner's unique-embedding
// semantics; GLUMIN's P2 walks both directions of every edge and reports
// twice the diamonds).  Reference kernels: GM_kernels/P1-GM.yml.cuh,
// P2-GM.yml.cuh, P3-GM.yml.cuh (set difference / intersection over edge
// tasks) and GM_LUT_kernels/P1.yml.cuh (LUT rows with AND-NOT popcounts,
// include/glumin/codegen_bitmap.cuh:357-407).  Vertex task v0, S = N(v0),
// R[i] = {j : S[j] in N(S[i])} (symmetric; S sorted by id):
//   3-star centred at v0:        sum_{a} sum_{j<a, j not in R[a]} |~R[a] & ~R[j] & [0, j)|
//   diamond, diagonal (v0, S[a]) with S[a] < v0:
//                                sum_{j in R[a]} |R[a] & ~R[j] & [0, j)|
//   tailed triangle, v0 the vertex of degree 3:
//                                sum_{a} sum_{j<a, j in R[a]} |~R[a] & ~R[j]|
// (a leaf k of the star is adjacent to neither a nor j; the two vertices
// off the diagonal are adjacent to S[a] but not to each other; the tail
// is adjacent to v0 only -- ~R[j] drops a and ~R[a] drops j).
enum { VM_STAR = 0, VM_DIAMOND = 1, VM_TAILED = 2 };

__device__ __forceinline__ u64 low_bits(int j, int w) {  // bits [0, j) of word w
  const int jw = j >> 6;
  return w < jw ? ~0ull : (w == jw ? ((1ull << (j & 63)) - 1ull) : 0ull);
}

// Waves claim rows a from *rowq; lanes take the partners j.  All lanes of
// every wave call.
template <int W, int WHICH>
__device__ u64 vmotif_lut_count(const u64 *rows, const int *x, int s, int v0, int *rowq, int lane) {
  const int weff = (s + 63) >> 6;
  const u64 last = (s & 63) ? ((1ull << (s & 63)) - 1ull) : ~0ull;  // valid bits of word weff - 1
  u64 acc = 0;
  for (;;) {
    int a = 0;
    if (lane == 0) a = atomicAdd(rowq, 1);
    a = __shfl(a, 0, 64);
    if (a >= s) break;
    if (WHICH == VM_DIAMOND && x[a] >= v0) break;  // x sorted: no later row has S[a] < v0
    u64 ra[W];
#pragma unroll
    for (int w = 0; w < W; ++w) ra[w] = w < weff ? rows[lut_word<W>(a, w)] : 0ull;
    const int jmax = WHICH == VM_DIAMOND ? s : a;
    uint32_t c = 0;
    for (int j = lane; j < jmax; j += 64) {
      const bool adj = (rows[lut_word<W>(a, j >> 6)] >> (j & 63)) & 1ull;
      if (adj != (WHICH != VM_STAR)) continue;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        if (w >= weff) continue;
        const u64 rj = rows[lut_word<W>(j, w)];
        u64 m;
        if (WHICH == VM_STAR) m = ~ra[w] & ~rj & low_bits(j, w);
        else if (WHICH == VM_DIAMOND) m = ra[w] & ~rj & low_bits(j, w);
        else m = ~ra[w] & ~rj & (w == weff - 1 ? last : ~0ull);
        c += (uint32_t)__popcll(m);
      }
    }
    acc += c;
  }
  return acc;
}

// Vertex tasks v0 with 3 <= d(v0) <= SMAX: the LUT over N(v0) in LDS.
template <int SMAX, int W, int NT, int WHICH>
__global__ __launch_bounds__(NT) void k_vmotif_lut(const int64_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ colidx,
                                                   const int32_t *__restrict__ tasks, int ntasks,
                                                   int *__restrict__ queue, u64 *__restrict__ out,
                                                   const int32_t *__restrict__ slots) {
  __shared__ CStage<SMAX, W, NT> S;
  const int tid = threadIdx.x;
  WaveFlags fl;
  fl.init(S.flagrow[tid >> 6], tid & 63);
  u64 acc = 0;
  for (;;) {
    if (tid == 0) S.item = atomicAdd(queue, 1);
    __syncthreads();
    const int t = S.item;
    if (t >= ntasks) break;
    const int v0 = tasks[t];
    const int64_t b = rowptr[v0];
    const int s = (int)(rowptr[v0 + 1] - b);
    for (int q = tid; q < s; q += NT) S.x[q] = colidx[b + q];
    block_lut_build<SMAX, W, NT>(rowptr, colidx, S, s, fl);
    const u64 part = vmotif_lut_count<W, WHICH>(S.rows, S.x, s, v0, &S.rowq, tid & 63);
    task_flush(acc, part, slots, v0, out, tid & 63);
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if ((tid & 63) == 0 && acc) atomicAdd(out, acc);
}

// Wave-cooperative |A ∩ B| (A: HBM scratch or graph list, B: graph list):
// the merge path when both exceed kHybridMin, else every element of the
// shorter list binary-searched in the longer (the hybrid rule).
struct VmWave {
  int As[kMergeTile];
  int Bs[kMergeTile + 1];
  int split[64];
};
__device__ __forceinline__ uint32_t wave_isect(const int32_t *A, int la, const int32_t *B, int lb, VmWave &M,
                                               int lane) {
  if (la > kHybridMin && lb > kHybridMin)
    return merge_walk(A, la, B, lb, M.As, M.Bs, M.split, lane, [](int, bool f) { return f; }, nullptr);
  const int32_t *T = la <= lb ? A : B, *L = la <= lb ? B : A;
  const int lt = min(la, lb), ll = max(la, lb);
  uint32_t c = 0;
  for (int k = lane; k < lt; k += 64) c += sorted_contains(L, ll, T[k]) ? 1u : 0u;
  return (uint32_t)wave_sum_i32((int)c);
}

// out[0..) <- {x in A : (x in B) == want_in, x < ub}, in order; returns the
// size (same on every lane).
__device__ __forceinline__ int wave_filter(const int32_t *A, int la, const int32_t *B, int lb, bool want_in, int ub,
                                           int32_t *out, int lane) {
  int n = 0;
  for (int k0 = 0; k0 < la; k0 += 64) {
    const int k = k0 + lane;
    bool keep = false;
    int x = 0;
    if (k < la) {
      x = A[k];
      keep = x < ub && sorted_contains(B, lb, x) == want_in;
    }
    const u64 m = __ballot(keep);
    if (keep) out[n + __popcll(m & lanemask_lt(lane))] = x;
    n += __popcll(m);
  }
  return n;
}

// Any degree, by set difference and intersection over the edge tasks of
// v0 (GLUMIN's GM kernels; a block per vertex task, its waves take the
// neighbours v1 in turn).  For v1 in N(v0), with GLUMIN's difference (no
// owner exclusion) and W0, W1 written to this wave's HBM slots:
//   3-star:  W0 = N(v0) \ N(v1) below v1; for v2 = W0[t]: t - |W0[0, t) ∩ N(v2)|
//   diamond (v1 < v0 only): W0 = N(v0) ∩ N(v1); for v2 = W0[t]: t - |W0[0, t) ∩ N(v2)|
//   tailed:  W0 = N(v0) ∩ N(v1) below v1, W1 = N(v0) \ N(v1) (v1 itself stays in
//            W1 and drops out as a neighbour of v2); for v2 in W0: |W1| - |W1 ∩ N(v2)|
// scratch: gridDim.x * 4 waves * 2 slots * md ints.
template <int WHICH>
__global__ __launch_bounds__(256) void k_vmotif_list(const int64_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colidx,
                                                     const int32_t *__restrict__ tasks, int ntasks,
                                                     int32_t *__restrict__ scratch, int md,
                                                     int *__restrict__ queue, u64 *__restrict__ out,
                                                     const int32_t *__restrict__ slots) {
  __shared__ VmWave M[4];
  __shared__ int sh_task, sh_next;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int32_t *W0 = scratch + ((size_t)blockIdx.x * 4 + wave) * 2 * (size_t)md;
  int32_t *W1 = W0 + md;
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
    const int32_t *N0 = colidx + rowptr[v0];
    const int d0 = (int)(rowptr[v0 + 1] - rowptr[v0]);
    u64 part = 0;
    for (;;) {
      int i = 0;
      if (lane == 0) i = atomicAdd(&sh_next, 1);
      i = __shfl(i, 0, 64);
      if (i >= d0) break;
      const int v1 = N0[i];
      if (WHICH == VM_DIAMOND && v1 >= v0) continue;
      const int32_t *N1 = colidx + rowptr[v1];
      const int d1 = (int)(rowptr[v1 + 1] - rowptr[v1]);
      const int n0 = wave_filter(N0, d0, N1, d1, WHICH != VM_STAR, WHICH == VM_DIAMOND ? INT32_MAX : v1, W0, lane);
      const int n1 = WHICH == VM_TAILED ? wave_filter(N0, d0, N1, d1, false, INT32_MAX, W1, lane) : 0;
      __threadfence_block();  // W0 / W1 written by other lanes of this wave
      wave_sync();
      uint32_t c = 0;
      for (int q = 0; q < n0; ++q) {
        const int v2 = W0[q];
        const int32_t *N2 = colidx + rowptr[v2];
        const int d2 = (int)(rowptr[v2 + 1] - rowptr[v2]);
        if (WHICH == VM_TAILED) c += (uint32_t)n1 - wave_isect(W1, n1, N2, d2, M[wave], lane);
        else c += (uint32_t)q - wave_isect(W0, q, N2, d2, M[wave], lane);
      }
      if (lane == 0) part += c;
      wave_sync();  // W0 / W1 reads done before the next v1 overwrites them
    }
    task_flush(acc, part, slots, v0, out, lane);
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

template <class K>
int grid_for(K kernel, int nt, size_t shmem, int num_cu, int64_t tasks) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, shmem) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return (int)std::max<int64_t>(1, std::min<int64_t>(tasks, (int64_t)per_cu * num_cu));
}

template <int SMAX, int W, int NT>
hipError_t run_vlut(const char *name, const DevCSR &g, const TaskList &tl, int k, int *queue, u64 *tlog, Launch &l) {
  if (!tl.count) return hipSuccess;
  auto kern = k_clique_vlut<SMAX, W, NT>;
  int grid = grid_for(kern, NT, 0, l.num_cu, tl.count);
  l.begin(name, tl.count);
  kern<<<grid, NT, 0, l.stream>>>(g.rowptr, g.colidx, tl.tasks, (int)tl.count, k - 1, k - 1, queue, l.d_out, tlog,
                                  l.slots);
  l.end();
  return hipGetLastError();
}

template <int SMAX, int W, int NT>
hipError_t run_vlut_rank(const char *name, const DevCSR &g, const TaskList &tl, int k, int *queue, u64 *tlog,
                         Launch &l) {
  if (!tl.count) return hipSuccess;
  auto kern = k_clique_vlut_rank<SMAX, W, NT>;
  int grid = grid_for(kern, NT, 0, l.num_cu, tl.count);
  l.begin(name, tl.count);
  kern<<<grid, NT, 0, l.stream>>>(g.rowptr, g.colidx, g.rrowptr, tl.tasks, (int)tl.count, k - 1, k - 1, queue,
                                  l.d_out, l.d_err, tlog, l.slots);
  l.end();
  return hipGetLastError();
}

// Rank-LUT bin with packed triangular rows (every task of the list).
template <int SMAX, int W, int NT>
hipError_t run_vlut_rank_tri(const char *name, const DevCSR &g, const TaskList &tl, int k, int *queue, u64 *tlog,
                             Launch &l) {
  if (!tl.count) return hipSuccess;
  auto kern = k_clique_vlut_rank_tri<SMAX, W, NT>;
  int grid = grid_for(kern, NT, 0, l.num_cu, tl.count);
  l.begin(name, tl.count);
  kern<<<grid, NT, 0, l.stream>>>(g.rowptr, g.colidx, g.rrowptr, tl.tasks, (int)tl.count, k - 1, SMAX, k - 1, 1,
                                  queue, l.d_out, l.d_err, tlog, l.slots);
  l.end();
  return hipGetLastError();
}

constexpr int kHubPieces = 64;  // most pieces one hub task of the triangular LDS kernel is split into

// Hub tasks (1024 < d+ <= kHubLutMax): those up to kTriLutMax with packed
// triangular rows in LDS, the rest with rows in HBM.  Both kernels walk the
// whole list and skip the other's tasks (queue and queue + 1).
hipError_t run_vlut_rank_hubs(const char *name_tri, const char *name_hbm, const DevCSR &g, const TaskList &tl, int k,
                              u64 *slots, int *queue, u64 *tlog, Launch &l) {
  if (!tl.count) return hipSuccess;
  auto tri = k_clique_vlut_rank_tri<kTriLutMax, kTriLutMax / 64, 1024>;
  // a hub task per block would leave most CUs idle behind the few largest
  // tasks: split each into pieces (rows a = piece mod nparts) so that the
  // list fills about two waves of resident blocks; per-task timing logs keep
  // one block per task
  const int resident = grid_for(tri, 1024, 0, l.num_cu, 1 << 30);
  const int nparts = tlog ? 1 : (int)std::min<int64_t>(kHubPieces, std::max<int64_t>(1, (2 * resident + tl.count - 1) / tl.count));
  int grid = grid_for(tri, 1024, 0, l.num_cu, tl.count * nparts);
  l.begin(name_tri, tl.count);
  tri<<<grid, 1024, 0, l.stream>>>(g.rowptr, g.colidx, g.rrowptr, tl.tasks, (int)tl.count, 1024 + 1,
                                   kTriLutMax, k - 1, nparts, queue, l.d_out, l.d_err, tlog, l.slots);
  l.end();
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  auto kern = k_clique_vlut_rank_hbm<kHubLutMax, 1024>;
  grid = std::min(grid_for(kern, 1024, 0, l.num_cu, tl.count), clique_hub_slots(l.num_cu));
  l.begin(name_hbm, tl.count);
  kern<<<grid, 1024, 0, l.stream>>>(g.rowptr, g.colidx, g.rrowptr, tl.tasks, (int)tl.count, kTriLutMax + 1, k - 1,
                                    slots, queue + 1, l.d_out, l.d_err, tlog, l.slots);
  l.end();
  return hipGetLastError();
}

template <int SMAX, int W, int NT>
hipError_t run_edge(const char *name, const DevCSR &g, const EdgeItems &ei, int k, int *scratch, int md,
                    int *queue, u64 *tlog, Launch &l) {
  if (!ei.count) return hipSuccess;
  auto kern = k_clique_edge<SMAX, W, NT>;
  int grid = grid_for(kern, NT, 0, l.num_cu, ei.count);
  l.begin(name, ei.count);
  kern<<<grid, NT, 0, l.stream>>>(g.rowptr, g.colidx, (const int2 *)ei.items, (int)ei.count, k, scratch, md, queue,
                                  l.d_out, tlog, l.slots);
  l.end();
  return hipGetLastError();
}

}  // namespace

// Resident blocks of the kernel that runs schedule list `li` (0-3 LUT bins,
// 4-7 rank LUT bins, 8-9 edge LUT, 10-11 edge lists): a task's share of
// that kernel's time is its wall time divided by this.
int clique_list_capacity(int num_cu, int li) {
  const int64_t inf = 1 << 30;
  switch (li) {
    case 0: return grid_for(k_clique_vlut<64, 1, 64>, 64, 0, num_cu, inf);
    case 1: return grid_for(k_clique_vlut<256, 4, 256>, 256, 0, num_cu, inf);
    case 2: return grid_for(k_clique_vlut<512, 8, 256>, 256, 0, num_cu, inf);
    case 3: return grid_for(k_clique_vlut<1024, 16, 1024>, 1024, 0, num_cu, inf);
    case 4: return grid_for(k_clique_vlut_rank<64, 1, 64>, 64, 0, num_cu, inf);
    case 5: return grid_for(k_clique_vlut_rank<256, 4, 256>, 256, 0, num_cu, inf);
    case 6: return grid_for(k_clique_vlut_rank_tri<512, 8, 256>, 256, 0, num_cu, inf);
    case 7: return grid_for(k_clique_vlut_rank<1024, 16, 1024>, 1024, 0, num_cu, inf);
    case 8: return grid_for(k_clique_edge<128, 2, 64>, 64, 0, num_cu, inf);
    case 9: return grid_for(k_clique_edge<1024, 16, 256>, 256, 0, num_cu, inf);
    case 10: return grid_for(k_clique_edge<128, 0, 64>, 64, 0, num_cu, inf);
    case 11: return grid_for(k_clique_edge<1024, 0, 256>, 256, 0, num_cu, inf);
    default:
      return std::min(grid_for(k_clique_vlut_rank_hbm<kHubLutMax, 1024>, 1024, 0, num_cu, inf),
                      clique_hub_slots(num_cu));
  }
}

int clique_hub_slots(int num_cu) { return num_cu; }

int clique_edge_grid(int num_cu) {
  int a = grid_for(k_clique_edge<128, 2, 64>, 64, 0, num_cu, 1 << 30);
  int b = grid_for(k_clique_edge<1024, 16, 256>, 256, 0, num_cu, 1 << 30);
  int c = grid_for(k_clique_edge<128, 0, 64>, 64, 0, num_cu, 1 << 30);
  int d = grid_for(k_clique_edge<1024, 0, 256>, 256, 0, num_cu, 1 << 30);
  return std::max(std::max(a, b), std::max(c, d));
}

hipError_t launch_clique(const DevCSR &g, const CliqueTasks &ct, int k, int *scratch, int md, Launch &l,
                         int *d_queues) {
  hipError_t e;
  const bool k4 = (k == 4);
  // edge items first: they carry the hub vertices
  if ((e = run_edge<1024, 16, 256>(k4 ? "k4_edge_lut_1024" : "k5_edge_lut_1024", g, ct.elut[1], k, scratch, md,
                                   d_queues + 8, ct.log_elut[1], l)))
    return e;
  if ((e = run_edge<128, 2, 64>(k4 ? "k4_edge_lut_128" : "k5_edge_lut_128", g, ct.elut[0], k, scratch, md,
                                d_queues + 9, ct.log_elut[0], l)))
    return e;
  if ((e = run_edge<1024, 0, 256>(k4 ? "k4_edge_list_1024" : "k5_edge_list_1024", g, ct.list[1], k, scratch, md,
                                  d_queues + 10, ct.log_list[1], l)))
    return e;
  if ((e = run_edge<128, 0, 64>(k4 ? "k4_edge_list_128" : "k5_edge_list_128", g, ct.list[0], k, scratch, md,
                                d_queues + 11, ct.log_list[0], l)))
    return e;
  if ((e = run_vlut_rank_hubs(k4 ? "k4_rlut_tri" : "k5_rlut_tri", k4 ? "k4_rlut_hbm" : "k5_rlut_hbm", g, ct.rlut_hbm,
                              k, ct.hub_slots, d_queues + 12, ct.log_rlut_hbm, l)))
    return e;
  if ((e = run_vlut_rank<1024, 16, 1024>(k4 ? "k4_rlut_1024" : "k5_rlut_1024", g, ct.rlut[3], k, d_queues + 4,
                                         ct.log_rlut[3], l)))
    return e;
  if ((e = run_vlut<1024, 16, 1024>(k4 ? "k4_lut_1024" : "k5_lut_1024", g, ct.lut[3], k, d_queues + 0, ct.log_lut[3],
                                    l)))
    return e;
  if ((e = run_vlut_rank_tri<512, 8, 256>(k4 ? "k4_rlut_512" : "k5_rlut_512", g, ct.rlut[2], k, d_queues + 5,
                                      ct.log_rlut[2], l)))
    return e;
  if ((e = run_vlut<512, 8, 256>(k4 ? "k4_lut_512" : "k5_lut_512", g, ct.lut[2], k, d_queues + 1, ct.log_lut[2], l)))
    return e;
  if ((e = run_vlut_rank<256, 4, 256>(k4 ? "k4_rlut_256" : "k5_rlut_256", g, ct.rlut[1], k, d_queues + 6,
                                      ct.log_rlut[1], l)))
    return e;
  if ((e = run_vlut<256, 4, 256>(k4 ? "k4_lut_256" : "k5_lut_256", g, ct.lut[1], k, d_queues + 2, ct.log_lut[1], l)))
    return e;
  if ((e = run_vlut_rank<64, 1, 64>(k4 ? "k4_rlut_64" : "k5_rlut_64", g, ct.rlut[0], k, d_queues + 7, ct.log_rlut[0],
                                    l)))
    return e;
  if (!ct.log_lut[0] && ct.lut[0].count) {  // per-task timing logs keep one task per wave
    constexpr int kPack = 8;
    auto kern = k_clique_vlut_packed<kPack>;
    const int grid = grid_for(kern, 128, 0, l.num_cu, (ct.lut[0].count + 2 * kPack - 1) / (2 * kPack));
    l.begin(k4 ? "k4_lut_64" : "k5_lut_64", ct.lut[0].count);
    kern<<<grid, 128, 0, l.stream>>>(g.rowptr, g.colidx, ct.lut[0].tasks, (int)ct.lut[0].count, k - 1, k - 1,
                                     d_queues + 3, l.d_out, l.slots);
    l.end();
    if ((e = hipGetLastError())) return e;
  } else if ((e = run_vlut<64, 1, 64>(k4 ? "k4_lut_64" : "k5_lut_64", g, ct.lut[0], k, d_queues + 3, ct.log_lut[0], l)))
    return e;
  return hipSuccess;
}

namespace {
template <int WHICH>
hipError_t launch_vmotif_which(const DevCSR &g, const TaskList tl[4], int32_t *scratch, int md, Launch &l,
                               int *q) {
  const char *names[3][4] = {{"vm_star_lut_64", "vm_star_lut_256", "vm_star_lut_1024", "vm_star_list"},
                             {"vm_diamond_lut_64", "vm_diamond_lut_256", "vm_diamond_lut_1024", "vm_diamond_list"},
                             {"vm_tailed_lut_64", "vm_tailed_lut_256", "vm_tailed_lut_1024", "vm_tailed_list"}};
  if (tl[3].count) {
    auto k = k_vmotif_list<WHICH>;
    const int grid = std::min(grid_for(k, 256, 0, l.num_cu, tl[3].count), vmotif_list_grid(l.num_cu));
    l.begin(names[WHICH][3], tl[3].count);
    k<<<grid, 256, 0, l.stream>>>(g.rowptr, g.colidx, tl[3].tasks, (int)tl[3].count, scratch, md, q + 3, l.d_out,
                                  l.slots);
    l.end();
    if (hipError_t e = hipGetLastError()) return e;
  }
  auto run = [&](auto kern, int nt, int i) -> hipError_t {
    if (!tl[i].count) return hipSuccess;
    const int grid = grid_for(kern, nt, 0, l.num_cu, tl[i].count);
    l.begin(names[WHICH][i], tl[i].count);
    kern<<<grid, nt, 0, l.stream>>>(g.rowptr, g.colidx, tl[i].tasks, (int)tl[i].count, q + i, l.d_out, l.slots);
    l.end();
    return hipGetLastError();
  };
  hipError_t e;
  if ((e = run(k_vmotif_lut<1024, 16, 1024, WHICH>, 1024, 2))) return e;
  if ((e = run(k_vmotif_lut<256, 4, 256, WHICH>, 256, 1))) return e;
  return run(k_vmotif_lut<64, 1, 64, WHICH>, 64, 0);
}
}  // namespace

int vmotif_list_grid(int num_cu) { return num_cu; }

hipError_t launch_vmotif(const DevCSR &g, int which, const TaskList tl[4], int32_t *scratch, int md, Launch &l,
                         int *d_queues) {
  switch (which) {
    case 0: return launch_vmotif_which<VM_STAR>(g, tl, scratch, md, l, d_queues);
    case 1: return launch_vmotif_which<VM_DIAMOND>(g, tl, scratch, md, l, d_queues);
    case 2: return launch_vmotif_which<VM_TAILED>(g, tl, scratch, md, l, d_queues);
    default: return hipErrorInvalidValue;
  }
}

// Per-vertex scheduler features (XGM_MLP_IN floats per vertex; zeros for
// vertices that are not k-clique tasks).  One thread per vertex.
__global__ void k_clique_features(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
                                  const int64_t *__restrict__ rrowptr, int64_t n, int k, float *__restrict__ F) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  float *f = F + v * 16;
  const int64_t b = rowptr[v];
  const int d = (int)(rowptr[v + 1] - b);
  if (d < k - 1) {
#pragma unroll
    for (int i = 0; i < 16; ++i) f[i] = 0.f;
    return;
  }
  const int32_t *X = colidx + b;
  double sum = 0;
  int mx = 0;
  for (int i = 0; i < d; ++i) {
    const int x = X[i];
    const int dx = (int)(rowptr[x + 1] - rowptr[x]);
    sum += dx;
    mx = max(mx, dx);
  }
  // sampled neighbour overlap: |N+(x) ∩ N+(v)| for up to 4 evenly spaced x
  const int ns = min(4, d);
  double ov = 0, dxs = 0;
  for (int q = 0; q < ns; ++q) {
    const int x = X[(int64_t)q * d / ns];
    const int32_t *Y = colidx + rowptr[x];
    const int dx = (int)(rowptr[x + 1] - rowptr[x]);
    int i = 0, j = 0, c = 0;
    while (i < d && j < dx) {
      const int a = X[i], bb = Y[j];
      c += (a == bb);
      i += (a <= bb);
      j += (bb <= a);
    }
    ov += c;
    dxs += dx;
  }
  const float rho = (float)(ov / ((double)ns * d));  // row density of the LUT
  const float r2 = (float)(ov / fmax(1.0, dxs));     // share of a neighbour's list inside N+(v)
  const float dd = (float)d;
  const float s = 1.f / 24.f;                         // keeps log2 features near [0, 1]
  f[0] = log2f(1.f + dd) * s * 2.f;
  f[1] = log2f(1.f + (float)sum) * s;
  f[2] = log2f(1.f + (float)mx) * s * 2.f;
  f[3] = r2;
  f[4] = rho;
  f[5] = log2f(1.f + dd * dd * rho * 0.5f) * s;
  f[6] = log2f(1.f + dd * dd * dd * rho * rho * rho / 6.f) * s;
  f[7] = log2f(1.f + dd * dd * dd * dd * rho * rho * rho * rho * rho * rho / 24.f) * s;
  f[8] = d <= 64 ? 1.f : 0.f;
  f[9] = (d > 64 && d <= 256) ? 1.f : 0.f;
  f[10] = (d > 256 && d <= 512) ? 1.f : 0.f;
  f[11] = (d > 512 && d <= 1024) ? 1.f : 0.f;
  f[12] = (float)((d + 63) >> 6) / 16.f;
  f[13] = log2f(1.f + (float)(rrowptr[v + 1] - rrowptr[v])) * s;
  f[14] = dd / 1024.f;
  f[15] = 1.f;
}

constexpr float kSchedMargin = 0.2f;

__global__ void k_sched_argmin(const int64_t *__restrict__ rowptr, int64_t n, const float *__restrict__ F,
                               const float *__restrict__ Y, int8_t *__restrict__ codes) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (F[v * 16 + 15] == 0.f) {
    codes[v] = -1;
    return;
  }
  const int d = (int)(rowptr[v + 1] - rowptr[v]);
  const float *y = Y + v * 5;
  // the model-free code of this task: the id LUT up to 64 rows (its packed
  // kernel), the rank LUT up to the hub LUT, the edge LUT above
  const int h = d <= 64 ? 0 : (d <= kHubLutMax ? 1 : 2);
  int best = h;
  float bv = y[h];
  for (int c = 0; c < 5; ++c) {
    // id-ordered LUT up to 1024 rows; rank LUT up to kHubLutMax (hub kernels above 1024)
    const bool ok = (c == 2 || c == 3) || (c == 0 ? d <= 1024 : d <= kHubLutMax);
    if (ok && y[c] < bv) { bv = y[c]; best = c; }
  }
  // leave the model-free code only for a predicted saving above the model's
  // error (kSchedMargin in log2(1 + 1000 us) units, about 13% of the cost)
  if (best != h && !(bv < y[h] - kSchedMargin)) best = h;
  // up to 64 rows the packed id-LUT kernel (8 tasks per wave) runs a task
  // faster than a rank-LUT block, in this bin or one bin up; the per-task
  // logs the model learned from time one task per kernel and do not see it
  if ((best == 1 || best == 4) && d <= 64) best = 0;
  codes[v] = (int8_t)best;
}

hipError_t launch_sched_argmin(const DevCSR &g, const float *F, const float *Y, int8_t *codes, hipStream_t s) {
  if (!g.n) return hipSuccess;
  k_sched_argmin<<<(unsigned)((g.n + 255) / 256), 256, 0, s>>>(g.rowptr, g.n, F, Y, codes);
  return hipGetLastError();
}

hipError_t launch_clique_features(const DevCSR &g, int k, float *F, hipStream_t s) {
  const int64_t n = g.n;
  if (!n) return hipSuccess;
  k_clique_features<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g.rowptr, g.colidx, g.rrowptr, n, k, F);
  return hipGetLastError();
}

}  // namespace xgm
