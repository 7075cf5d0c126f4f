// Flattened segment machinery shared by the vertex-task kernels.
//
// A task probes (or aggregates) the elements of many short sorted lists
// N(v), v in some vertex list.  Rather than giving each list its own lanes
// (most lists are far shorter than 64), the non-empty lists are
// concatenated into one index space [0, L): a wave walks 64 consecutive
// indices per window.  Lane -> list is resolved per window without any
// cross-lane reduction: the lanes that hold the ends of the next 64 lists
// scatter a per-window stamp into a 64-entry LDS flag row at the position
// where their list ends, one ballot over the row marks the lanes that open
// a new list, and mbcnt turns it into each lane's list index.  The first
// list of the next window is the current one plus the number of lists that
// end inside this window (one more ballot).  A wave owns a contiguous range
// of windows and maps U windows before issuing their U global loads.
#pragma once

#include <climits>

#include "device_common.h"

namespace xgm {

template <int PCAP, bool IDX = true>
struct Segs {
  static constexpr bool kIdx = IDX;
  int2 os[PCAP + 1];       // .x = first element (index into colidx; nnz < 2^31), .y = start in [0, L); os[ns].y = L
  int idx[IDX ? PCAP : 1]; // position of the owning vertex in the vertex list (if kept)
};

// Scan of (len, keep) over verts[k0, k1) that compacts the non-empty lists
// into t starting at slot ns_in with flattened start L_in.  One wave.
template <class Seg, class Keep>
__device__ __forceinline__ void wave_segments_range(Seg &t, const int64_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx, const int *verts, int k0,
                                                    int k1, int bound, Keep keep, int lane, int &ns, int &L) {
  for (int k = k0; k < k1; k += 64) {
    const int kk = k + lane;
    int len = 0;
    int64_t vb = 0;
    if (kk < k1) {
      const int v = verts[kk];
      vb = rowptr[v];
      len = (int)(rowptr[v + 1] - vb);
      if (!keep(v, len)) len = 0;
      else if (bound >= 0) len = lower_bound_dev(colidx + vb, len, bound);
    }
    const int incl = wave_incl_scan(len, lane) + L;
    const u64 m = __ballot(len > 0);
    if (len > 0) {
      const int pos = ns + __popcll(m & lanemask_lt(lane));
      t.os[pos] = make_int2((int)vb, incl - len);
      if constexpr (Seg::kIdx) t.idx[pos] = kk;
    }
    ns += __popcll(m);
    L = __shfl(incl, 63, 64);
  }
}

// wave_segments_range with every list verts[k] cut to its first cut[k]
// elements (precomputed cut positions).
template <class Seg>
__device__ __forceinline__ void wave_segments_range_cut(Seg &t, const int64_t *__restrict__ rowptr, const int *verts,
                                                        const int *cut, int k0, int k1, int lane, int &ns, int &L) {
  for (int k = k0; k < k1; k += 64) {
    const int kk = k + lane;
    int len = 0;
    int64_t vb = 0;
    if (kk < k1) {
      vb = rowptr[verts[kk]];
      len = cut[kk];
    }
    const int incl = wave_incl_scan(len, lane) + L;
    const u64 m = __ballot(len > 0);
    if (len > 0) {
      const int pos = ns + __popcll(m & lanemask_lt(lane));
      t.os[pos] = make_int2((int)vb, incl - len);
      if constexpr (Seg::kIdx) t.idx[pos] = kk;
    }
    ns += __popcll(m);
    L = __shfl(incl, 63, 64);
  }
}

// Same as wave_segments_range over a virtual vertex list: vert(k) gives
// the k-th vertex and keep(k, v, len) selects it.
template <class Seg, class VertAt, class KeepK>
__device__ __forceinline__ void wave_segments_fn(Seg &t, const int64_t *__restrict__ rowptr, int k0, int k1,
                                                 VertAt vert, KeepK keep, int lane, int &ns, int &L) {
  for (int k = k0; k < k1; k += 64) {
    const int kk = k + lane;
    int len = 0;
    int64_t vb = 0;
    if (kk < k1) {
      const int v = vert(kk);
      vb = rowptr[v];
      len = (int)(rowptr[v + 1] - vb);
      if (!keep(kk, v, len)) len = 0;
    }
    const int incl = wave_incl_scan(len, lane) + L;
    const u64 m = __ballot(len > 0);
    if (len > 0) {
      const int pos = ns + __popcll(m & lanemask_lt(lane));
      t.os[pos] = make_int2((int)vb, incl - len);
      if constexpr (Seg::kIdx) t.idx[pos] = kk;
    }
    ns += __popcll(m);
    L = __shfl(incl, 63, 64);
  }
  if (lane == 0) t.os[ns] = make_int2(0, L);
}

// Build the table for verts[k0, k1) with one wave.  keep(v, len) selects
// lists; bound >= 0 cuts every list to its elements < bound.
template <class Seg, class Keep>
__device__ __forceinline__ void wave_segments(Seg &t, const int64_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ colidx, const int *verts, int k0,
                                              int k1, int bound, Keep keep, int lane, int &ns_out,
                                              int &L_out) {
  int ns = 0, L = 0;
  wave_segments_range(t, rowptr, colidx, verts, k0, k1, bound, keep, lane, ns, L);
  if (lane == 0) t.os[ns] = make_int2(0, L);
  ns_out = ns;
  L_out = L;
}

// Second half of the block-cooperative builds below: with t.os[q].x and
// tmp[q] holding the offset and length of list q < n, wave 0 compacts the
// non-empty lists in place and writes the flattened starts.
template <class Seg>
__device__ __forceinline__ void block_segments_compact(Seg &t, const int *tmp, int n, int k0, int *sh_ns,
                                                       int *sh_L) {
  const int tid = threadIdx.x, lane = tid & 63;
  __syncthreads();
  if (tid < 64) {
    int ns = 0, L = 0;
    for (int k = 0; k < n; k += 64) {
      const int kk = k + lane;
      int len = 0, vb = 0;
      if (kk < n) {
        len = tmp[kk];
        vb = t.os[kk].x;
      }
      const int incl = wave_incl_scan(len, lane) + L;
      const u64 m = __ballot(len > 0);
      if (len > 0) {  // in place: pos <= kk and the whole chunk was read above
        const int pos = ns + __popcll(m & lanemask_lt(lane));
        t.os[pos] = make_int2(vb, incl - len);
        if constexpr (Seg::kIdx) t.idx[pos] = k0 + kk;
      }
      ns += __popcll(m);
      L = __shfl(incl, 63, 64);
    }
    if (lane == 0) {
      t.os[ns] = make_int2(0, L);
      *sh_ns = ns;
      *sh_L = L;
    }
  }
  __syncthreads();
}

// Block-cooperative build: every thread loads the (offset, length) of its
// share of verts[k0, k1) (k1 - k0 <= PCAP), then wave 0 compacts the
// non-empty lists in LDS.  tmp holds PCAP ints.  All threads call; the
// result is visible to all threads on return.
template <int NT, class Seg, class Keep>
__device__ __forceinline__ void block_segments(Seg &t, int *tmp, const int64_t *__restrict__ rowptr,
                                               const int32_t *__restrict__ colidx, const int *verts, int k0,
                                               int k1, int bound, Keep keep, int *sh_ns, int *sh_L) {
  const int n = k1 - k0;
  for (int q = threadIdx.x; q < n; q += NT) {
    const int v = verts[k0 + q];
    const int64_t vb = rowptr[v];
    int len = (int)(rowptr[v + 1] - vb);
    if (!keep(v, len)) len = 0;
    else if (bound >= 0) len = lower_bound_dev(colidx + vb, len, bound);
    t.os[q].x = (int)vb;
    tmp[q] = len;
  }
  block_segments_compact(t, tmp, n, k0, sh_ns, sh_L);
}

// As block_segments with every list verts[k] cut to its first cut[k]
// elements (precomputed cut positions instead of binary searches).
template <int NT, class Seg>
__device__ __forceinline__ void block_segments_cut(Seg &t, int *tmp, const int64_t *__restrict__ rowptr,
                                                   const int *verts, const int *cut, int k0, int k1, int *sh_ns,
                                                   int *sh_L) {
  const int n = k1 - k0;
  for (int q = threadIdx.x; q < n; q += NT) {
    t.os[q].x = (int)rowptr[verts[k0 + q]];
    tmp[q] = cut[k0 + q];
  }
  block_segments_compact(t, tmp, n, k0, sh_ns, sh_L);
}

// Block-cooperative build over a virtual vertex list (see wave_segments_fn).
template <int NT, class Seg, class VertAt, class KeepK>
__device__ __forceinline__ void block_segments_fn(Seg &t, int *tmp, const int64_t *__restrict__ rowptr, int k0,
                                                  int k1, VertAt vert, KeepK keep, int *sh_ns, int *sh_L) {
  const int n = k1 - k0;
  for (int q = threadIdx.x; q < n; q += NT) {
    const int v = vert(k0 + q);
    const int64_t vb = rowptr[v];
    int len = (int)(rowptr[v + 1] - vb);
    if (!keep(k0 + q, v, len)) len = 0;
    t.os[q].x = (int)vb;
    tmp[q] = len;
  }
  block_segments_compact(t, tmp, n, k0, sh_ns, sh_L);
}

struct KeepAll {
  __device__ bool operator()(int, int) const { return true; }
};

// Per-wave state of the window mapping: a 64-entry LDS flag row (set to -1
// once) and a running stamp.
struct WaveFlags {
  int *row;
  int stamp;
  __device__ __forceinline__ void init(int *r, int lane) {
    row = r;
    stamp = 0;
    row[lane] = -1;
    wave_sync();
  }
};

// Lane view of one window: segment j, element w (valid if inside the
// range), and the mask of lanes that open a new segment (lane 0 included).
struct WindowLane {
  int j;
  int w;
  bool valid;
  bool next_valid;  // the element after this lane's is inside the range
  u64 heads;
};

// Walk flattened elements [lo, hi) with one wave, U windows per step.
// f(const WindowLane (&)[U]) is called by all 64 lanes for every group of
// U windows (invalid lanes included, so f may use ballots; their w is INV),
// after the U global loads of the group were issued.
//
// The segment holding the next window's first element (index jl, its
// offset delta x - y and its end) is kept in scalar registers, so a window
// that lies inside it (the common case for long lists) maps its lanes
// without touching LDS.  A window that crosses ends reads the next 64 table
// entries with one LDS access, marks the lanes that open a segment in the
// flag row, takes each lane's offset from the entry's lane by a cross-lane
// permute and the next window's segment from the same entries (readlane).
template <int U, bool PIPE = false, int INV = 0, class Seg, class F>
__device__ __forceinline__ void wave_walk_group(const Seg &t, const int32_t *__restrict__ colidx, int ns,
                                                int lo, int hi, int lane, WaveFlags &fl, F &&f) {
  // the range and table size are the same on every lane (scalar registers)
  lo = __builtin_amdgcn_readfirstlane(lo);
  hi = __builtin_amdgcn_readfirstlane(hi);
  ns = __builtin_amdgcn_readfirstlane(ns);
  if (lo >= hi) return;
  // segment holding lo: the last j with os[j].y <= lo
  int jl;
  {
    int a = 0, b = ns;
    while (b - a > 1) {
      const int mid = (a + b) >> 1;
      if (t.os[mid].y <= lo) a = mid; else b = mid;
    }
    // wave-uniform: held in scalar registers, so the per-window branch below
    // is a scalar branch instead of two exec-masked paths
    jl = __builtin_amdgcn_readfirstlane(a);
  }
  // delta (x - y: element i of the segment is colidx[delta + i]) and end of
  // segment jl; jl < ns here, so jl + 1 <= ns is a table entry
  int cur_d, cur_e;
  {
    const int2 c = t.os[jl];
    cur_d = __builtin_amdgcn_readfirstlane(c.x - c.y);
    cur_e = __builtin_amdgcn_readfirstlane(t.os[jl + 1].y);
  }
  // map the U windows of a group at base and issue their loads.  PIPE
  // loads every lane (invalid lanes read colidx[0], which exists because the
  // range is not empty) so the loads of the next group can be in flight
  // while f handles the current one: the same number of loads is
  // outstanding on every path, and f waits for the older group only.
  auto map_group = [&](int base, WindowLane (&r)[U]) {
    int pos[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = base + 64 * u;
      const int i = b + lane;
      r[u].valid = i < hi;
      r[u].next_valid = i + 1 < hi;
      pos[u] = 0;
      // invariant: os[jl].y <= b < cur_e = os[jl + 1].y (no empty segments)
      if (cur_e >= b + 64) {
        r[u].j = jl;
        r[u].heads = 1ull;
        if (r[u].valid) pos[u] = cur_d + i;
        if (cur_e == b + 64 && b + 64 < hi) {
          // the next window opens segment jl + 1, which is not the end
          // sentinel (its start b + 64 < hi <= L), so jl + 2 <= ns
          ++jl;
          const int2 c = t.os[jl];
          cur_d = __builtin_amdgcn_readfirstlane(c.x - c.y);
          cur_e = __builtin_amdgcn_readfirstlane(t.os[jl + 1].y);
        }
        continue;
      }
      // entries jl + 1 + lane: the segments that start after b (the sentinel
      // os[ns] = (0, L) included; lanes past it get a start beyond any window)
      const int k = jl + 1 + lane;
      const int2 o = (k <= ns) ? t.os[k] : make_int2(0, INT_MAX);
      const int p = o.y - b;  // >= 1: segment jl holds b
      const int dk = o.x - o.y;
      ++fl.stamp;
      if (p > 0 && p < 64) fl.row[p] = fl.stamp;
      wave_sync();
      const u64 heads = __ballot(fl.row[lane] == fl.stamp);
      const int nh = __popcll(heads & lanemask_le(lane));  // segments opened at or before this lane
      const int dsh = __shfl(dk, nh > 0 ? nh - 1 : 0, 64);
      r[u].j = jl + nh;
      r[u].heads = heads | 1ull;
      if (r[u].valid) pos[u] = (nh ? dsh : cur_d) + i;
      // next window: segment jl + cnt, cnt = segments starting in (b, b + 64]
      // (>= 1: os[jl + 1].y = cur_e < b + 64)
      const int cnt = __popcll(__ballot(p <= 64));
      if (b + 64 < hi) {
        jl += cnt;
        if (cnt < 64) {
          cur_d = __builtin_amdgcn_readlane(dk, cnt - 1);
          cur_e = __builtin_amdgcn_readlane(o.y, cnt);
        } else {  // 64 one-element segments: the next entries were not read
          const int2 c = t.os[jl];
          cur_d = __builtin_amdgcn_readfirstlane(c.x - c.y);
          cur_e = __builtin_amdgcn_readfirstlane(t.os[jl + 1].y);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (PIPE) {
        const int w = colidx[(uint32_t)pos[u]];
        r[u].w = r[u].valid ? w : INV;
      } else {
        r[u].w = r[u].valid ? colidx[(uint32_t)pos[u]] : INV;
      }
    }
  };
  if constexpr (!PIPE) {
    for (int base = lo; base < hi; base += 64 * U) {
      WindowLane r[U];
      map_group(base, r);
      f(r);
    }
  } else {
    WindowLane r[U];
    map_group(lo, r);
    for (int base = lo;; base += 64 * U) {
      // the group past the range maps lanes that are all invalid; its state
      // updates are guarded by b + 64 < hi
      WindowLane nx[U];
      map_group(base + 64 * U, nx);
      f(r);
      if (base + 64 * U >= hi) break;
#pragma unroll
      for (int u = 0; u < U; ++u) r[u] = nx[u];
    }
  }
}

// wave_walk_group with f(const WindowLane&) called for every window.
template <int U, bool PIPE = false, class Seg, class F>
__device__ __forceinline__ void wave_walk(const Seg &t, const int32_t *__restrict__ colidx, int ns,
                                          int lo, int hi, int lane, WaveFlags &fl, F &&f) {
  wave_walk_group<U, PIPE>(t, colidx, ns, lo, hi, lane, fl, [&](const WindowLane (&r)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) f(r[u]);
  });
}

// Contiguous, 64-aligned share [lo, hi) of [0, L) for wave `wave` of `nw`.
__device__ __forceinline__ void wave_share(int L, int wave, int nw, int &lo, int &hi) {
  const int per = ((L + nw - 1) / nw + 63) & ~63;
  lo = min(L, wave * per);
  hi = min(L, lo + per);
}

// Per-segment reduction of a hit bit inside one window: returns, in the
// last lane of each segment run, the number of hits in that run (0 in
// other lanes).  Lets a kernel issue one LDS atomic per segment per window.
__device__ __forceinline__ int segment_run_hits(u64 hits, const WindowLane &r, int lane) {
  if (!r.valid) return 0;
  const u64 le = r.heads & lanemask_le(lane);
  const int first = 63 - __clzll(le);
  const bool last = (lane == 63) || ((r.heads >> (lane + 1)) & 1ull) || !r.next_valid;
  if (!last) return 0;
  const u64 run = lanemask_le(lane) & ~lanemask_lt(first);
  return __popcll(hits & run);
}

}  // namespace xgm
