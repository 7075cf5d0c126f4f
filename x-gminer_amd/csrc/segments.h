// Flattened segment machinery shared by the vertex-task kernels.
//
// A task probes (or aggregates) the elements of many short sorted lists
// N(v), v in some vertex list.  Rather than giving each list its own lanes
// (most lists are far shorter than 64), the lists are concatenated into one
// index space [0, L): a wave walks 64 consecutive indices per window, and
// each lane finds its list from a 64-bit mask of list ends inside the
// window.  A wave owns a contiguous range of windows, so the first list of
// the next window follows from the last lane of the current one (one LDS
// read) instead of a fresh binary search, and U windows are mapped before
// their U global loads are issued together.
#pragma once

#include <climits>

#include "device_common.h"

namespace xgm {

template <int PCAP>
struct Segs {
  uint32_t off[PCAP];  // first element of each non-empty list (index into colidx; nnz < 2^31)
  int pre[PCAP];       // inclusive prefix of the lengths
  int idx[PCAP];       // position of the owning vertex in the vertex list
};

// Build the table for verts[k0, k1) with one wave.  keep(v, len) selects
// lists; bound >= 0 cuts every list to its elements < bound (one binary
// search per list).  Returns the number of non-empty lists and L.
template <int PCAP, class Keep>
__device__ __forceinline__ void wave_segments(Segs<PCAP> &t, const int64_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ colidx, const int *verts, int k0,
                                              int k1, int bound, Keep keep, int lane, int &ns_out,
                                              int &L_out) {
  int ns = 0, carry = 0;
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
    const int incl = wave_incl_scan(len, lane) + carry;
    const u64 m = __ballot(len > 0);
    if (len > 0) {
      const int pos = ns + __popcll(m & lanemask_lt(lane));
      t.off[pos] = (uint32_t)vb;
      t.pre[pos] = incl;
      t.idx[pos] = kk;
    }
    ns += __popcll(m);
    carry = __shfl(incl, 63, 64);
  }
  ns_out = ns;
  L_out = carry;
}

struct KeepAll {
  __device__ bool operator()(int, int) const { return true; }
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

// Walk flattened elements [lo, hi) with one wave (lo a multiple of 64 or
// the start of the range), U windows per step.  f(const WindowLane&) is
// called by all 64 lanes for every window (invalid lanes included, so f
// may use ballots).
template <int U, int PCAP, class F>
__device__ __forceinline__ void wave_walk(const Segs<PCAP> &t, const int32_t *__restrict__ colidx, int ns,
                                          int lo, int hi, int lane, F &&f) {
  if (lo >= hi) return;
  int jl = upper_bound_dev(t.pre, ns, lo);  // segment holding element lo
  for (int base = lo; base < hi; base += 64 * U) {
    WindowLane r[U];
    int64_t pos[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = base + 64 * u;
      const int e = (jl + lane < ns) ? t.pre[jl + lane] : INT_MAX;
      const int p = e - b;  // >= 1 for in-range windows
      const u64 ends = wave_or_u64((p > 0 && p < 64) ? (1ull << p) : 0ull);
      const int j = jl + __popcll(ends & lanemask_le(lane));
      const int i = b + lane;
      r[u].valid = i < hi;
      r[u].next_valid = i + 1 < hi;
      r[u].j = j;
      r[u].heads = ends | 1ull;
      pos[u] = 0;
      if (r[u].valid) {
        const int st = j > 0 ? t.pre[j - 1] : 0;
        pos[u] = (int64_t)t.off[j] + (i - st);
      }
      // first segment of the next window
      if (u + 1 < U || base + 64 * U < hi) {
        const int j63 = __shfl(j, 63, 64);
        jl = (j63 < ns && t.pre[j63] <= b + 64) ? j63 + 1 : j63;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) r[u].w = r[u].valid ? colidx[pos[u]] : 0;
#pragma unroll
    for (int u = 0; u < U; ++u) f(r[u]);
  }
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
