// Flattened segment machinery shared by the vertex-task kernels.
//
// A task probes (or aggregates) the elements of many short sorted lists
// N(v), v in some vertex list.  Rather than giving each list its own lanes
// (most lists are far shorter than 64), the lists are concatenated into one
// index space [0, L): a wave takes 64 consecutive indices per step, and each
// lane finds its list from a 64-bit mask of list ends inside the window.
#pragma once

#include <climits>

#include "device_common.h"

namespace xgm {

template <int PCAP>
struct Segs {
  int64_t off[PCAP];  // first element of each non-empty list (index into colidx)
  int pre[PCAP];      // inclusive prefix of the lengths
  int idx[PCAP];      // position of the owning vertex in the vertex list
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
      t.off[pos] = vb;
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

// Lane view of one window [base, base+64): segment j, element value w (if
// i < L), and the mask of lanes that open a new segment inside the window.
struct WindowLane {
  int j;
  int w;
  bool valid;
  u64 heads;  // bit l set if lane l starts a segment (lane 0 always set)
};

template <int PCAP>
__device__ __forceinline__ WindowLane window_lane(const Segs<PCAP> &t, const int32_t *__restrict__ colidx,
                                                  int ns, int L, int base, int lane) {
  WindowLane r;
  const int jlo = upper_bound_dev(t.pre, ns, base);
  const int e = (jlo + lane < ns) ? t.pre[jlo + lane] : INT_MAX;
  const int p = e - base;
  const u64 ends = wave_or_u64(p < 64 ? (1ull << p) : 0ull);
  r.j = jlo + __popcll(ends & lanemask_le(lane));
  r.heads = ends | 1ull;
  const int i = base + lane;
  r.valid = i < L;
  r.w = 0;
  if (r.valid) {
    const int start = r.j > 0 ? t.pre[r.j - 1] : 0;
    r.w = colidx[t.off[r.j] + (i - start)];
  }
  return r;
}

// Per-segment reduction of a hit bit inside one window: returns, in the
// last lane of each segment run, the number of hits in that run (0 in
// other lanes).  Lets a kernel issue one LDS atomic per segment per window.
__device__ __forceinline__ int segment_run_hits(u64 hits, u64 heads, bool valid_next_in_run, int lane) {
  // the run of lane l starts at the highest head <= l
  const u64 le = heads & lanemask_le(lane);
  const int first = 63 - __clzll(le);
  const bool last = (lane == 63) || ((heads >> (lane + 1)) & 1ull) || !valid_next_in_run;
  if (!last) return 0;
  const u64 run = lanemask_le(lane) & ~lanemask_lt(first);
  return __popcll(hits & run);
}

}  // namespace xgm
